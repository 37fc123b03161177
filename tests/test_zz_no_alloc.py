"""The allocation-free continued fraction (ocean_gpu_spectrum_noalloc) against
ocean_gpu_spectrum, through the C-ABI: the same kernel on the same inputs, so the
outputs must be identical, on two calls in a row (the second with a smaller input,
and no imag_a, so nothing of the first call's staging may leak into it)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_DP = C.POINTER(C.c_double)


def _noalloc():
    from ocean_amd import core
    return core._lib().ocean_gpu_spectrum_noalloc  # argument types bound by core


def _call(f, a, b, ia, n, kpref, gam, el, eh, ne, lit):
    out = np.zeros((ne + 1, 5))
    keep = [np.ascontiguousarray(x, np.float64) for x in (a, b) + ((ia,) if ia is not None else ())]
    rc = f(keep[0].ctypes.data_as(_DP), keep[1].ctypes.data_as(_DP),
           keep[2].ctypes.data_as(_DP) if ia is not None else None, n, kpref, gam, el, eh, ne, lit,
           out.ctypes.data_as(_DP))
    return rc, out


def test_noalloc_spectrum_matches_allocating_entry_point():
    from ocean_amd import core
    f = _noalloc()
    rng = np.random.default_rng(17)
    # first call: 300 levels with imag_a, 4000 energies; second: 40 levels, none, 700
    cases = [(300, 4000, True, 1), (40, 700, False, 0)]
    for n, ne, with_ia, lit in cases:
        a = rng.uniform(0.2, 1.0, n + 1)
        b = rng.uniform(0.1, 0.5, n + 1)
        b[0] = 0.0
        ia = rng.uniform(-0.01, 0.01, n + 1) if with_ia else None
        print("noalloc case", n, ne, with_ia, lit, flush=True)
        rc, got = _call(f, a, b, ia, n, 1.3e-3, 0.02, -0.2, 1.5, ne, lit)
        assert rc == 0
        ref = core.spectrum(a, b, n, 1.3e-3, 0.02, -0.2, 1.5, ne, imag_a=ia, literal_cmplx=bool(lit))
        assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_noalloc_spectrum_refuses_sizes_past_its_array():
    f = _noalloc()
    a = np.zeros(2)
    b = np.zeros(2)
    # 3 (n + 1) + 5 (ne + 1) doubles must fit in 2^19: ne = 2^17 does not
    rc, _ = _call(f, a, b, None, 1, 1.0, 0.02, -0.2, 1.5, 1 << 17, 1)
    from ocean_amd import core
    assert rc == 1001, core._lib().ocean_gpu_strerror(rc)
