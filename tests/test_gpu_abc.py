"""Non-Hermitian (bi-Lanczos) Haydock recursion (ocean_gpu_haydock_abc) and the
continued fraction with complex b and c (ocean_gpu_spectrum_abc) against the oracle.

The problems are made non-Hermitian the ways the backwards action is the Hermitian
conjugate for: a complex-symmetric spin-orbit matrix (random, or the reference's
soprep with core-hole lifetimes) and complex band energies (a lifetime term).  Tolerances: the coefficients of 30 FP64 steps 1e-9
relative to the largest coefficient (the two recursions sum their dots in different
orders and the bi-Lanczos recursion amplifies rounding more than the Hermitian one);
the fraction kernel 1e-10 of peak on the same coefficients; spectra of the two
recursions 1e-8 of peak through the FP64 fraction, 1e-3 in the mixed precision."""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import to_problem

pytestmark = pytest.mark.gpu

from ocean_amd import core, setup_host as sh, synth  # noqa: E402

E2H = orc.EV2HARTREE

CASES = {
    # K edge, four alphas on one spin (VALU action kernels)
    "k_small": dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 4), lc=0, seed=41),
    # odd counts, two spins
    "k_ragged_nspn2": dict(nb=11, kmesh=(2, 3, 3), xmesh=(5, 3, 3), lc=0, nspn=2, seed=42),
    # L2,3: twelve alphas (matrix-core projection and adjoint) with its own
    # spin-orbit term made complex-symmetric
    "l23": dict(nb=13, kmesh=(3, 2, 2), xmesh=(3, 3, 5), lc=1, seed=43, xi_eV=3.0),
    # the branch's physical use: the L2,3 spin-orbit term with core-hole lifetimes
    # (core_broaden.ipt), S = S_r + i S_i from the reference's soprep, both parts real
    # symmetric, so the backwards action is exactly H^dagger; real band energies
    "l23_lifetime": dict(nb=13, kmesh=(3, 2, 2), xmesh=(3, 3, 5), lc=1, seed=44, xi_eV=3.0),
}
LIFETIME_EV = {"l23_lifetime": (0.6, 0.9)}


def _inputs(name, energy_lifetime=True):
    kw = CASES[name]
    inp = synth.make_inputs(**kw)
    rng = np.random.default_rng(kw["seed"] + 200)
    na = inp.energy.shape[2]
    if name in LIFETIME_EV:
        sr, si, _ = orc.soprep(kw["lc"], kw["xi_eV"], LIFETIME_EV[name])
        assert np.abs(si).max() > 0 and np.array_equal(si, si.T) and np.array_equal(sr, sr.T)
        inp.mult.somelr = np.asfortranarray(sr)
        inp.mult.someli = np.asfortranarray(si)
        energy_lifetime = False
    else:
        s = rng.standard_normal((na, na)) + 1j * rng.standard_normal((na, na))
        s = 2e-3 * (s + s.T)
        if inp.mult.have_spin_orbit:
            s = s + inp.mult.somelr  # keep the real (symmetric) part of the L2,3 term
        inp.mult.somelr = np.asfortranarray(s.real)
        inp.mult.someli = np.asfortranarray(s.imag)
    inp.mult.have_spin_orbit = True
    if energy_lifetime:
        inp.energy = np.asfortranarray(inp.energy + 1j * 0.01 * rng.uniform(size=inp.energy.shape))
    inp.allow = np.asfortranarray((rng.uniform(size=inp.allow.shape) > 0.2).astype(np.float64))
    z = rng.standard_normal(inp.energy.shape) + 1j * rng.standard_normal(inp.energy.shape)
    psi0 = sh.normalise_psi0(np.asfortranarray(z), inp.nk, inp.celvol)[0]
    return inp, psi0


def _coef_err(got, ref, n):
    errs = []
    for g, r in zip(got, ref):
        errs.append(np.abs(g[: n + 1] - r[: n + 1]).max())
    scale = max(np.abs(r[: n + 1]).max() for r in ref)
    return max(errs) / scale


@pytest.mark.parametrize("name", list(CASES))
def test_haydock_abc_matches_oracle(name):
    inp, psi0 = _inputs(name)
    P = to_problem(inp)
    n = 30
    print("haydock_abc", name, flush=True)
    ref = orc.haydock_abc(P, psi0, n)
    assert ref[-1] == 0
    with core.OceanGPU.from_inputs(inp) as g:
        got = g.haydock_abc(psi0, n)
    err = _coef_err(got, ref[:6], n)
    print(name, "coefficient err", err, flush=True)
    assert err < 1e-9
    gam, el, eh, ne = 0.5 * E2H, -5.0 * E2H, 40.0 * E2H, 400
    for lit in (True, False):
        # the fraction kernel on the same coefficients
        s_got = core.spectrum_abc(*got, n, 1.0, gam, el, eh, ne, literal_cmplx=lit)
        s_ref = orc.write_core_abc(*got, n, 1.0, gam, el, eh, ne, literal_cmplx=lit)
        for col in (1, 3):
            e = np.abs(s_got[:, col] - s_ref[:, col]).max() / np.abs(s_ref[:, col]).max()
            assert e < 1e-10, (lit, col, e)
    # the recursion through the FP64 fraction (the single-precision one rounds the
    # products b c, so coefficients 1e-11 apart can land 1e-7 apart there)
    s_got = core.spectrum_abc(*got, n, 1.0, gam, el, eh, ne, literal_cmplx=False)
    s_ref = orc.write_core_abc(*ref[:6], n, 1.0, gam, el, eh, ne, literal_cmplx=False)
    for col in (1, 3):
        e = np.abs(s_got[:, col] - s_ref[:, col]).max() / np.abs(s_ref[:, col]).max()
        print(name, "spectrum column", col, "err of peak", e, flush=True)
        assert e < 1e-8, (col, e)


def test_haydock_abc_hermitian_problem_matches_haydock():
    """A real-symmetric spin-orbit matrix and real energies: the backwards action is
    the forward one and the recursion reduces to the Hermitian one."""
    inp = synth.make_inputs(**CASES["k_small"])
    rng = np.random.default_rng(7)
    na = inp.energy.shape[2]
    s = rng.standard_normal((na, na))
    s = 2e-3 * (s + s.T)
    inp.mult.somelr = np.asfortranarray(s)
    inp.mult.someli = np.asfortranarray(np.zeros_like(s))
    inp.mult.have_spin_orbit = True
    psi0 = inp.psi0
    n = 30
    with core.OceanGPU.from_inputs(inp) as g:
        ra, ia, rb, ib, rc, ic = g.haydock_abc(psi0, n)
        ha, hb, _ = g.haydock(psi0, n)
    assert np.abs(ra[:n] - ha[:n]).max() < 1e-9 * np.abs(ha).max()
    assert np.abs(rb[1:] - hb[1:]).max() < 1e-9 * np.abs(hb).max()
    assert np.abs(rc[1:] - hb[1:]).max() < 1e-9 * np.abs(hb).max()
    assert np.abs(ib).max() < 1e-9 * np.abs(hb).max() and np.abs(ic).max() < 1e-9 * np.abs(hb).max()
    gam, el, eh, ne = 0.5 * E2H, -5.0 * E2H, 40.0 * E2H, 300
    z = np.zeros(n + 1)
    for lit in (True, False):
        a = core.spectrum(ha, hb, n, 1.0, gam, el, eh, ne, literal_cmplx=lit)
        b = core.spectrum_abc(ha, z, hb, z, hb, z, n, 1.0, gam, el, eh, ne, literal_cmplx=lit)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()


def test_haydock_abc_mixed_precision_spectrum():
    inp, psi0 = _inputs("l23")
    P = to_problem(inp)
    n = 30
    ref = orc.haydock_abc(P, psi0, n)
    with core.OceanGPU.from_inputs(inp, precision=core.MIXED) as g:
        got = g.haydock_abc(psi0, n)
    gam, el, eh, ne = 0.5 * E2H, -5.0 * E2H, 40.0 * E2H, 400
    s_got = core.spectrum_abc(*got, n, 1.0, gam, el, eh, ne)[:, 1]
    s_ref = orc.write_core_abc(*ref[:6], n, 1.0, gam, el, eh, ne)[:, 1]
    err = np.abs(s_got - s_ref).max() / np.abs(s_ref).max()
    print("mixed spectrum err of peak", err, flush=True)
    assert err < 1e-3


def test_haydock_abc_nan_and_refusal():
    inp, psi0 = _inputs("k_small")
    bad = psi0.copy()
    bad[0, 0, 0] = np.nan
    inp.allow[0, 0, 0] = 1.0
    with core.OceanGPU.from_inputs(inp) as g:
        *coef, done, rc = g.haydock_abc(bad, 6, return_status=True)
    assert rc == -1 and done == 1
    assert np.isnan(coef[0][0]) and np.all(coef[0][1:] == 0.0)
    kw = dict(nb=4, kmesh=(2, 2, 2), xmesh=(2, 2, 2), lc=0, seed=3)
    inp = synth.make_inputs(**kw)
    with core.OceanGPU.from_inputs(inp, nrhs=3) as g:
        z = np.asfortranarray(np.ones(g.shape, complex))
        rc = g.haydock_abc(z, 4, return_status=True)[-1]
    assert rc == 9963
