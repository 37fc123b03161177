"""Batched form: many problems of uneven size in one launch per stage.

The batch reuses the single-problem kernel bodies with the same partitions, so
with the default tiles a batched Haydock recursion must reproduce every
problem's stand-alone coefficients bit for bit; after autotuning (which may
change the adjoint's x-split and hence its summation order) H.psi must agree to
1e-12 relative and a/b to 1e-10 over the first 50 steps.
"""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import relerr, to_problem

pytestmark = pytest.mark.gpu

from ocean_amd import core, synth  # noqa: E402

MEMBERS = [
    dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 4), lc=0, seed=21),
    dict(nb=24, kmesh=(2, 3, 4), xmesh=(3, 4, 5), lc=0, seed=22),
    dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 4), lc=1, seed=23, xi_eV=6.0),
    dict(nb=8, kmesh=(2, 2, 4), xmesh=(3, 3, 3), lc=1, seed=24, nspn=2, xi_eV=3.0),
    dict(nb=32, kmesh=(6, 6, 6), xmesh=(4, 4, 4), lc=0, seed=25),
    dict(nb=12, kmesh=(3, 3, 3), xmesh=(3, 3, 3), lc=0, seed=26, w_scale=1.3),
]


@pytest.fixture(scope="module")
def members():
    return [synth.make_inputs(**m) for m in MEMBERS]


def _open(inps, precision=core.FP64):
    return [core.OceanGPU.from_inputs(i, precision=precision) for i in inps]


@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_batch_haydock_reproduces_single_recursions(members, precision):
    print("batch haydock, precision", precision, flush=True)
    gs = _open(members, precision)
    niter = 10  # three captured cycles plus one direct iteration
    single = [g.haydock(i.psi0, niter) for g, i in zip(gs, members)]
    bt = core.OceanBatch(gs)
    got, ms = bt.haydock([i.psi0 for i in members], niter)
    print("groups", bt.groups(), "ms", ms, flush=True)
    for q, ((a, b, ai), (ra, rb, rai)) in enumerate(zip(got, single)):
        assert np.array_equal(a, ra) and np.array_equal(b, rb) and np.array_equal(ai, rai), q
    bt.close()
    for g in gs:
        g.close()


def test_batch_after_autotune_matches_oracle(members):
    gs = _open(members)
    bt = core.OceanBatch(gs)
    groups = bt.autotune(reps=2)
    print("tuned groups", groups, flush=True)
    rng = np.random.default_rng(5)
    psis, vin, vout = [], [], []
    for g in gs:
        z = np.asfortranarray(rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
        psis.append(z)
        vin.append(g.vec_new(z))
        vout.append(g.vec_new())
    ms = bt.xact(vin, vout, nrep=3)
    assert ms > 0
    for g, i, z, vo in zip(gs, members, psis, vout):
        ref = orc.xact(to_problem(i), z)
        assert relerr(g.vec_get(vo), ref) < 1e-12
    niter = 24
    got, _ = bt.haydock([i.psi0 for i in members], niter)
    for (a, b, _), i in zip(got, members):
        ra, rb, _, ierr = orc.haydock(to_problem(i), i.psi0, niter)
        assert ierr == 0
        assert np.max(np.abs(a[:niter] - ra[:niter]) / np.abs(ra[:niter])) < 1e-10
        assert np.max(np.abs(b[1:niter + 1] - rb[1:niter + 1]) / np.abs(rb[1:niter + 1])) < 1e-10
    bt.close()
    for g in gs:
        g.close()


def test_load_bloch_dev_equals_host_load(members):
    """u handed over in device memory gives the same action as the host load."""
    import ctypes as C
    inp = members[0]
    g1 = core.OceanGPU.from_inputs(inp)
    g2 = core.OceanGPU.from_inputs(inp)
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library already loaded
    planes = [np.ascontiguousarray(inp.u.real.ravel(order="F")),
              np.ascontiguousarray(inp.u.imag.ravel(order="F"))]
    ptrs = []
    for a in planes:
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        ptrs.append(p)
    g2.load_bloch_dev(ptrs[0].value, ptrs[1].value)
    for p in ptrs:
        hip.hipFree(p)
    z = inp.psi0
    assert np.array_equal(g1.xact(z), g2.xact(z))
    g1.close()
    g2.close()


def test_batch_gmres_matches_single_solves_and_oracle(members):
    """ocean_gpu_batch_gmres: every problem's recycled GMRES (own energies, shared
    settings) run together gives the stand-alone ocean_gpu_gmres rows bit for bit (the
    same kernels on each problem's own stream), and the oracle's spectrum."""
    gs = _open(members)
    E = [np.linspace(-2.0 + q, 30.0 - q, 4 + q % 3) for q in range(len(members))]
    single = [g.gmres(i.psi0, e, 0.5, 13.605, 1e-14, 20) for g, i, e in zip(gs, members, E)]
    bt = core.OceanBatch(gs)
    rows, ms = bt.gmres([i.psi0 for i in members], E, 0.5, 13.605, 1e-14, 20)
    print("batched GMRES ms", ms, flush=True)
    for q, (r, s1) in enumerate(zip(rows, single)):
        assert np.array_equal(r, s1), q
    for q in (0, 2, 3):
        i = members[q]
        ref = orc.gmres_do_recycle(to_problem(i), i.psi0, E[q], 0.5, 13.605, 1e-14, 20, i.kpref)
        pk = np.abs(ref[:, 2]).max()
        assert np.abs(rows[q][:, 2] - ref[:, 2]).max() / pk < 1e-8, q
    # per-problem codes: a starting vector holding a NaN never meets the stopping test,
    # so that problem stops with 1005 after its 50 restarts while the others finish
    starts = [i.psi0.copy() for i in members]
    starts[1][0, 0, 0] = np.nan
    rows2, _, rcs = bt.gmres(starts, [e[:1] for e in E], 0.5, 13.605, 1e-14, 20, return_status=True)
    assert rcs[1] == 1005 and all(rcs[q] == 0 for q in range(len(members)) if q != 1), list(rcs)
    bt.close()
    for g in gs:
        g.close()


def test_single_problem_autotune_keeps_parity(members):
    """ocean_gpu_autotune picks one problem's tiles with the batch autotuner; the
    action with the picked tiles still matches the oracle (FP64 and mixed)."""
    for q in (0, 2):
        i = members[q]
        P = to_problem(i)
        for prec, tol in ((core.FP64, 1e-12), (core.MIXED, 1e-5)):
            with core.OceanGPU.from_inputs(i, precision=prec) as g:
                before = g.tiles()
                after = g.autotune(reps=2)
                print("member", q, "precision", prec, "tiles", before, "->", after, flush=True)
                assert after[1] >= i.nb and after[0] in (64, 128, 256, 512, 1024)
                assert relerr(g.xact(i.psi0), orc.xact(P, i.psi0)) < tol
