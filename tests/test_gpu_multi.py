"""Right-hand sides that share one Hamiltonian (ocean_gpu_create_multi): each block of
the stacked H.psi against the oracle's H.psi of that block's own vector on the same
inputs, in both precisions and directions and for the term subsets; the stacked
Haydock recursions against the oracle's recursion of each vector and against the
one-right-hand-side GPU recursion; a NaN in one right-hand side; the refusals.

Tolerances as for one right-hand side: FP64 H.psi 1e-12 relative, mixed 1e-5 of the
largest element, Haydock a/b 1e-10 relative over 50 steps, spectra 1e-8 of peak."""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import relerr, to_problem

pytestmark = pytest.mark.gpu

from ocean_amd import core, setup_host as sh, synth  # noqa: E402

E2H = orc.EV2HARTREE

# (make_inputs arguments, right-hand sides, random spin-orbit matrix and allow mask)
CASES = {
    # K edge, three polarisations: 12 alphas on one spin (the many-alpha kernels:
    # matrix-core projection and long-range sums, the staged spin-orbit epilogue)
    "k_r3": (dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 4), lc=0, seed=21), 3, True),
    # odd band and x counts: row tails of every kernel
    "k_r3_ragged": (dict(nb=23, kmesh=(2, 3, 3), xmesh=(5, 3, 3), lc=0, seed=22), 3, True),
    # nx 512 (several projection x-tiles) and nb 72 (two 64-band blocks of the
    # matrix-core long-range sums)
    "k_r3_wide": (dict(nb=72, kmesh=(2, 2, 2), xmesh=(8, 8, 8), lc=0, seed=23), 3, True),
    # two spins: 6 and 12 alphas per spin
    "k_nspn2_r3": (dict(nb=8, kmesh=(2, 2, 4), xmesh=(3, 3, 3), lc=0, nspn=2, seed=24), 3, True),
    "k_nspn2_r6": (dict(nb=8, kmesh=(2, 2, 4), xmesh=(3, 3, 3), lc=0, nspn=2, seed=25), 6, False),
    # L2,3 with its own spin-orbit term, two polarisations on two spins
    "l23_nspn2_r2": (dict(nb=8, kmesh=(2, 2, 4), xmesh=(3, 3, 3), lc=1, nspn=2, seed=26, xi_eV=3.0), 2, False),
    # 24 alphas on one spin: six K-edge polarisations, two L2,3 ones (three column
    # tiles in both matrix-core products)
    "k_r6": (dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 4), lc=0, seed=27), 6, True),
    "k_r6_wide": (dict(nb=72, kmesh=(2, 2, 2), xmesh=(8, 8, 8), lc=0, seed=28), 6, True),
    "l23_r2": (dict(nb=13, kmesh=(3, 2, 2), xmesh=(3, 3, 5), lc=1, seed=29, xi_eV=4.0), 2, False),
}


def _inputs(name):
    kw, R, so = CASES[name]
    inp = synth.make_inputs(**kw)
    if so:
        # the K edge has no spin-orbit term: a random Hermitian S over its four alphas
        # exercises the per-right-hand-side mixing; a sparse allow mask the replication
        rng = np.random.default_rng(kw["seed"] + 100)
        na = inp.energy.shape[2]
        s = rng.standard_normal((na, na)) + 1j * rng.standard_normal((na, na))
        s = (s + s.conj().T) * 1e-3
        inp.mult.somelr = np.asfortranarray(s.real)
        inp.mult.someli = np.asfortranarray(s.imag)
        inp.mult.have_spin_orbit = True
        inp.allow = np.asfortranarray((rng.uniform(size=inp.allow.shape) > 0.2).astype(np.float64))
    return inp, R


def _vecs(R, shape, seed):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(R)]


@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
@pytest.mark.parametrize("name", list(CASES))
def test_multi_xact_blocks_match_oracle(name, precision):
    inp, R = _inputs(name)
    print("multi xact", name, "precision", precision, flush=True)
    P = to_problem(inp)
    psis = _vecs(R, inp.energy.shape, 7)
    na0 = inp.energy.shape[2]
    tol = 1e-12 if precision == core.FP64 else 1e-5
    with core.OceanGPU.from_inputs(inp, precision=precision, nrhs=R) as g:
        assert g.shape == (inp.nb, inp.nk, R * na0)
        stacked = np.concatenate(psis, axis=2)
        runs = [dict(backwards=False), dict(backwards=True)]
        if precision == core.FP64:
            runs += [dict(hflag=hf) for hf in ((1, 0, 0, 0, 0, 0), (0, 0, 0, 0, 1, 0), (0, 0, 0, 0, 0, 1))]
        for kw in runs:
            got = g.xact(stacked, **kw)
            for r in range(R):
                ref = orc.xact(P, psis[r], **kw)
                err = relerr(got[:, :, r * na0:(r + 1) * na0], ref)
                assert err < tol, (name, kw, r, err)


def _psi0s(inp, R, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(R):
        z = np.asfortranarray(rng.standard_normal(inp.energy.shape) + 1j * rng.standard_normal(inp.energy.shape))
        out.append(sh.normalise_psi0(z, inp.nk, inp.celvol)[0])
    return out


@pytest.mark.parametrize("name", ["k_r3", "k_r6"])
def test_multi_haydock_matches_oracle_and_single_recursions(name):
    inp, R = _inputs(name)
    P = to_problem(inp)
    n = 50
    psi0s = _psi0s(inp, R, 9)
    with core.OceanGPU.from_inputs(inp, nrhs=R) as g:
        a, b, ai = g.haydock_multi(psi0s, n)
    with core.OceanGPU.from_inputs(inp) as g1:
        single = [g1.haydock(p, n) for p in psi0s]
    gam, el, eh, ne = 0.5 * E2H, -5.0 * E2H, 40.0 * E2H, 500
    for r in range(R):
        ra, rb, _, ierr = orc.haydock(P, psi0s[r], n)
        assert ierr == 0
        ea = np.max(np.abs(a[r, :n] - ra[:n]) / np.abs(ra[:n]))
        eb = np.max(np.abs(b[r, 1:n + 1] - rb[1:n + 1]) / np.abs(rb[1:n + 1]))
        print("rhs", r, "a", ea, "b", eb, flush=True)
        assert ea < 1e-10 and eb < 1e-10
        sa, sb, _ = single[r]
        assert np.max(np.abs(a[r, :n] - sa[:n]) / np.abs(sa[:n])) < 1e-10
        assert np.max(np.abs(b[r, 1:n + 1] - sb[1:n + 1]) / np.abs(sb[1:n + 1])) < 1e-10
        got = core.spectrum(a[r], b[r], n, 1.0, gam, el, eh, ne)[:, 2]
        ref = orc.write_core(ra, rb, n, 1.0, gam, el, eh, ne)[:, 2]
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-8


def test_multi_haydock_mixed_spectra_close_to_oracle():
    inp, R = _inputs("k_r3_wide")
    P = to_problem(inp)
    n = 40
    psi0s = _psi0s(inp, R, 10)
    with core.OceanGPU.from_inputs(inp, precision=core.MIXED, nrhs=R) as g:
        a, b, _ = g.haydock_multi(psi0s, n)
    gam, el, eh, ne = 0.5 * E2H, -5.0 * E2H, 40.0 * E2H, 500
    for r in range(R):
        ra, rb, _, _ = orc.haydock(P, psi0s[r], n)
        got = core.spectrum(a[r], b[r], n, 1.0, gam, el, eh, ne)[:, 2]
        ref = orc.write_core(ra, rb, n, 1.0, gam, el, eh, ne)[:, 2]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("mixed rhs", r, "spectrum err of peak", err, flush=True)
        assert err < 1e-4


def test_multi_haydock_nan_in_one_right_hand_side():
    """A NaN in one right-hand side's start stops that recursion's record at its first
    step (rc -1, as OCEAN_hay_ab_twoterm, OCEAN2/OCEAN_haydock.f90:912-920); the
    others run to niter, their coefficients unchanged, since the action mixes no
    alphas across right-hand sides."""
    inp, R = _inputs("k_r3")
    psi0s = _psi0s(inp, R, 11)
    with core.OceanGPU.from_inputs(inp, nrhs=R) as g:
        a_ok, b_ok, _ = g.haydock_multi(psi0s, 12)
        bad = [p.copy() for p in psi0s]
        bad[1][0, 0, 0] = np.nan
        a, b, _, done, rc = g.haydock_multi(bad, 12, return_status=True)
    assert rc == -1
    assert list(done) == [12, 1, 12]
    for r in (0, 2):
        assert np.array_equal(a[r], a_ok[r]) and np.array_equal(b[r], b_ok[r])


def test_multi_refusals():
    kw = dict(nb=4, kmesh=(2, 2, 2), xmesh=(2, 2, 2), lc=0, seed=3)
    inp = synth.make_inputs(**kw)
    # 8 and 16 alphas on one spin, 24 per spin on two spins: no kernel instance
    for nrhs, nspn in ((2, 1), (4, 1), (12, 2)):
        inp_s = inp if nspn == 1 else synth.make_inputs(**dict(kw, nspn=2))
        with pytest.raises(core.OceanGPUError) as e:
            core.OceanGPU.from_inputs(inp_s, nrhs=nrhs)
        assert e.value.code == 9963
    with core.OceanGPU.from_inputs(inp, nrhs=3) as g:
        rc = g.haydock(np.concatenate([inp.psi0] * 3, axis=2), 4, return_status=True)[-1]
        assert rc == 9963
        z = np.asfortranarray(np.ones(g.shape, complex))
        with pytest.raises(core.OceanGPUError):
            g.gmres(z, [1.0], 0.5, 13.6, 1e-10, 8)
        bt = core.OceanBatch([g])  # batched actions take it; batched recursions do not
        with pytest.raises(core.OceanGPUError) as e:
            bt.haydock([z], 4)
        assert e.value.code == 9963
        bt.close()
