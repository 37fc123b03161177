"""GPU (HIP C-ABI) vs the oracle on identical seeded inputs.

Tolerances: the FP64 drop-in is held to 1e-12 relative on H.psi, Haydock a/b to
1e-10 relative over the first 50 steps, spectra to 1e-8 of peak.
"""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import relerr, to_problem

pytestmark = pytest.mark.gpu

from ocean_amd import synth  # noqa: E402

CASES = {
    "k_tiny": dict(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=3),
    "k_odd": dict(nb=16, kmesh=(3, 4, 2), xmesh=(4, 3, 5), lc=0, seed=4),
    "l23_so": dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 4), lc=1, seed=5, xi_eV=6.0),
    "k_6cube": dict(nb=32, kmesh=(6, 6, 6), xmesh=(6, 6, 6), lc=0, seed=6),
    "nspn2": dict(nb=8, kmesh=(2, 2, 4), xmesh=(3, 3, 3), lc=1, seed=7, nspn=2, xi_eV=3.0),
    # odd band count and odd x count: the row tails of both projection forms, the
    # padded even row stride of the FP32 slab, the last partial 32-point MFMA tile
    "l23_ragged": dict(nb=13, kmesh=(3, 2, 2), xmesh=(3, 3, 5), lc=1, seed=8, xi_eV=4.0),
    "k_ragged": dict(nb=23, kmesh=(2, 3, 3), xmesh=(5, 3, 3), lc=0, seed=9),
}


def _gpu(inp, **kw):
    from ocean_amd import core
    return core.OceanGPU.from_inputs(inp, **kw)


@pytest.fixture(scope="module")
def k16():
    return synth.workload("k_nb16")


@pytest.mark.parametrize("name", list(CASES))
def test_xact_matches_oracle(name):
    inp = synth.make_inputs(**CASES[name])
    print("case", name, flush=True)
    P = to_problem(inp)
    g = _gpu(inp)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)
    psi = np.asfortranarray(psi)
    for back in (False, True):
        ref = orc.xact(P, psi, backwards=back)
        got = g.xact(psi, backwards=back)
        assert relerr(got, ref) < 1e-12, (name, back, relerr(got, ref))
    # individual terms through hflag
    for hf in ((1, 0, 0, 0, 0, 0), (0, 0, 0, 0, 1, 0), (0, 0, 0, 0, 0, 1)):
        ref = orc.xact(P, psi, hflag=hf)
        got = g.xact(psi, hflag=hf)
        assert relerr(got, ref) < 1e-12, (name, hf)
    g.close()


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_xact_close_to_oracle(name):
    """The mixed form (FP32 u, FFT and W; FP32 matrix cores for the many-alpha
    projection) against the FP64 oracle: H.psi to 1e-5 of its largest element, both
    directions (FP32 rounding of u and of sums over up to a few hundred terms)."""
    inp = synth.make_inputs(**CASES[name])
    P = to_problem(inp)
    from ocean_amd import core
    print("mixed xact case", name, flush=True)
    with _gpu(inp, precision=core.MIXED) as g:
        rng = np.random.default_rng(11)
        psi = np.asfortranarray(rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
        for back in (False, True):
            got = g.xact(psi, backwards=back)
            ref = orc.xact(P, psi, backwards=back)
            print("  backwards", back, "relerr", relerr(got, ref), flush=True)
            assert relerr(got, ref) < 1e-5, (name, back, relerr(got, ref))


def test_xact_k_nb16(k16):
    P = to_problem(k16)
    g = _gpu(k16)
    ref = orc.xact(P, k16.psi0)
    got = g.xact(k16.psi0)
    assert relerr(got, ref) < 1e-12
    g.close()


def test_haydock_k_nb16(k16):
    P = to_problem(k16)
    g = _gpu(k16)
    niter = k16.params["niter"]
    a_ref, b_ref, ia_ref, ierr = orc.haydock(P, k16.psi0, niter)
    a, b, ai = g.haydock(k16.psi0, niter)
    assert ierr == 0
    ra = np.abs(a[:50] - a_ref[:50]) / np.abs(a_ref[:50])
    rb = np.abs(b[1:51] - b_ref[1:51]) / np.abs(b_ref[1:51])
    print("max rel a", ra.max(), "b", rb.max())
    assert ra.max() < 1e-10 and rb.max() < 1e-10
    # spectra (literal write_core) within 1e-8 of peak
    p = k16.params
    E2H = orc.EV2HARTREE
    sp_ref = orc.write_core(a_ref, b_ref, niter, k16.kpref, p["gam"] * E2H, p["el"] * E2H,
                            p["eh"] * E2H, p["ne"])
    from ocean_amd import core
    sp = core.spectrum(a, b, niter, k16.kpref, p["gam"] * E2H, p["el"] * E2H, p["eh"] * E2H, p["ne"])
    peak = np.abs(sp_ref[:, 2]).max()
    err = np.abs(sp[:, 2] - sp_ref[:, 2]).max() / peak
    print("spectrum err/peak", err)
    assert err < 1e-8
    g.close()


def test_spectrum_kernel_matches_oracle():
    from ocean_amd import core
    rng = np.random.default_rng(2)
    n = 80
    a = rng.uniform(0.2, 1.0, n + 1)
    b = rng.uniform(0.1, 0.5, n + 1)
    b[0] = 0
    for lit in (True, False):
        ref = orc.write_core(a, b, n, 1.3e-3, 0.02, -0.2, 1.5, 1500, literal_cmplx=lit)
        got = core.spectrum(a, b, n, 1.3e-3, 0.02, -0.2, 1.5, 1500, literal_cmplx=lit)
        pk = np.abs(ref[:, 2]).max()
        assert np.abs(got[:, 2] - ref[:, 2]).max() / pk < 1e-12
        assert np.abs(got[:, 3] - ref[:, 3]).max() / np.abs(ref[:, 3]).max() < 1e-12
        assert np.allclose(got[:, 0], ref[:, 0], rtol=0, atol=1e-12)


def _fe_lanc(site="01"):
    import os
    gold = os.path.join(os.path.dirname(__file__), "golden")
    lines = open(os.path.join(gold, f"abslanc_Fe.0001_1s_{site}")).read().split("\n")
    n, kpref = int(lines[0].split()[0]), float(lines[0].split()[1])
    a, b = np.zeros(n + 1), np.zeros(n + 1)
    a[0] = float(lines[1])
    for i in range(1, n + 1):
        a[i], b[i] = (float(t) for t in lines[1 + i].split()[:2])
    eig = np.array([float(l.split()[2]) for l in lines[n + 2:] if l.strip()])
    return n, kpref, a, b, eig


FE_NVAL = 16.0 / 5.423513587 ** 3
FE_EPS = 1.0e5


def test_write_core_nine_columns_matches_oracle():
    """ocean_gpu_write_core vs write_core_full: Spect(0)/SPKK(0) and the cross sections
    to 1e-12 of peak, the gamfcn-broadened columns to 1e-9 (device atan/log)."""
    from ocean_amd import core
    n, kpref, a, b, _ = _fe_lanc()
    for old in (False, True):
        for lit in (True, False):
            kw = dict(ebase_eV=7112.0, old_broaden=old, nval=FE_NVAL, eps=FE_EPS, ucvol=159.5,
                      literal_cmplx=lit)
            ref = orc.write_core_full(a, b, n, kpref, 0.0184, -1.5, 2.5, 400, **kw)
            got = core.write_core(a, b, n, kpref, 0.0184, -1.5, 2.5, 400, **kw)
            assert got.shape == ref.shape == (401, 9)
            assert np.allclose(got[:, 0], ref[:, 0], rtol=1e-15, atol=0)
            for c in (2, 4, 5, 6, 7, 8):
                assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-12 * np.abs(ref[:, c]).max(), (old, lit, c)
            for c in (1, 3):
                assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-9 * np.abs(ref[:, c]).max(), (old, lit, c)


def test_write_core_voigt_matches_oracle():
    """The Voigt form (fine grid reused along the output energies in the reference)
    evaluated independently per output energy on the device."""
    from ocean_amd import core
    n, kpref, a, b, _ = _fe_lanc("02")
    for (el, eh, ne, sig) in ((-0.3, 1.6, 57, 0.012), (-0.1, 0.2, 30, 0.05)):
        kw = dict(ebase_eV=1.0, nval=FE_NVAL, eps=FE_EPS, gam_gauss=sig)
        ref = orc.write_core_full(a, b, n, kpref, 0.0184, el, eh, ne, **kw)
        got = core.write_core(a, b, n, kpref, 0.0184, el, eh, ne, **kw)
        assert got.shape == ref.shape == (ne + 1, 5)
        assert np.allclose(got[:, 0], ref[:, 0], rtol=1e-15, atol=1e-12)
        for c in (2, 4):
            assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-11 * np.abs(ref[:, c]).max(), c
        for c in (1, 3):
            assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-9 * np.abs(ref[:, c]).max(), c


def test_tridiag_eigs_match_fixture_block():
    """Sturm bisection on the device vs the abslanc eigenvalue block (elsch) and LAPACK."""
    from ocean_amd import core
    for site in ("01", "02"):
        n, _, a, b, eig = _fe_lanc(site)
        w = core.tridiag_eigs(a, b, n)
        assert np.abs(w - np.sort(eig)).max() < 1e-8
        assert np.abs(w - orc.redtrid_eigs(a, b, n)).max() < 1e-12
    # a zero coupling splits the matrix; repeated eigenvalues come out twice
    a = np.array([1.0, 1.0, 0.5, 1.0])
    b = np.array([0.0, 0.0, 0.3, 0.0])
    w = core.tridiag_eigs(a, b, 3)
    assert np.abs(w - orc.redtrid_eigs(a, b, 3)).max() < 1e-14


def test_vector_ops(k16):
    g = _gpu(k16)
    rng = np.random.default_rng(5)
    p = np.asfortranarray(rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
    q = np.asfortranarray(rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
    vp, vq = g.vec_new(p), g.vec_new(q)
    r, i = g.vec_dot(vp, vq)
    rr, ii = orc.psi_dot(p, q)
    assert abs(r - rr) < 1e-10 * abs(rr) and abs(i - ii) < 1e-10 * max(abs(ii), 1)
    assert abs(g.vec_nrm(vp) - orc.psi_nrm(p)) < 1e-12 * orc.psi_nrm(p)
    g.vec_axpy(0.3 - 0.7j, vp, vq)
    assert relerr(g.vec_get(vq), q + (0.3 - 0.7j) * p) < 1e-14
    g.vec_scal(2.5, vp)
    assert relerr(g.vec_get(vp), 2.5 * p) < 1e-15
    vo = g.vec_new()
    g.xact_vec(vp, vo)
    assert relerr(g.vec_get(vo), orc.xact(to_problem(k16), 2.5 * p)) < 1e-12
    # axmy / axmz / divide in the reference's operation order (device FMA contraction aside)
    P2, Q2 = g.vec_get(vp), g.vec_get(vq)
    vy = g.vec_new(Q2)
    g.vec_axmy(0.7 - 1.3j, vp, vy)
    assert relerr(g.vec_get(vy), orc.psi_axmz(P2, Q2, 0.7, -1.3)) < 1e-15
    vz = g.vec_new()
    g.vec_axmz(-0.2 + 0.9j, vp, vz, vq)
    assert relerr(g.vec_get(vz), orc.psi_axmz(P2, Q2, -0.2, 0.9)) < 1e-15
    for kw in (dict(rval=1.7), dict(ival=-0.6), dict(rval=0.4, ival=2.2)):
        vd = g.vec_new(P2)
        g.vec_divide(vd, **kw)
        assert relerr(g.vec_get(vd), orc.psi_divide(P2, **kw)) < 1e-15, kw
    g.close()


def test_gmres_small():
    inp = synth.make_inputs(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=9)
    P = to_problem(inp)
    g = _gpu(inp)
    E = np.linspace(-2.0, 30.0, 6)
    ref = orc.gmres_do_recycle(P, inp.psi0, E, 0.5, 13.605, 1e-10, 20, inp.kpref)
    got = g.gmres(inp.psi0, E, 0.5, 13.605, 1e-10, 20)
    print(ref[:, :3], got[:, :3])
    pk = np.abs(ref[:, 2]).max()
    assert np.abs(got[:, 2] - ref[:, 2]).max() / pk < 1e-8
    assert np.array_equal(got[:, 4], ref[:, 4])
    g.close()


def test_haydock_early_exit(k16):
    """The device loop stops where the reference's early exit would (replayed by the
    oracle over the same coefficients) and keeps the coefficients up to there."""
    g = _gpu(k16)
    niter, spacing, ne = k16.params["niter"], 10, 500
    E2H = orc.EV2HARTREE
    gam, el, eh = 0.5 * E2H, -5.0 * E2H, 40.0 * E2H
    a, b, ai = g.haydock(k16.psi0, niter)
    rel = []
    for it in range(spacing + 1, niter + 1):
        s1 = orc.calc_spect_core(a, b, it, k16.kpref, gam, el, eh, ne, ai)[2]
        s2 = orc.calc_spect_core(a, b, it - spacing, k16.kpref, gam, el, eh, ne, ai)[2]
        rel.append(2 * np.abs(s1 - s2).sum() / (np.abs(s1).sum() + np.abs(s2).sum()))
    rel = np.array(rel)
    # a threshold well away from every relArea value that stops the loop mid-way
    cands = np.sort(rel)
    thr = None
    for lo, hi in zip(cands[:-1], cands[1:]):
        t = np.sqrt(lo * hi)
        if hi / lo < 1.01:
            continue
        stop = orc.haydock_early_exit(a, b, niter, spacing, t, k16.kpref, gam, el, eh, ne, ai)
        if spacing + 2 < stop < niter - 10:
            thr = t
            break
    assert thr is not None
    ae, be, aie, done = g.haydock_early(k16.psi0, niter, spacing, thr, gam, el, eh, ne)
    print("threshold", thr, "stop", stop, "device", done, flush=True)
    assert done == stop
    assert np.array_equal(ae[:done], a[:done]) and np.array_equal(be[:done + 1], b[:done + 1])
    assert not ae[done:].any() and not be[done + 1:].any()
    g.close()


def test_ospect_kernels_match_oracle():
    """ocean_gpu_ospect (core with and without the anti-resonant term, and the valence
    branch) vs the oracle's o_spect restatement on the Fe Lanczos coefficients."""
    from ocean_amd import core
    n, kpref, a, b, _ = _fe_lanc("01")
    E2H = orc.EV2HARTREE
    for ebase in (0.0, 3.0):
        got = core.ospect(a, b, n, kpref, 300, -20.0, 60.0, 0.5, ebase_eV=ebase, eps0=FE_EPS, nval=FE_NVAL)
        ref = orc.o_spect_core(a, b, n, kpref, 300, -20.0 * E2H, 60.0 * E2H, 0.5 * E2H, ebase * E2H, FE_NVAL,
                               FE_EPS)
        assert got.shape == ref.shape == (301, 5)
        assert np.allclose(got[:, 0], ref[:, 0], rtol=1e-15, atol=1e-12)
        for c in (2,):
            assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-12 * np.abs(ref[:, c]).max(), (ebase, c)
        for c in (1, 3, 4):
            assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-9 * np.abs(ref[:, c]).max(), (ebase, c)
    got = core.ospect(a, b, 60, kpref, 200, 1.0, 60.0, 0.3, ucvol=159.5, is_val=True)
    ref = orc.o_spect_val(a, b, 60, kpref, 159.5, 200, 1.0 * E2H, 60.0 * E2H, 0.3 * E2H)
    assert got.shape == ref.shape == (200, 8)
    _check_opcons(got, ref)


def _check_opcons(got, ref):
    # n and kappa come from acos(eps1/|eps|)/2, which loses digits where eps is nearly
    # real, so they are held to 1e-11 of |N| = sqrt|eps| rather than of themselves
    absN = np.hypot(ref[:, 3], ref[:, 4]).max()
    for c in range(8):
        scale = absN if c in (3, 4) else np.abs(ref[:, c]).max()
        assert np.abs(got[:, c] - ref[:, c]).max() <= 1e-11 * scale, c


def test_fortran_ospect_run_directory(tmp_path):
    """ocean_gpu_ospect.x in a run directory holding the Fe abslanc files: both runs of
    runlist, the photon-weighted TotAvg, and an opcons run from oplanc."""
    import os
    import shutil
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "ocean_amd", "fortran", "ocean_gpu_ospect.x")
    if not os.path.exists(exe):
        pytest.fail("ocean_gpu_ospect.x not built (run __graft_entry__.build())")
    gold = os.path.join(os.path.dirname(__file__), "golden")
    for site in ("01", "02"):
        shutil.copy(os.path.join(gold, f"abslanc_Fe.0001_1s_{site}"), tmp_path)
    shutil.copy(os.path.join(gold, "abslanc_Fe.0001_1s_01"), tmp_path / "oplanc")
    (tmp_path / "spect.in").write_text("400 -20.0 60.0 0.5 0.0\n")
    (tmp_path / "epsilon").write_text("100000\n")
    (tmp_path / "nval.h").write_text(f"{FE_NVAL!r}\n")
    acell = 5.423513587
    (tmp_path / "avecsinbohr.ipt").write_text(f"{acell} 0 0\n0 {acell} 0\n0 0 {acell}\n")
    (tmp_path / "runlist").write_text("3\n26 1 0 Fe 1s 1 1 XAS\n26 1 0 Fe 1s 1 2 XAS\n26 1 0 Fe 1s 1 1 VAL\n")
    (tmp_path / "do_avg.ipt").write_text(".true.\n")
    for ph, w in ((1, 0.25), (2, 0.75)):
        (tmp_path / f"photon{ph}").write_text(f"quad\ncartesian 1 0 0\nend\ncartesian 0 1 0\nend\n8990\n{w}\n")
    print("fortran o_spect case: Fe abslanc x2 + oplanc, ne=400", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    E2H = orc.EV2HARTREE
    tot = 0
    for site, w in (("01", 0.25), ("02", 0.75)):
        n, kpref, a, b, _ = _fe_lanc(site)
        ref = orc.o_spect_core(a, b, n, kpref, 400, -20.0 * E2H, 60.0 * E2H, 0.5 * E2H, 0.0, FE_NVAL, FE_EPS)
        txt = np.loadtxt(tmp_path / f"absspct_Fe.0001_1s_{site}", comments="#")
        assert txt.shape == (401, 5)
        for c in range(1, 5):  # eight printed digits
            assert np.abs(txt[:, c] - ref[:, c]).max() <= 1e-7 * np.abs(ref[:, c]).max(), (site, c)
        tot = tot + w * ref[:, [4, 2, 3, 1]]
    avg = np.loadtxt(tmp_path / "TotAvg")
    assert avg.shape == (400, 5)
    assert np.abs(avg[:, 1:] - tot[:400]).max() <= 1e-9 * np.abs(tot).max()  # gamfcn columns: device atan/log
    n, kpref, a, b, _ = _fe_lanc("01")
    ref = orc.o_spect_val(a, b, n, kpref, acell ** 3, 400, -20.0 * E2H, 60.0 * E2H, 0.5 * E2H)
    txt = np.loadtxt(tmp_path / "opcons", comments="#")
    assert txt.shape == (400, 8)
    _check_opcons(txt, ref)


@pytest.mark.parametrize("lc,lvl,lvh,nproj,xi,lt", [(0, 0, 3, (1, 1, 1, 1), 0.0, None),
                                                     (1, 0, 2, (2, 1, 2), 6.0, None),
                                                     (1, 1, 3, (1, 2, 1), 3.0, (0.3, 0.55)),
                                                     (0, 0, 1, (2, 1), 0.0, (0.4, 0.1)),
                                                     (1, 0, 2, (1, 2, 1), 2.0, "tdlda")])
def test_mult_setup_matches_oracle(lc, lvl, lvh, nproj, xi, lt):
    """ocean_gpu_mult_setup (nbsemhsetup2 blocks and soprep on the device) against the
    oracle's quadrature restatement of nbsemhsetup2 and its jimel-based soprep."""
    from ocean_amd import core
    g = core.OceanGPU(4, (2, 2, 2), (2, 2, 2), lc, 1, 100.0, 1.0, lvl, lvh, nproj, lvl, lvh)
    tdlda = lt == "tdlda"
    lt = None if tdlda else lt
    rng = np.random.default_rng(31 + lc + lvl)
    slater, refs = [], []
    for l in range(lvl, lvh + 1):
        npl = nproj[l - lvl]
        fk, gk, sf, sg = {}, {}, {}, {}
        for k in ([0] if tdlda else range(2, min(2 * lc, 2 * l) + 1, 2)):
            m = rng.uniform(0.5, 2.0, (npl, npl))
            fk[k], sf[k] = 0.5 * (m + m.T), rng.uniform(0.7, 0.9)
        for k in range(abs(lc - l), lc + l + 1, 2):
            m = rng.uniform(0.5, 2.0, (npl, npl))
            gk[k], sg[k] = 0.5 * (m + m.T), rng.uniform(0.7, 0.9)
        slater.append(({k: fk[k] * sf[k] for k in fk}, {k: gk[k] * sg[k] for k in gk}))
        refs.append(orc.nbsemhsetup2(lc, l, npl, fk, sf, gk, sg, scfac=0.95, tdlda=tdlda) * orc.EV2HARTREE)
    X, sr, si, have_so = g.mult_setup(slater, 0.95, xi, lt, tdlda)
    for got, ref in zip(X, refs):
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), np.abs(got - ref).max()
    rsr, rsi, rhave = orc.soprep(lc, xi, lt)
    assert np.abs(sr - rsr).max() <= 1e-16 + 1e-14 * np.abs(rsr).max()
    assert np.abs(si - rsi).max() <= 1e-16 + 1e-14 * np.abs(rsi).max() and have_so == rhave
    g.close()
