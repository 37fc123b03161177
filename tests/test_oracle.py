"""Pin the oracle: reference fixtures, closed forms and brute force (CPU only).

Each test names what it pins.  Retyping the oracle's formula would pin nothing,
so every check compares against an independent quantity: the reference's own
stored outputs (TEST/Fe/Results), a closed form, a dense matrix built another
way, or a second implementation (the C restatement, setup_host's closed forms).
"""
import math
import os

import numpy as np
import pytest

import ocean_oracle as orc
from helpers import relerr, to_problem
from ocean_amd import setup_host as sh
from ocean_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- fixtures
def load_lanc(fn):
    lines = open(fn).read().split("\n")
    n, kpref = lines[0].split()
    n = int(n)
    a = np.zeros(n + 1)
    b = np.zeros(n + 1)
    a[0] = float(lines[1])
    for i in range(1, n + 1):
        x = lines[1 + i].split()
        a[i], b[i] = float(x[0]), float(x[1])
    eig = np.array([float(l.split()[2]) for l in lines[n + 2:] if l.strip()])
    return n, float(kpref), a, b, eig


def load_spct(fn):
    lines = open(fn).read().split("\n")
    gam = float(lines[0].split("gam=")[1].split()[0])
    d = np.array([[float(t) for t in l.split()] for l in lines[2:] if l.strip()])
    return gam, d


@pytest.mark.parametrize("site", ["01", "02"])
@pytest.mark.parametrize("literal", [True, False])
def test_write_core_reproduces_fe_fixture(site, literal):
    """write_core (OCEAN_haydock.f90:1801-1894) against TEST/Fe/Results abslanc -> absspct."""
    n, kpref, a, b, _ = load_lanc(os.path.join(GOLD, f"abslanc_Fe.0001_1s_{site}"))
    gam, d = load_spct(os.path.join(GOLD, f"absspct_Fe.0001_1s_{site}"))
    h2e = orc.HARTREE2EV_OLD  # the fixture predates the 2022 CODATA constant
    ne = d.shape[0] - 1
    el, eh = d[0, 0] / h2e, d[-1, 0] / h2e
    out = orc.write_core(a, b, n, kpref, gam, el, eh, ne, literal_cmplx=literal, hartree2ev=h2e)
    pk = np.abs(d[:, 2]).max()
    # limited by the fixture's printed precision (10-decimal a/b, 8-digit columns)
    assert np.abs(out[:, 2] - d[:, 2]).max() / pk < 1.5e-6
    # columns 2 and 4 (Spect, SPKK) of this old fixture carry the gamfcn
    # energy-dependent broadening (jdamp = 1); the Gamma-only columns are 3 and 5
    assert np.abs(out[:, 4] - d[:, 4]).max() / np.abs(d[:, 4]).max() < 1.5e-6
    assert np.allclose(out[:, 0], d[:, 0], rtol=0, atol=2e-6 * np.abs(d[:, 0]).max())


@pytest.mark.parametrize("site", ["01", "02"])
def test_redtrid_matches_fixture_eigenvalues(site):
    """redtrid (OCEAN_haydock.f90:2273-2316) against the eigenvalue block of abslanc."""
    n, _, a, b, eig = load_lanc(os.path.join(GOLD, f"abslanc_Fe.0001_1s_{site}"))
    ours = orc.redtrid_eigs(a, b, n)
    assert eig.shape == ours.shape == (n + 1,)
    assert np.abs(np.sort(eig) - ours).max() < 1e-8


def test_continued_fraction_one_level_is_lorentzian():
    """One Lanczos level with b(1) = 0 reduces to kpref * G / ((e - a0)^2 + G^2)."""
    a = np.array([0.3, 0.0])
    b = np.array([0.0, 0.0])
    out = orc.write_core(a, b, 1, 2.5, 0.01, -0.2, 0.8, 500, literal_cmplx=False)
    e = out[:, 0] / orc.HARTREE2EV
    ref = 2.5 * 0.01 / ((e - 0.3) ** 2 + 0.01 ** 2)
    assert relerr(out[:, 2], ref) < 1e-12


def test_continued_fraction_constant_chain_is_semicircle():
    """Constant a, b: the terminator makes the fraction exact, G = (z-a-sqrt((z-a)^2-4b^2))/(2b^2)."""
    n, a0, b0, gam, kp = 37, 0.25, 0.1, 0.004, 1.0
    a = np.full(n + 1, a0)
    b = np.full(n + 1, b0)
    b[0] = 0
    out = orc.write_core(a, b, n, kp, gam, -0.1, 0.6, 700, literal_cmplx=False)
    e = out[:, 0] / orc.HARTREE2EV
    z = e + 1j * gam
    s = np.sqrt((z - a0) ** 2 - 4 * b0 ** 2)
    s = np.where(s.imag > 0, s, -s)
    G = (z - a0 - s) / (2 * b0 ** 2)  # retarded branch
    delta = 1.0 / G
    ref = kp * np.abs(delta.imag) / np.abs(delta) ** 2
    assert relerr(out[:, 2], ref) < 1e-10


def test_literal_cmplx_rounding_is_single_precision_sized():
    """The kind-less cmplx() sites change the spectrum at the 1e-7 level, not more."""
    rng = np.random.default_rng(0)
    n = 100
    a = rng.uniform(0.5, 1.2, n + 1)
    b = rng.uniform(0.2, 0.6, n + 1)
    s1 = orc.write_core(a, b, n, 1.0, 0.02, 0.0, 2.0, 800, literal_cmplx=True)
    s2 = orc.write_core(a, b, n, 1.0, 0.02, 0.0, 2.0, 800, literal_cmplx=False)
    d = np.abs(s1[:, 2] - s2[:, 2]).max() / np.abs(s2[:, 2]).max()
    assert 1e-10 < d < 1e-5


# ---------------------------------------------------------------- FFT / lr
def test_fft_convention_matches_dense_dft():
    """FFTW_BACKWARD over (k1,k2,k3), k3 fastest (Modules/FFT_wrapper.f90:67-72)."""
    km = (3, 4, 2)
    F = orc.dft_matrix_backward(km)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(24) + 1j * rng.standard_normal(24)
    assert relerr(orc.fft_backward(v, km), F @ v) < 1e-13
    assert relerr(orc.fft_forward(v, km), np.conj(F) @ v) < 1e-13


def _tiny(**kw):
    base = dict(nb=4, kmesh=(2, 2, 2), xmesh=(2, 2, 2), lc=0, seed=1)
    base.update(kw)
    return synth.make_inputs(**base)


def test_lr_act_equals_dense_operator():
    """H_lr = -sum_x U_x^H (1/N) F_- diag(W_x) F_+ U_x built from explicit matrices."""
    inp = _tiny(kmesh=(2, 3, 2))
    P = to_problem(inp)
    nb, nk, na, nx = inp.nb, inp.nk, inp.nalpha, inp.nx
    F = orc.dft_matrix_backward(inp.kmesh)
    Hk = np.zeros((nb * nk, nb * nk), complex)
    for x in range(nx):
        U = np.zeros((nk, nb * nk), complex)
        for k in range(nk):
            U[k, k * nb:(k + 1) * nb] = inp.u[:, k, x, 0]
        C = np.conj(F) @ np.diag(inp.W[:, x]) @ F / nk
        Hk -= np.conj(U).T @ C @ U
    rng = np.random.default_rng(2)
    psi = np.asfortranarray(rng.standard_normal((nb, nk, na)) + 1j * rng.standard_normal((nb, nk, na)))
    got = orc.lr_act(P, psi)
    for a in range(na):
        ref = (Hk @ psi[:, :, a].ravel(order="F")).reshape((nb, nk), order="F")
        assert relerr(got[:, :, a], ref) < 1e-13
    assert relerr(orc.lr_kernel_loops(P, psi), got) < 1e-13


def _obf_case(no=5, **kw):
    inp = _tiny(**kw)
    ob = synth.make_obf(inp, no, seed=7)
    tau0 = inp.tau + inp.xshift / np.array(inp.xmesh, float)
    phs, xoff, tau_c, xshift = orc.obf_phase(inp.xmesh, inp.kmesh, tau0)
    P = to_problem(inp)
    P.W = ob.W
    # the Bloch functions the OBFs stand for (the sum OCEAN_obf_lrLOAD forms)
    u = np.einsum("nok,ox->nkx", ob.obf2u, ob.obf) * phs[None, :, :]
    P2 = to_problem(inp)
    P2.W = ob.W
    P2.u = np.asfortranarray(u[:, :, :, None])
    return inp, ob, phs, xoff, P, P2


def test_lr_act_obf2_is_scaled_traditional_action():
    """lr_act_obf2 = -N_k lr_act with u = sum_o obf2u obf phs: the OBF path's
    unnormalised transform pair and its + sign against the pinned Bloch path."""
    inp, ob, phs, _, P, P2 = _obf_case(kmesh=(2, 3, 2), xmesh=(3, 2, 2), nb=3)
    rng = np.random.default_rng(5)
    shp = (inp.nb, inp.nk, inp.nalpha)
    psi = np.asfortranarray(rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    got = orc.lr_act_obf2(P, psi, ob.obf, ob.obf2u, phs)
    assert relerr(got, -inp.nk * orc.lr_act(P2, psi)) < 1e-12


def test_obf_phase_closed_forms():
    """Even meshes: bloch_phase's rotation; odd kmesh(1): xshift(1) from xmesh(2)."""
    xm, km, tau = (3, 4, 2), (2, 4, 2), np.array([0.9, 0.27, 0.41])
    phs, xoff, tau_c, xshift = orc.obf_phase(xm, km, tau)
    nk, nx = int(np.prod(km)), int(np.prod(xm))
    rot, tau_b, xs_b = orc.bloch_phase(np.ones((1, nk, nx, 1), complex), xm, km, tau)
    assert np.abs(phs - rot[0, :, :, 0]).max() < 1e-14
    assert np.array_equal(xshift, xs_b) and np.allclose(tau_c, tau_b)
    km = (3, 4, 2)
    xm = (3, 5, 2)
    phs, xoff, tau_c, xshift = orc.obf_phase(xm, km, tau)
    assert xshift[0] == math.floor(5 * (0.9 - 0.5)) == 2   # xmesh(2), not xmesh(1) (-> 1)
    assert abs(tau_c[0] - (0.9 - 2 / 3)) < 1e-15
    # the product's offsets agree with the restatement
    xo, tc, xs = sh.obf_x_offsets(xm, km, tau)
    assert np.array_equal(xo, xoff) and np.array_equal(xs, xshift) and np.allclose(tc, tau_c)
    # a plane wave: phs(q, x) at q = (1, 0, 0) is exp(2 pi i xoff_1 / (X_1 K_1))
    q = 1 * km[1] * km[2]
    assert np.abs(phs[q] - np.exp(2j * np.pi * xoff[:, 0] / (xm[0] * km[0]))).max() < 1e-13


def test_xact_obf_drops_multiplet_and_is_hermitian():
    """xact_obf = xact with the multiplet off, SO on and W -> -N_k W on the OBF
    Bloch functions; the operator is Hermitian."""
    inp, ob, phs, _, P, P2 = _obf_case(lc=1, xi_eV=6.0, nb=3, kmesh=(2, 2, 2), xmesh=(2, 2, 3))
    P2.W = -inp.nk * ob.W
    rng = np.random.default_rng(6)
    shp = (inp.nb, inp.nk, inp.nalpha)
    psi = np.asfortranarray(rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    for back in (False, True):
        got = orc.xact_obf(P, psi, ob.obf, ob.obf2u, phs, backwards=back)
        ref = orc.xact(P2, psi, backwards=back, hflag=(1, 0, 0, 0, 0, 1))
        assert relerr(got, ref) < 1e-12
    # without the long-range term the multiplet stays (nothing overwrites it)
    hf = (0, 0, 0, 0, 1, 1)
    assert relerr(orc.xact_obf(P, psi, ob.obf, ob.obf2u, phs, hflag=hf), orc.xact(P, psi, hflag=hf)) < 1e-14
    N = int(np.prod(shp))
    H = np.zeros((N, N), complex)
    for j in range(N):
        e = np.zeros(N, complex)
        e[j] = 1.0
        H[:, j] = orc.xact_obf(P, e.reshape(shp, order="F"), ob.obf, ob.obf2u, phs).ravel(order="F")
    assert np.abs(H - np.conj(H.T)).max() < 1e-12 * np.abs(H).max()


@pytest.mark.parametrize("lc", [0, 1])
def test_soprep_lifetimes_closed_form(lc):
    """someli = -<L.S> l_alpha - l_beta: the j = lc + 1/2 states (2lc + 2 per valence
    spin) get -life_time(1), the j = lc - 1/2 states (2lc) get -life_time(2)."""
    lt = (0.3, 0.55)
    _, si, have = orc.soprep(lc, 0.0, lt)
    assert have
    ev = np.sort(np.linalg.eigvalsh(si)) / orc.EV2HARTREE
    want = np.sort([-lt[1]] * (2 * (2 * lc)) + [-lt[0]] * (2 * (2 * lc + 2)))
    assert np.abs(ev - want).max() < 1e-13


def test_fg_combo_equals_dense_operator():
    """H_fg = conj(C) (mul K) C^T with K = -mpm (direct) + X^T block (exchange)."""
    inp = _tiny(lc=1, seed=4)
    P = to_problem(inp)
    m = inp.mult
    nb, nk, na = inp.nb, inp.nk, inp.nalpha
    mul = inp.inter_scale / (nk * inp.celvol)
    plist = [(l, mm, 0) for l in range(m.lmin, m.lmax + 1) for mm in range(-l, l + 1)]
    Pn = len(plist)
    N = nb * nk * na
    Cm = np.zeros((N, Pn * na), complex)
    for a in range(na):
        for j, (l, mm, nu) in enumerate(plist):
            col = np.zeros((nb, nk, na), complex)
            col[:, :, a] = m.c[:, :, nu, mm + m.lmax, l - m.lmin, 0]
            Cm[:, a * Pn + j] = col.ravel(order="F")
    K = np.zeros((Pn * na, Pn * na), complex)
    for a in range(na):
        for j, (l, mm, nu) in enumerate(plist):
            K[a * Pn + j, a * Pn + j] -= m.mpm[0, 0, l - m.lmin]
    for l in range(m.lmin, m.lmax + 1):
        X = m.X[l - m.lvl]
        idx = [(a, (l, mm, 0)) for a in range(na) for mm in range(-l, l + 1)]
        for ii, (a, p) in enumerate(idx):
            for jj, (b, q) in enumerate(idx):
                K[a * Pn + plist.index(p), b * Pn + plist.index(q)] += X[jj, ii]
    H = np.conj(Cm) @ (mul * K) @ Cm.T
    rng = np.random.default_rng(3)
    psi = np.asfortranarray(rng.standard_normal((nb, nk, na)) + 1j * rng.standard_normal((nb, nk, na)))
    got = orc.fg_combo(P, psi, inp.inter_scale).ravel(order="F")
    assert relerr(got, H @ psi.ravel(order="F")) < 1e-12
    assert np.abs(H - np.conj(H.T)).max() < 1e-14 * np.abs(H).max()


@pytest.mark.parametrize("lc,nspn,xi", [(0, 1, 0.0), (1, 1, 6.0), (1, 2, 3.0)])
def test_full_action_is_hermitian(lc, nspn, xi):
    """H = E + A(H_lr + H_fg)A + S is Hermitian for real W, Hermitian K and S."""
    inp = _tiny(lc=lc, nspn=nspn, xi_eV=xi, kmesh=(2, 2, 2), xmesh=(2, 2, 3), nb=3)
    H = orc.dense_H(to_problem(inp))
    assert np.abs(H - np.conj(H.T)).max() < 1e-13 * np.abs(H).max()


def test_backwards_flag_conjugates_e_and_s():
    inp = _tiny(lc=1, xi_eV=6.0, nb=3)
    inp.energy = inp.energy + 1j * 0.01  # a lifetime term makes the conjugation visible
    P = to_problem(inp)
    Hf = orc.dense_H(P)
    Hb = orc.dense_H(P, backwards=True)
    Pn = to_problem(inp, e0=True)
    Pn.mult.someli = Pn.mult.someli  # zero here; E's imaginary part carries the check
    d = Hf - Hb
    diag = np.diag(d)
    assert np.allclose(diag.imag, 2 * 0.01, atol=1e-14)


# ---------------------------------------------------------------- Haydock / GMRES
def test_haydock_matches_dense_lanczos():
    """a/b from the recursion equal Lanczos on the dense matrix (first 12 steps)."""
    inp = _tiny(lc=1, xi_eV=4.0, nb=4, kmesh=(2, 2, 2), xmesh=(2, 2, 2))
    P = to_problem(inp)
    H = orc.dense_H(P)
    v = inp.psi0.ravel(order="F").copy()
    n = 12
    ra, rb, _, ierr = orc.haydock(P, inp.psi0, n)
    assert ierr == 0
    # textbook Lanczos with full reorthogonalisation on the dense matrix
    Q = [v / np.linalg.norm(v)]
    aa, bb = [], [0.0]
    for j in range(n):
        w = H @ Q[-1]
        aa.append(np.vdot(Q[-1], w).real)
        w = w - aa[-1] * Q[-1] - (bb[-1] * Q[-2] if j > 0 else 0)
        for q in Q:
            w = w - np.vdot(q, w) * q
        bb.append(np.linalg.norm(w))
        Q.append(w / bb[-1])
    assert np.allclose(ra[:n], aa, rtol=1e-11, atol=1e-13)
    assert np.allclose(rb[1:n + 1], bb[1:], rtol=1e-10)


def test_haydock_spectrum_equals_eigen_sum():
    """Full-length recursion: the continued fraction equals sum_l |<v_l|psi0>|^2 Lorentzians."""
    inp = _tiny(nb=2, kmesh=(2, 1, 1), xmesh=(2, 1, 1))  # N = 2*2*4 = 16
    P = to_problem(inp)
    H = orc.dense_H(P)
    N = H.shape[0]
    ra, rb, _, _ = orc.haydock(P, inp.psi0, N)
    lam, V = np.linalg.eigh(H)
    w = np.abs(V.conj().T @ inp.psi0.ravel(order="F")) ** 2
    gam = 0.01
    out = orc.write_core(ra, rb, N, 1.0, gam, lam.min() - 0.1, lam.max() + 0.1, 400, literal_cmplx=False)
    e = out[:, 0] / orc.HARTREE2EV
    ref = np.sum(w[None, :] * gam / ((e[:, None] - lam[None, :]) ** 2 + gam ** 2), axis=1)
    assert relerr(out[:, 2], ref) < 1e-6


def test_gmres_matches_dense_solve():
    """-Im<b|x> from GMRES at each energy equals the direct solve of (E+iG-H)x=b."""
    inp = _tiny(nb=3, seed=9, xmesh=(3, 1, 1))
    P = to_problem(inp)
    E = np.array([-1.0, 5.0, 12.0, 12.5, 25.0])
    rows = orc.gmres_do_recycle(P, inp.psi0, E, 0.5, 13.605, 1e-14, 12, inp.kpref)
    H = orc.dense_H(P)
    b = (P.allow * inp.psi0).ravel(order="F")
    for r in rows:
        z = r[0] * orc.EV2HARTREE + 1j * 0.5 * orc.EV2HARTREE
        x = np.linalg.solve(z * np.eye(H.shape[0]) - H, b)
        d = np.vdot(b, x)
        assert abs(-d.imag * inp.kpref - r[2]) < 1e-6 * abs(r[2])
        assert abs((1 - d.real * inp.kpref) - r[1]) < 1e-6


# ---------------------------------------------------------------- setup
def test_gaunt_quadrature_equals_3j_closed_form():
    """threey (newthreey.f90) by Gauss-Legendre quadrature vs Wigner-3j closed forms."""
    x, w = orc.gauss_legendre(16)
    prefs = orc.getprefs(5)
    worst = 0.0
    for l1 in range(0, 4):
        for l3 in range(0, 4):
            for k in range(abs(l1 - l3), min(l1 + l3, 5) + 1):
                for m1 in range(-l1, l1 + 1):
                    for m3 in range(-l3, l3 + 1):
                        for cj in (False, True):
                            q = (m3 - m1) if cj else (m1 - m3)
                            if abs(q) > k:
                                continue
                            a = orc.threey(l1, m1, k, q, l3, m3, cj, x, w, prefs)
                            b1 = orc.gaunt_closed(l1, m1, k, q, l3, m3, cj)
                            b2 = sh.gaunt(l1, m1, k, q, l3, m3, cj)
                            worst = max(worst, abs(a - b1), abs(a - b2))
    assert worst < 1e-13


@pytest.mark.parametrize("lc,lv", [(0, 0), (0, 1), (0, 3), (1, 0), (1, 1), (1, 2), (1, 3)])
def test_exchange_matrix_quadrature_vs_closed_form(lc, lv):
    """nbsemhsetup2 (OCEAN_multiplet.f90:2007-2181) restated with quadrature vs setup_host."""
    rng = np.random.default_rng(lc * 10 + lv)
    fk = {k: np.array([[rng.uniform(0.5, 2)]]) for k in range(2, min(2 * lc, 2 * lv) + 1, 2)}
    gk = {k: np.array([[rng.uniform(0.5, 2)]]) for k in range(abs(lc - lv), lc + lv + 1, 2)}
    ones = {k: 1.0 for k in range(0, 8)}
    a = orc.nbsemhsetup2(lc, lv, 1, fk, ones, gk, ones)
    b = sh.exchange_matrix(lc, lv, 1, fk, gk)
    assert np.abs(a - b).max() < 1e-12
    assert np.abs(a - np.conj(a.T)).max() < 1e-12


def test_spin_orbit_closed_form_splitting():
    """-xi <L.S>: j = 3/2 at -xi/2, j = 1/2 at +xi; splitting 1.5 xi (OCEAN_soprep :572-690)."""
    xi = 6.0
    sr, si, have = orc.soprep(1, xi)
    sr2, si2, have2 = sh.spin_orbit(1, xi)
    assert have and have2
    assert np.abs(sr - sr2).max() < 1e-15 and np.abs(si).max() == 0
    # restrict to one valence spin: the 6 core states (spin x m)
    idx = [a for a, lab in enumerate(orc.alpha_labels(1)) if lab[2] == -0.5]
    ev = np.sort(np.linalg.eigvalsh(sr[np.ix_(idx, idx)])) / orc.EV2HARTREE
    assert np.allclose(ev, [-3.0] * 4 + [6.0] * 2, atol=1e-12)
    assert np.abs(orc.soprep(0, xi)[0]).max() == 0.0  # K edge: no orbital moment


def test_jimel_matches_ladder_operators():
    for j in (0.5, 1.0, 2.0):
        ms = np.arange(-j, j + 1)
        for m in ms:
            for mp in ms:
                for i, M in enumerate(sh._lm_matrix(int(j)) if j >= 1 else sh._lm_matrix_half()):
                    ref = M[int(round(m + j)), int(round(mp + j))]
                    assert abs(orc.jimel(j, m, mp, i + 1) - ref) < 1e-14


def test_intval_matches_interp_and_caps():
    xt = np.array([0.5, 1.0, 2.0, 4.0])
    yt = np.array([3.0, 2.0, 1.5, 0.5])
    x = np.linspace(0.0, 5.0, 51)
    y = orc.intval(xt, yt, x)
    assert np.allclose(y, np.interp(x, xt, yt), atol=1e-15)
    ext = orc.intval(xt, yt, np.array([0.0, 5.0]), "ext", "ext")
    assert np.allclose(ext, [3.0 + (0.0 - 0.5) / 0.5 * (2.0 - 3.0), 1.5 + (5.0 - 2.0) / 2.0 * (0.5 - 1.5)])


def test_populate_w_closed_forms_and_host_agreement():
    """lr_populate_W2: r = 0 gives ptab(1); beyond the table 1/(eps r); host builder agrees."""
    rtab, ptab = synth.w_table()
    W = orc.populate_W(np.eye(3) * 6.0, (1, 1, 1), (1, 1, 1), (0, 0, 0), (0, 0, 0), rtab, ptab, 4.0)
    assert W.shape == (1, 1) and W[0, 0] == ptab[0]
    W = orc.populate_W(np.eye(3) * 40.0, (2, 1, 1), (1, 1, 1), (0, 0, 0), (0, 0, 0), rtab, ptab, 4.0)
    assert abs(W[1, 0] - 1.0 / (4.0 * 40.0)) < 1e-15  # R = -1 cell, r = 40 bohr
    tau = np.array([0.13, 0.27, 0.41])
    u = np.zeros((1, 27, 64, 1), complex)
    _, tau_c, xs = orc.bloch_phase(u, (4, 4, 4), (3, 3, 3), tau)
    a = orc.populate_W(np.eye(3) * 6.0, (3, 3, 3), (4, 4, 4), tau_c, xs, rtab, ptab, 4.0)
    b = sh.build_w(np.eye(3) * 6.0, (3, 3, 3), (4, 4, 4), tau_c, xs, rtab, ptab, 4.0)
    assert np.abs(a - b).max() < 1e-15


def test_w_hat_weights_reconstruct_w():
    """dW/dptab weights: W = sum of the two hat weights times ptab inside the table."""
    rtab, ptab = synth.w_table()
    tau = np.array([0.1, 0.2, 0.3])
    u = np.zeros((1, 8, 27, 1), complex)
    _, tau_c, xs = orc.bloch_phase(u, (3, 3, 3), (2, 2, 2), tau)
    W = orc.populate_W(np.eye(3) * 3.0, (2, 2, 2), (3, 3, 3), tau_c, xs, rtab, ptab, 1e30)
    i0, w0, w1 = orc.populate_W_weights(np.eye(3) * 3.0, (2, 2, 2), (3, 3, 3), tau_c, xs, rtab)
    assert np.abs(w0 * ptab[i0] + w1 * ptab[i0 + 1] - W).max() < 1e-14


def test_bloch_phase_closed_form_and_host_agreement():
    """lrLOAD phase: with no shift, u(x,k) -> exp(2 pi i sum_d x_d k_d / (X_d K_d)) u."""
    xm, km = (2, 3, 2), (2, 2, 3)
    nk, nx = 12, 12
    u = np.ones((1, nk, nx, 1), complex)
    up, tau_c, xs = orc.bloch_phase(u, xm, km, (0.0, 0.0, 0.5))  # odd K3 centres at tau - 1/2
    assert np.all(xs == 0)
    kidx = [(a, b, c) for a in range(2) for b in range(2) for c in range(3)]
    xidx = [(i, j, l) for l in range(2) for j in range(3) for i in range(2)]
    for k, kv in enumerate(kidx):
        for x, xv in enumerate(xidx):
            ph = 2 * np.pi * sum(xv[d] * kv[d] / (xm[d] * km[d]) for d in range(3))
            assert abs(up[0, k, x, 0] - np.exp(1j * ph)) < 1e-14
    rng = np.random.default_rng(0)
    ur = rng.standard_normal((3, nk, nx, 1)) + 1j * rng.standard_normal((3, nk, nx, 1))
    a = orc.bloch_phase(ur, xm, km, (0.6, 0.1, 0.8))
    b = sh.bloch_phase(ur, xm, km, (0.6, 0.1, 0.8))
    assert np.abs(a[0] - b[0]).max() < 1e-14 and np.allclose(a[1], b[1]) and np.all(a[2] == b[2])


def test_pnorm_kpref():
    """core_pnorm (OCEAN_psi.f90:5589-5638): kpref = 4 pi |psi|^2 / (N_k Omega^2)."""
    rng = np.random.default_rng(0)
    p = rng.standard_normal((3, 4, 4)) + 1j * rng.standard_normal((3, 4, 4))
    q, kp = orc.core_pnorm(p, 4, 216.0)
    assert abs(np.linalg.norm(q) - 1) < 1e-14
    ref = 4 * np.pi * np.linalg.norm(p) ** 2 / (4 * 216.0 ** 2)
    assert abs(kp - ref) < 1e-13 * ref
    q2, kp2 = sh.normalise_psi0(p, 4, 216.0)
    assert abs(kp - kp2) < 1e-13 * ref and np.abs(q - q2).max() < 1e-15


def test_psi_dotter_selects_matching_spins():
    """psi_dotter (OCEAN_psi.f90:5887-6031): only core spin == valence spin channels are set."""
    rng = np.random.default_rng(1)
    nb, nk, lc, nptot = 3, 2, 1, 4
    mels = rng.standard_normal((nptot, 2 * lc + 1)) + 1j * rng.standard_normal((nptot, 2 * lc + 1))
    cks = rng.standard_normal((nptot, nb * nk, 1)) + 1j * rng.standard_normal((nptot, nb * nk, 1))
    psi = orc.psi_dotter(mels, cks, lc, nb, nk, 1)
    for a, (cs, cm, vs) in enumerate(orc.alpha_labels(lc)):
        if cs != vs:
            assert np.all(psi[:, :, a] == 0)
        else:
            k, n = 1, 2
            ref = np.conj(np.sum(mels[:, int(cm) + lc] * cks[:, k * nb + n, 0]))
            assert abs(psi[n, k, a] - ref) < 1e-14


def test_nbsemkcmel_bode_rule_converges():
    """nbsemkcmel (OCEAN2/nbsemkcmel.f90:8-209): trim at rmax, Bode-rule cmel on the log grid."""
    rmax = 1.5
    nr = 1201
    r0 = 1e-3
    dl = math.log(rmax / r0) / 800.0
    rad = r0 * np.exp(dl * np.arange(nr))
    rv = np.linspace(1e-4, 10.0, 4000)
    v = 1.0 / (1.0 + rv)
    phi = np.exp(-rad)[:, None]
    vtrim, cmel, nmel = orc.nbsemkcmel(v, rv, rmax, rad, phi)
    vrun = v[np.searchsorted(rv, rmax)]
    mask = rad <= rmax
    f = np.exp(-2 * rad) * (np.interp(rad, rv, v - np.where(rv < rmax, vrun, v)) ) * rad ** 2
    ref = np.trapezoid(np.where(mask, f, 0.0), rad) * orc.HARTREE2EV
    assert abs(cmel[0, 0] - ref) < 2e-3 * abs(ref)
    assert np.all(vtrim[rv < rmax] == vrun) and np.all(vtrim[rv >= rmax] == v[rv >= rmax])


# ---------------------------------------------------------------- the C restatement
@pytest.fixture(scope="module")
def c_oracle_mod():
    import c_oracle
    c_oracle.build()
    return c_oracle


@pytest.mark.parametrize("kw", [
    dict(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=3),
    dict(nb=6, kmesh=(3, 4, 2), xmesh=(4, 3, 2), lc=0, seed=4),
    dict(nb=6, kmesh=(2, 2, 2), xmesh=(2, 2, 2), lc=1, seed=5, xi_eV=6.0),
    dict(nb=4, kmesh=(2, 2, 4), xmesh=(3, 1, 3), lc=1, seed=7, nspn=2, xi_eV=3.0),
    dict(nb=4, kmesh=(5, 3, 6), xmesh=(2, 1, 2), lc=0, seed=8),
])
def test_c_restatement_equals_numpy_oracle(c_oracle_mod, kw):
    inp = synth.make_inputs(**kw)
    P = to_problem(inp)
    cs = c_oracle_mod.CSystem(inp)
    rng = np.random.default_rng(1)
    psi = np.asfortranarray(rng.standard_normal(cs.shape) + 1j * rng.standard_normal(cs.shape))
    for back in (False, True):
        assert relerr(cs.xact(psi, backwards=back, nthreads=2), orc.xact(P, psi, backwards=back)) < 1e-13


def test_c_restatement_haydock(c_oracle_mod):
    inp = synth.make_inputs(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=1, seed=2, xi_eV=5.0)
    P = to_problem(inp)
    ra, rb, _, rc = c_oracle_mod.CSystem(inp).haydock(inp.psi0, 30, nthreads=2)
    a, b, _, _ = orc.haydock(P, inp.psi0, 30)
    assert rc == 0
    assert np.abs(ra[:30] - a[:30]).max() < 1e-12 * np.abs(a).max()
    assert np.abs(rb[1:31] - b[1:31]).max() < 1e-12 * np.abs(b).max()


# ---------------------------------------------------------------------------
# differentiation form: the oracle's adjoint gradient against finite differences
# ---------------------------------------------------------------------------
def test_spectral_grad_matches_central_differences():
    from grad_common import oracle_problem, target_params
    from ocean_amd import synth

    inp = synth.make_inputs(nb=4, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=31, inter_scale=0.9)
    E = np.linspace(-2.0, 32.0, 6)
    kp = inp.kpref
    tp, ts, tg, td = target_params(inp)
    target = orc.spectral_sigma(oracle_problem(inp, tp, ts, td), inp.psi0, E, tg, kp)
    p0, s0, g0, d0 = inp.ptab.copy(), inp.inter_scale, 0.5, 0.0
    P0 = oracle_problem(inp, p0, s0, d0)
    ww = orc.populate_W_weights(inp.avec, inp.kmesh, inp.xmesh, inp.tau, inp.xshift, inp.rtab, s0)
    G = orc.spectral_grad(P0, inp.psi0, E, g0, kp, target, wweights=ww, nptab=len(p0))

    def loss(p=p0, s=s0, g=g0, d=d0):
        sig = orc.spectral_sigma(oracle_problem(inp, p, s, d), inp.psi0, E, g, kp)
        return float(np.sum((sig - target) ** 2))

    h = 1e-6
    assert abs(G["loss"] - loss()) <= 1e-12 * abs(G["loss"])
    fd = {
        "g_gamma_eV": (loss(g=g0 + h) - loss(g=g0 - h)) / (2 * h),
        "g_scissor_eV": (loss(d=d0 + h) - loss(d=d0 - h)) / (2 * h),
        "g_scale": (loss(s=s0 + h) - loss(s=s0 - h)) / (2 * h),
    }
    for k, v in fd.items():
        assert abs(G[k] - v) <= 2e-5 * abs(v) + 1e-12, (k, G[k], v)
    # the knots that carry most of the W gradient
    for i in np.argsort(-np.abs(G["g_ptab"]))[:4]:
        dp = np.zeros_like(p0)
        dp[i] = h
        v = (loss(p=p0 + dp) - loss(p=p0 - dp)) / (2 * h)
        assert abs(G["g_ptab"][i] - v) <= 2e-5 * abs(v) + 1e-12, (i, G["g_ptab"][i], v)
    # knots beyond every distance carry no gradient
    far = inp.rtab > orc._w_geometry(inp.avec, inp.kmesh, inp.xmesh, inp.tau, inp.xshift)[0].max() + 0.1
    assert np.all(G["g_ptab"][far] == 0.0)


# ---------------------------------------------------------------------------
# early exit of the Haydock loop
# ---------------------------------------------------------------------------
def test_calc_spect_core_is_write_core_at_midpoints():
    """calc_spect_core (:1991-2027) is write_core's fraction (pinned by the Fe fixture)
    on the midpoint grid e = el + (eh - el)(2 ie - 1)/(2 ne)."""
    n, kpref, a, b, _ = load_lanc(os.path.join(GOLD, "abslanc_Fe.0001_1s_01"))
    el, eh, ne, gam = -1.5, 1.7, 400, 0.018
    cs = orc.calc_spect_core(a, b, n, kpref, gam, el, eh, ne)
    h = (eh - el) / (2 * ne)
    wc = orc.write_core(a, b, n, kpref, gam, el + h, eh - h, ne - 1)
    assert np.allclose(cs[0], wc[:, 0], rtol=1e-13, atol=0)
    assert np.abs(cs[2] - wc[:, 2]).max() <= 1e-9 * np.abs(wc[:, 2]).max()
    assert np.abs(cs[1] - wc[:, 4]).max() <= 1e-9 * np.abs(wc[:, 4]).max()


def test_early_exit_on_a_constant_chain_stops_at_the_first_pair():
    """With a_i = a, b_i = b the square-root terminator makes every truncation the same
    fraction, so relArea vanishes (up to the single-precision rounding of the kind-less
    cmplx sites, ~1e-8) from the first check and the loop stops at spacing + 2."""
    n = 60
    a = np.full(n + 1, 0.4)
    b = np.full(n + 1, 0.25)
    b[0] = 0.0
    for spacing in (3, 10):
        assert orc.haydock_early_exit(a, b, n, spacing, 1e-6, 1.0, 0.02, -0.5, 1.5, 300) == spacing + 2
    # and a threshold of zero never stops it
    assert orc.haydock_early_exit(a, b, n, 5, 0.0, 1.0, 0.02, -0.5, 1.5, 300) == n


# ------------------------------------------------------------ write_core (nine columns)
def _fe_fixture(site):
    n, kpref, a, b, _ = load_lanc(os.path.join(GOLD, f"abslanc_Fe.0001_1s_{site}"))
    gam, d = load_spct(os.path.join(GOLD, f"absspct_Fe.0001_1s_{site}"))
    h2e = orc.HARTREE2EV_OLD
    return n, kpref, a, b, gam, d, h2e


FE_NVAL = 16.0 / 5.423513587 ** 3  # TEST/Fe/Fe.in: zion 16 (fe.fhi), one atom, acell^3
FE_EPS = 1.0e5                     # TEST/Fe/Fe.in: diemac


@pytest.mark.parametrize("site", ["01", "02"])
def test_write_core_gamfcn_columns_reproduce_fe_fixture(site):
    """write_core with oldXASbroaden (OCEAN_haydock.f90:1801-1894): Spect and SPKK use
    gam0 + gamfcn(E) (CNBSE/zcmsnmult/zf90/gamfcn.f90) at the Fe run's valence density
    and dielectric constant; Spect(0)/SPKK(0) keep gam0.  The fixture's first five
    columns are reproduced to the printed precision."""
    n, kpref, a, b, gam, d, h2e = _fe_fixture(site)
    ne = d.shape[0] - 1
    el, eh = d[0, 0] / h2e, d[-1, 0] / h2e
    out = orc.write_core_full(a, b, n, kpref, gam, el, eh, ne, old_broaden=True, nval=FE_NVAL, eps=FE_EPS,
                              hartree2ev=h2e)
    assert out.shape == (ne + 1, 9)
    assert np.abs(out[:, 0] - d[:, 0]).max() < 1e-6
    for c in (1, 2):
        assert np.abs(out[:, c] - d[:, c]).max() < 1e-6 * np.abs(d[:, c]).max(), c
    for c in (3, 4):
        assert np.abs(out[:, c] - d[:, c]).max() < 1e-6 * np.abs(d[:, c]).max(), c
    # the broadened columns differ from the unbroadened ones only above E_F (E > 0)
    below = out[:, 0] / h2e < -1e-3
    assert np.array_equal(out[below, 1], out[below, 2])
    assert np.abs(out[~below, 1] - out[~below, 2]).max() > 1e-3 * np.abs(out[:, 2]).max()


def test_gamfcn_vanishes_below_fermi_level_and_grows_above():
    """gamfcn.f90:13-15: no decay channel below E_F.  Plasmon emission needs
    k^2 - kf^2 > 2 w_q, with w_q >= wp = sqrt(4 pi n) (the q -> 0 pole of a metal), so the
    rate is exactly zero up to E = wp, then positive and rising just above it."""
    wp = math.sqrt(4 * math.pi * FE_NVAL)
    assert orc.gamfcn(-0.01, FE_NVAL, FE_EPS) == 0.0
    assert orc.gamfcn(0.98 * wp, FE_NVAL, FE_EPS) == 0.0
    g = [orc.gamfcn(e * wp, FE_NVAL, FE_EPS) for e in (1.3, 1.8, 2.5)]
    assert 0.0 < g[0] < g[1] < g[2]


def test_write_core_cross_section_columns():
    """Columns 6-9 of write_core (OCEAN_haydock.f90:1855-1880): sigma = E alpha V Spect(0)
    with E in Hartree (a.u.^2), times bohr^2 in cm^2; the two-photon columns
    2 pi alpha^2 V Spect(0) (a.u.^5) and that times bohr^4 cm^4 and the atomic time."""
    n, kpref, a, b, gam, d, h2e = _fe_fixture("01")
    V = 5.423513587 ** 3
    out = orc.write_core_full(a, b, n, kpref, gam, -0.5, 1.0, 40, ebase_eV=3.0, ucvol=V)
    E_h = out[:, 0] * orc.EV2HARTREE
    assert np.allclose(out[:, 5], E_h * orc.ALPHA * V * out[:, 2], rtol=1e-14, atol=0)
    assert np.allclose(out[:, 6], out[:, 5] * (0.529177210544e-8) ** 2, rtol=1e-12, atol=0)
    assert np.allclose(out[:, 7], 2 * np.pi * orc.ALPHA ** 2 * V * out[:, 2], rtol=1e-14, atol=0)
    assert np.allclose(out[:, 8], out[:, 7] * (0.529177210544e-8) ** 4 * 2.4188843265864e-17, rtol=1e-12,
                       atol=0)
    # printed energy: ebase + Hartree2eV * (el + (eh - el) ie / ne)
    assert np.allclose(out[:, 0], 3.0 + orc.HARTREE2EV * np.linspace(-0.5, 1.0, 41), rtol=0, atol=1e-12)


def test_write_core_voigt_is_gaussian_convolved_lorentzian():
    """write_core_Voigt (OCEAN_haydock.f90:1896-1989) for a one-level fraction (b1 = 0:
    an exact Lorentzian of half width gam0) is pi kpref times scipy's Voigt profile,
    up to the +-3 sigma truncation and the Riemann sum on the fine grid; the printed
    energy is the last fine energy, ee + nstep estep."""
    from scipy.special import voigt_profile
    a0, gam0, sig, kpref = 0.2, 0.02, 0.03, 1.7
    a = np.array([a0, 0.0])
    b = np.array([0.0, 0.0])
    el, eh, ne = -0.1, 0.5, 60
    out = orc.write_core_full(a, b, 1, kpref, gam0, el, eh, ne, gam_gauss=sig, nval=FE_NVAL, eps=FE_EPS,
                              literal_cmplx=False)
    assert out.shape == (ne + 1, 5)
    ee = el + (eh - el) * np.arange(ne + 1) / ne
    ref = np.pi * kpref * voigt_profile(ee - a0, sig, gam0)
    peak = ref.max()
    assert np.abs(out[:, 2] - ref).max() < 6e-3 * peak
    # a half-step shift of the grid would already break the bound
    shifted = np.pi * kpref * voigt_profile(ee - a0 + 0.5 * (eh - el) / ne, sig, gam0)
    assert np.abs(out[:, 2] - shifted).max() > 2e-2 * peak
    de = (eh - el) / ne
    estep = min(gam0, sig) / 2
    nratio = int((de - estep) / estep) + 1
    estep = de / nratio
    nstep = int(2.99999 * sig / estep) + 1
    assert np.allclose(out[:, 0], orc.HARTREE2EV * (ee + nstep * estep), rtol=0, atol=1e-10)
    # below E_F the gamfcn columns equal the plain ones
    below = ee + nstep * estep < -1e-9
    assert np.array_equal(out[below, 1], out[below, 2])


def test_write_core_voigt_reuse_equals_direct_fine_grid():
    """The reference reuses the fine grid from one output energy to the next (shift by
    nratio).  Its sums equal a direct evaluation of every fine energy ee + ii estep
    when the fine energy is formed the same way (point q's ee plus its offset), which
    this rebuilds independently from the reuse bookkeeping."""
    n, kpref, a, b, gam, d, h2e = _fe_fixture("01")
    el, eh, ne, sig = -0.2, 0.3, 12, 0.01
    out = orc.write_core_full(a, b, n, kpref, gam, el, eh, ne, gam_gauss=sig, nval=FE_NVAL, eps=FE_EPS)
    de = (eh - el) / ne
    estep = min(gam, sig) / 2
    nratio = int((de - estep) / estep) + 1
    estep = de / nratio
    nstep = int(2.99999 * sig / estep) + 1
    pref = estep / (sig * 2.506628274631)
    def writer(p, ii):  # point p's own loop writes ii >= nstep - nratio; the rest came from p - 1
        if p == 0 or ii >= nstep - nratio:
            return p, ii
        return writer(p - 1, ii + nratio)

    for p in (0, 1, 5, ne):
        acc = np.zeros(2)
        for ii in range(-nstep, nstep + 1):
            q, o = writer(p, ii)
            e = el + (eh - el) * (2 * q) / (2 * ne) + o * estep
            dl = orc._cf_delta(a, b, np.zeros_like(a), n, np.array([e]), np.array([gam]), True)[0]
            acc[0] += math.exp(-ii * ii * estep * estep / (2 * sig * sig)) * pref * kpref * abs(dl.imag) / abs(dl) ** 2
        assert abs(acc[0] - out[p, 2]) <= 1e-12 * abs(out[p, 2]), p


# ------------------------------------------------------------ o_spect
def test_o_spect_core_reproduces_fe_fixture():
    """o_spect's core branch on the Fe abslanc (ebase 0: no anti-resonant term) gives the
    fixture's Spect, Spect(0) and, in both SPKK columns, its broadened SPKK."""
    n, kpref, a, b, gam, d, h2e = _fe_fixture("01")
    ne = d.shape[0] - 1
    el, eh = d[0, 0] / h2e, d[-1, 0] / h2e
    out = orc.o_spect_core(a, b, n, kpref, ne, el, eh, gam, 0.0, FE_NVAL, FE_EPS)
    assert out.shape == (ne + 1, 5)
    assert np.abs(out[:, 0] * h2e / orc.HARTREE2EV - d[:, 0]).max() < 1e-6
    for c, fc in ((1, 1), (2, 2), (3, 3), (4, 3)):
        assert np.abs(out[:, c] - d[:, fc]).max() < 1e-6 * np.abs(d[:, fc]).max(), c


def test_o_spect_core_anti_resonant_closed_form():
    """One level a0 (b1 = 0) with ebase > 0.01: eps = -kpref/(E - a0 + i g) + kpref/(E + a0 +
    2 ebase + i g), the resonant pole minus its mirror; below E_F gamfcn adds nothing."""
    a0, g, kp, eb = -0.3, 0.02, 0.7, 0.2
    a = np.array([a0, 0.0])
    b = np.zeros(2)
    out = orc.o_spect_core(a, b, 1, kp, 50, -0.6, -0.01, g, eb, FE_NVAL, FE_EPS)
    E = -0.6 + 0.59 * np.arange(51) / 50
    eps = -kp / (E - a0 + 1j * g) + kp / (E + a0 + 2 * eb + 1j * g)
    assert np.allclose(out[:, 0], (E + eb) * orc.HARTREE2EV, rtol=1e-15, atol=1e-12)
    for c in (1, 2):
        assert np.allclose(out[:, c], eps.imag, rtol=1e-12, atol=1e-14)
    for c in (3, 4):
        assert np.allclose(out[:, c], -eps.real, rtol=1e-12, atol=1e-14)
    # without the anti-resonant term (ebase below 0.01) only the resonant pole remains
    out0 = orc.o_spect_core(a, b, 1, kp, 50, -0.6, -0.01, g, 0.005, FE_NVAL, FE_EPS)
    assert np.allclose(out0[:, 2], (-kp / (E - a0 + 1j * g)).imag, rtol=1e-12, atol=1e-14)


def test_o_spect_val_lorentz_oscillator():
    """One level: eps(w) = 1 + 2 a0 f / (a0^2 - (w + i g)^2), f = 2 kpref ucvol; then
    (n + i kappa)^2 = eps, R = |(N - 1)/(N + 1)|^2, loss = Im eps / |eps|^2, and mu is
    proportional to w kappa."""
    a0, g, kp, V = 0.35, 0.01, 0.02, 40.0
    a = np.array([a0, 0.0])
    b = np.zeros(2)
    ne, el, eh = 80, 0.05, 0.8
    out = orc.o_spect_val(a, b, 1, kp, V, ne, el, eh, g)
    w = el + (eh - el) * (2 * np.arange(1, ne + 1) - 1) / (2 * ne)
    f = 2 * kp * V
    eps = 1 + 2 * a0 * f / (a0 ** 2 - (w + 1j * g) ** 2)
    assert np.allclose(out[:, 0], w * orc.HARTREE2EV, rtol=1e-15)
    assert np.allclose(out[:, 1] + 1j * out[:, 2], eps, rtol=1e-12, atol=1e-13)
    N = out[:, 3] + 1j * out[:, 4]
    assert np.allclose(N ** 2, eps, rtol=1e-12, atol=1e-13)
    assert np.allclose(out[:, 6], np.abs((N - 1) / (N + 1)) ** 2, rtol=1e-12)
    assert np.allclose(out[:, 7], eps.imag / np.abs(eps) ** 2, rtol=1e-12)
    ratio = out[:, 5] / (w * out[:, 4])
    assert np.allclose(ratio, ratio[0], rtol=1e-12)
