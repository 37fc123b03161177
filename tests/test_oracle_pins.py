"""Pins of the end-to-end conventions no reference fixture reaches, by what the
physics fixes rather than by restating the reference's index arithmetic.

  * The exchange/direct blocks of nbsemhsetup2 (OCEAN2/OCEAN_multiplet.f90:2007-2181)
    read through the action's decode of the packed index (nu fastest, then valence
    m, then alpha; OCEAN_multiplet.f90:1901-1907) with the alpha order of
    :343-355 (core spin slowest, core m, valence spin fastest): the direct part
    conserves both spins, the exchange part needs core spin = valence spin on both
    sides, and both conserve m_v - m_c (the Coulomb operator commutes with L_z of
    the electron-hole pair).  At the K edge the exchange block is known in closed
    form: 2l+1 eigenvalues 2 G^l / (2l+1) on the spin-matched pairs, one per
    valence m, all others zero.
  * The alpha order against the valence spin the long-range term selects
    (val_spin, long_range.f90:790-805): with the spin-2 Bloch functions zero,
    H_lr vanishes exactly on the alphas whose valence spin is +1/2.
  * The x-shift / phase / W geometry chain (OCEAN_bloch.f90:94-203,
    long_range.f90:1749-1841): translating the crystal by one x-grid step (the
    raw Bloch functions rolled by one point, the core site moved by one step) is a
    unitary change of gauge, so the long-range operator keeps its eigenvalues.
"""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import to_problem

from ocean_amd import setup_host as sh, synth  # noqa: E402


def alpha_spins(lc):
    """(2 cms, m_c, 2 vms) per alpha in the order of OCEAN_multiplet.f90:343-355:
    core spin outermost, then core m, valence spin innermost."""
    out = []
    for scs in (-1, 1):
        for mc in range(-lc, lc + 1):
            for svs in (-1, 1):
                out.append((scs, mc, svs))
    return out


def decode(ii, lc, lv, npl):
    """(alpha, m_v, nu) of packed index ii as the action reads it."""
    nu = ii % npl
    mv = (ii // npl) % (2 * lv + 1) - lv
    a = ii // (npl * (2 * lv + 1))
    return a, mv, nu


def _blocks(lc, lv, npl, seed, part):
    rng = np.random.default_rng(seed)
    fk, sf, gk, sg = {}, {}, {}, {}
    for k in range(2, min(2 * lc, 2 * lv) + 1, 2):
        m = rng.uniform(0.5, 2.0, (npl, npl))
        fk[k], sf[k] = (0.5 * (m + m.T)) * (part != "g"), 1.0
    for k in range(abs(lc - lv), lc + lv + 1, 2):
        m = rng.uniform(0.5, 2.0, (npl, npl))
        gk[k], sg[k] = (0.5 * (m + m.T)) * (part != "f"), 1.0
    return orc.nbsemhsetup2(lc, lv, npl, fk, sf, gk, sg)


@pytest.mark.parametrize("lc", [0, 1])
@pytest.mark.parametrize("lv", [0, 1, 2, 3])
@pytest.mark.parametrize("npl", [1, 2])
def test_exchange_blocks_obey_selection_rules(lc, lv, npl):
    labels = alpha_spins(lc)
    for part in ("f", "g"):
        X = _blocks(lc, lv, npl, 7 + 10 * lc + lv, part)
        mham = X.shape[0]
        assert mham == 4 * (2 * lc + 1) * (2 * lv + 1) * npl
        assert np.abs(X.imag).max() <= 1e-14 * max(np.abs(X).max(), 1e-300)  # Gaunt factors are real
        assert np.abs(X - X.T).max() <= 1e-13 * max(np.abs(X).max(), 1e-300)
        scale = np.abs(X).max()
        nonzero = 0
        for i1 in range(mham):
            a1, mv1, _ = decode(i1, lc, lv, npl)
            sc1, mc1, sv1 = labels[a1]
            for i2 in range(mham):
                a2, mv2, _ = decode(i2, lc, lv, npl)
                sc2, mc2, sv2 = labels[a2]
                if part == "f":
                    allowed = sc1 == sc2 and sv1 == sv2
                else:
                    allowed = sc1 == sv1 and sc2 == sv2
                allowed = allowed and (mv1 - mc1 == mv2 - mc2)
                if abs(X[i1, i2]) > 1e-12 * scale:
                    nonzero += 1
                    assert allowed, (part, i1, i2, labels[a1], mv1, labels[a2], mv2)
        if scale > 0:
            assert nonzero > 0


@pytest.mark.parametrize("lv", [0, 1, 2, 3])
def test_kedge_exchange_closed_form(lv):
    """K edge, one projector: X = G^l |s><s| summed over valence m, with the pair
    state |s> = |core down, v down> + |core up, v up> of one valence m: 2l+1
    eigenvalues 2 G / (2l+1), the rest zero, and each eigenvector of the nonzero
    ones lives on the two spin-matched alphas of a single m_v."""
    G = 1.7
    X = orc.nbsemhsetup2(0, lv, 1, {}, {}, {lv: np.array([[G]])}, {lv: 1.0})
    w, V = np.linalg.eigh(X)
    top = w > 1e-12
    assert top.sum() == 2 * lv + 1
    assert np.allclose(w[top], 2 * G / (2 * lv + 1), rtol=1e-13, atol=0)
    assert np.abs(w[~top]).max() < 1e-13
    labels = alpha_spins(0)
    # the nonzero eigenspace is spanned by one spin-matched pair state per m_v
    P = V[:, top] @ V[:, top].conj().T
    for mv in range(-lv, lv + 1):
        idx = [ii for ii in range(X.shape[0]) if decode(ii, 0, lv, 1)[1] == mv and
               labels[decode(ii, 0, lv, 1)[0]][0] == labels[decode(ii, 0, lv, 1)[0]][2]]
        assert len(idx) == 2
        s = np.zeros(X.shape[0])
        s[idx] = 1.0 / np.sqrt(2.0)
        assert abs(s @ P @ s - 1.0) < 1e-12  # the pair state lies in the eigenspace
        assert np.allclose(X @ s, 2 * G / (2 * lv + 1) * s, atol=1e-13)


@pytest.mark.parametrize("lc", [0, 1])
def test_long_range_valence_spin_of_alpha(lc):
    """nspn = 2 with the spin-2 Bloch functions zero: H_lr psi is exactly zero on
    the alphas whose valence spin is +1/2 (alpha order of OCEAN_multiplet.f90:343-355)
    and nonzero on the others."""
    inp = synth.make_inputs(nb=4, kmesh=(2, 2, 2), xmesh=(2, 3, 2), lc=lc, nspn=2, seed=12, xi_eV=1.0)
    inp.u[..., 1] = 0.0
    P = to_problem(inp, e0=False, do_mult=False)
    rng = np.random.default_rng(4)
    psi = np.asfortranarray(rng.standard_normal(inp.energy.shape) + 1j * rng.standard_normal(inp.energy.shape))
    h = orc.xact(P, psi, hflag=(1, 0, 0, 0, 0, 0))
    for a, (_, _, sv) in enumerate(alpha_spins(lc)):
        if sv == 1:
            assert not h[:, :, a].any(), a
        else:
            assert np.abs(h[:, :, a]).max() > 1e-6 * np.abs(h).max(), a


def _roll_x(u_raw, xmesh, axis):
    """u'(x) = u(x - one grid step along axis), x index ix fastest."""
    nb, nk, nx, ns = u_raw.shape
    t = u_raw.reshape((nb, nk) + tuple(xmesh) + (ns,), order="F")
    t = np.roll(t, 1, axis=2 + axis)
    return np.asfortranarray(t.reshape((nb, nk, nx, ns), order="F"))


def lr_eigenvalues(u_raw, tau, kmesh, xmesh, avec, rtab, ptab, eps):
    u, tau_c, xshift = orc.bloch_phase(u_raw, xmesh, kmesh, tau)
    W = orc.populate_W(avec, kmesh, xmesh, tau_c, xshift, rtab, ptab, eps)
    nb, nk = u.shape[0], u.shape[1]
    P = orc.Problem(nb, tuple(kmesh), tuple(xmesh), 0, 1, float(np.linalg.det(avec)), 1.0, u, W,
                    np.zeros((nb, nk, 4), complex, order="F"), np.ones((nb, nk, 4), order="F"), None,
                    e0=False, do_mult=False, long_range=True)
    return np.linalg.eigvalsh(orc.dense_H(P, hflag=(1, 0, 0, 0, 0, 0)))


@pytest.mark.parametrize("kmesh,xmesh,tau", [((2, 2, 2), (4, 3, 3), (0.13, 0.27, 0.41)),
                                             ((3, 2, 1), (3, 4, 2), (0.52, 0.08, 0.77))])
def test_one_grid_step_translation_keeps_lr_spectrum(kmesh, xmesh, tau):
    rng = np.random.default_rng(21)
    nb, nk, nx = 3, int(np.prod(kmesh)), int(np.prod(xmesh))
    u_raw = np.zeros((nb, nk, nx, 1), complex, order="F")
    for k in range(nk):
        q, _ = np.linalg.qr(rng.standard_normal((nx, nb)) + 1j * rng.standard_normal((nx, nb)))
        u_raw[:, k, :, 0] = q.T
    avec = np.array([[6.0, 0.3, 0.0], [0.0, 5.5, 0.2], [0.1, 0.0, 6.3]])
    rtab, ptab = synth.w_table()
    ref = lr_eigenvalues(u_raw, tau, kmesh, xmesh, avec, rtab, ptab, 4.0)
    assert np.abs(ref).max() > 0
    for d in range(3):
        t2 = np.array(tau, float)
        t2[d] += 1.0 / xmesh[d]
        got = lr_eigenvalues(_roll_x(u_raw, xmesh, d), t2, kmesh, xmesh, avec, rtab, ptab, 4.0)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), d
    # and the host/device setup path (setup_host.bloch_phase) agrees with the oracle's
    u1, tc1, xs1 = orc.bloch_phase(u_raw, xmesh, kmesh, tau)
    u2, tc2, xs2 = sh.bloch_phase(u_raw, xmesh, kmesh, np.asarray(tau))
    assert np.allclose(u1, u2, rtol=0, atol=1e-14) and np.array_equal(np.asarray(xs1), np.asarray(xs2))


def plane_wave_lr(kmesh, xmesh, tau, avec, conj_phase=False):
    """(dense H_lr of the oracle for one constant band, closed-form plane-wave matrix).

    With u_raw = N_x^-1/2 for every k the phased Bloch functions are the plane waves
    exp(i k.r) on the supercell grid, so <k|H_lr|k'> must be the Fourier sum
    -(N_k N_x)^-1 sum_r exp(-2 pi i (k - k').r) W(r - tau) over the grid points r at
    which lr_populate_W2 evaluated W (positions from its own geometry, r = tau + alf)."""
    kmesh, xmesh = tuple(kmesh), tuple(xmesh)
    nk, nx = int(np.prod(kmesh)), int(np.prod(xmesh))
    u_raw = np.full((1, nk, nx, 1), 1.0 / np.sqrt(nx), complex, order="F")
    u, tau_c, xshift = orc.bloch_phase(u_raw, xmesh, kmesh, tau)
    if conj_phase:  # a planted mistake: the supercell phase with the opposite sign
        u = np.asfortranarray(np.conj(u))
    rtab, ptab = synth.w_table()
    W = orc.populate_W(avec, kmesh, xmesh, tau_c, xshift, rtab, ptab, 4.0)
    P = orc.Problem(1, kmesh, xmesh, 0, 1, float(np.linalg.det(avec)), 1.0, u, W,
                    np.zeros((1, nk, 4), complex, order="F"), np.ones((1, nk, 4), order="F"), None,
                    e0=False, do_mult=False)
    H = orc.dense_H(P, hflag=(1, 0, 0, 0, 0, 0))[:nk, :nk]  # alpha 0 block
    _, _, cart = orc._w_geometry(avec, kmesh, xmesh, tau_c, xshift)
    # the physical positions: W was evaluated at r - tau for these grid points r
    pos = np.einsum("ij,kxj->kxi", np.linalg.inv(np.asarray(avec)), cart) + np.asarray(tau)[None, None, :]
    kv = np.array([(a / kmesh[0], b / kmesh[1], c / kmesh[2]) for a in range(kmesh[0])
                   for b in range(kmesh[1]) for c in range(kmesh[2])])
    Hc = np.zeros((nk, nk), complex)
    for k in range(nk):
        for kp in range(nk):
            ph = np.exp(-2j * np.pi * np.einsum("i,kxi->kx", kv[k] - kv[kp], pos))
            Hc[k, kp] = -np.sum(ph * W) / (nk * nx)
    return H, Hc


@pytest.mark.parametrize("kmesh,xmesh,tau", [((2, 2, 2), (4, 3, 3), (0.13, 0.27, 0.41)),
                                             ((3, 2, 1), (3, 4, 2), (0.52, 0.08, 0.77)),
                                             ((4, 1, 3), (2, 3, 5), (0.9, 0.45, 0.05))])
def test_plane_waves_give_the_fourier_sum_of_w(kmesh, xmesh, tau):
    avec = np.array([[6.0, 0.3, 0.0], [0.0, 5.5, 0.2], [0.1, 0.0, 6.3]])
    H, Hc = plane_wave_lr(kmesh, xmesh, tau, avec)
    assert np.abs(H - Hc).max() < 1e-13 * np.abs(Hc).max()
    # planted mistake: the opposite phase sign breaks it
    Hb, _ = plane_wave_lr(kmesh, xmesh, tau, avec, conj_phase=True)
    assert np.abs(Hb - Hc).max() > 1e-3 * np.abs(Hc).max()
