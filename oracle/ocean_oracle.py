"""CPU restatement of OCEAN2's core-level BSE hot path (the oracle).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product (``ocean_amd``) never imports it; the GPU path fails loudly
when its HIP library is missing instead of falling back here.

Every function restates one reference routine and cites it (paths relative to
the OCEAN 3.2.1 tree).  Arrays use the reference's index order (math order
below) and are allocated Fortran-ordered so that their memory equals the
reference's column-major layout:

    u[n, k, x, s]         re/im_bloch_state(nb, nk, nx, nspn)   long_range.f90:186-187
    W[k, x]               W(nk, nx)                            long_range.f90:188
    psi[n, k, a]          OCEAN_vector %r/%i (nb, nk, nalpha)   OCEAN_psi.f90:64-65
    c[n, k, nu, m, l, s]  mpcr/mpci(nb, nk, npmax, -lmax:lmax, lmin:lmax, nspn)
                                                               OCEAN_multiplet.f90:401-402

k index: k1 outer, k3 inner (OCEAN_bloch.f90:171-175); x index: ix fastest
(OCEAN_bloch.f90:133-156); alpha = (core spin, core m, valence spin), valence
spin fastest (OCEAN_multiplet.f90:343-355).  Energies are in Hartree.

Pinning (what checks each function; "parity unpinned" where none does):
  write_core           TEST/Fe/Results abslanc->absspct fixture, Lorentzian closed form
  write_core_full      same fixture (Spect/SPKK columns with gamfcn at the Fe run's own
                       valence density 16/acell^3 and diemac 1e5), cross-section columns
                       against the CODATA products; Voigt form against scipy's
                       voigt_profile for a one-level fraction (within the +-3 sigma
                       truncation) and its fine-grid reuse against direct evaluation
  gamfcn               the same fixture (its only consumer), zero below the Fermi level
  calc_spect_core      write_core at the midpoint grid
  energies_core_xas    the allow mask's defining inequality on a seeded directory
                       (parity of the branch itself is by restatement only)
  o_spect_core         the Fe absspct fixture (Spect, Spect(0) and SPKK columns), the
                       resonant-minus-anti-resonant Lorentzian closed form for one level
  o_spect_val          the Lorentz-oscillator closed form of eps for one level; n + i kappa
                       squared equals eps, R and the loss function from their definitions
  redtrid              TEST/Fe/Results abslanc eigenvalue block
  haydock              dense Lanczos/eigendecomposition brute force (tests/test_oracle.py)
  haydock_abc          reduction to haydock on a Hermitian problem; the three-term relations
                       of both sides with the dense H and backwards operator; closed forms
                       of a(0), c(1), b(1), a(1) in the moments <v|A^k|v> (tests/test_oracle_abc.py)
  write_core_abc       write_core for real c = b; the (0,0) resolvent element of the
                       tridiagonal matrix when b(n) c(n) = 0
  lr_act               dense operator brute force, loop-literal lr_kernel restatement
  lr_act_obf2          -N_k lr_act with u = sum_o obf2u obf phs (the Bloch functions the
                       reference's own use_fake_obf path builds, OCEAN_obf.f90:84-96),
                       Hermiticity of the masked operator
  obf_phase            bloch_phase's closed form for even meshes; the odd-kmesh(1)
                       xshift by hand
  xact_obf             xact with the multiplet term off, plus the two terms added after
  obf_lrload_bloch     one basis function (plain normalisation) and a hand-summed
                       two-function, one-point case
  fg_combo/new_soact   dense operator brute force, Hermiticity
  gaunt/threey         Wigner-3j closed form vs Gauss-Legendre quadrature restatement
  soprep               closed-form <L.S> eigenvalues
  populate_W/intval    closed forms of the geometry and the interpolant
  bloch_phase          closed form (phase of a plane wave)
  gmres_do_recycle     dense direct solve of (E + iG - H) x = b
  spectral_grad        central finite differences of spectral_sigma (no reference
                       gradient exists; the loss itself rests on the pins above)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# OCEAN_constants.f90:13-20 (2022 CODATA)
EV2HARTREE = 0.036749322175666
HARTREE2EV = 27.211386245980
# OCEAN_constants.f90:25 (commented older constant, used by the TEST/Fe outputs)
HARTREE2EV_OLD = 27.21138386


# ----------------------------------------------------------------------------
# small helpers
# ----------------------------------------------------------------------------
def fzeros(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def intval(xtab, ytab, x, lopt="cap", hopt="cap"):
    """Linear interpolation with 'cap'/'ext' ends.

    CNBSE/zcmsnmult/zf90/intval.f90:8-57.  The bracket search starts at ii=1 and
    advances while xtab(ii+1) < x.
    """
    xtab = np.asarray(xtab, np.float64)
    ytab = np.asarray(ytab, np.float64)
    x = np.asarray(x, np.float64)
    n = xtab.shape[0]
    below = x < xtab[0]
    above = x > xtab[n - 1]
    idx = np.searchsorted(xtab, x, side="left")
    ii = np.clip(idx - 1, 0, n - 2)
    if lopt == "ext":
        ii = np.where(below, 0, ii)
    if hopt == "ext":
        ii = np.where(above, n - 2, ii)
    if (lopt == "err" and np.any(below)) or (hopt == "err" and np.any(above)):
        raise ValueError("intval: out of table range")
    rat = (x - xtab[ii]) / (xtab[ii + 1] - xtab[ii])
    y = ytab[ii] + rat * (ytab[ii + 1] - ytab[ii])
    if lopt == "cap":
        y = np.where(below, ytab[0], y)
    if hopt == "cap":
        y = np.where(above, ytab[n - 1], y)
    return y


def getomega(avec):
    """Cell volume |det avec|, avec(:, i) = lattice vector i.

    CNBSE/zcmsnmult/zf90/getomega.f90:8-26.
    """
    a = np.asarray(avec, np.float64)
    om = 0.0
    for i in range(3):
        j = (i + 1) % 3
        k = (j + 1) % 3
        om += a[i, 0] * a[j, 1] * a[k, 2]
        om -= a[i, 0] * a[k, 1] * a[j, 2]
    return abs(om)


def alpha_labels(lc):
    """(cms, cml, vms) per alpha, valence spin fastest.

    OCEAN_multiplet.f90:343-355.
    """
    out = []
    for icms in (-1, 1):
        for icml in range(-lc, lc + 1):
            for ivms in (-1, 1):
                out.append((0.5 * icms, float(icml), 0.5 * ivms))
    return out


def val_spin_of_alpha(lc, nspn):
    """0-based valence-spin index per alpha.  long_range.f90:793-805."""
    na = 4 * (2 * lc + 1)
    if nspn == 1:
        return np.zeros(na, dtype=np.int64)
    return np.array([a % 2 for a in range(na)], dtype=np.int64)


# ----------------------------------------------------------------------------
# 3-D FFT over the k-mesh with the reference's sign/normalisation conventions
# ----------------------------------------------------------------------------
def fft_backward(v, kmesh, axis=0):
    """FFTW_BACKWARD (exp(+2 pi i ...)), unnormalised, over (k1,k2,k3), k3 fastest.

    Modules/FFT_wrapper.f90:67-72 (plan over dims (k3,k2,k1) reversed),
    :185-215 (FFT_wrapper_single with norm=.false.).
    """
    k1, k2, k3 = kmesh
    v = np.moveaxis(v, axis, -1)
    sh = v.shape[:-1]
    t = v.reshape(sh + (k1, k2, k3))
    t = np.fft.ifftn(t, axes=(-3, -2, -1)) * (k1 * k2 * k3)
    return np.moveaxis(t.reshape(sh + (k1 * k2 * k3,)), -1, axis)


def fft_forward(v, kmesh, axis=0):
    """FFTW_FORWARD (exp(-2 pi i ...)), unnormalised.  Modules/FFT_wrapper.f90:185-215."""
    k1, k2, k3 = kmesh
    v = np.moveaxis(v, axis, -1)
    sh = v.shape[:-1]
    t = v.reshape(sh + (k1, k2, k3))
    t = np.fft.fftn(t, axes=(-3, -2, -1))
    return np.moveaxis(t.reshape(sh + (k1 * k2 * k3,)), -1, axis)


def dft_matrix_backward(kmesh):
    """Dense matrix of fft_backward (brute-force check of the convention)."""
    k1, k2, k3 = kmesh
    idx = np.array([(a, b, c) for a in range(k1) for b in range(k2) for c in range(k3)], float)
    ph = (np.outer(idx[:, 0], idx[:, 0]) / k1 + np.outer(idx[:, 1], idx[:, 1]) / k2
          + np.outer(idx[:, 2], idx[:, 2]) / k3)
    return np.exp(2j * np.pi * ph)


# ----------------------------------------------------------------------------
# Setup restatements
# ----------------------------------------------------------------------------
def bloch_phase(u_raw, xmesh, kmesh, tau, xshift_override=(1, 1, 1)):
    """Centre the core site and apply the supercell phase to u.

    OCEAN_bloch.f90:54-219 (OCEAN_bloch_lrLOAD).  Returns (u, tau_new, xshift).
    u_raw[n, k, x, s] is the content of con.u2.dat (OCEAN_bloch.f90:998-1115).
    """
    xmesh = tuple(int(v) for v in xmesh)
    kmesh = tuple(int(v) for v in kmesh)
    tau = np.array(tau, np.float64)
    xshift = np.zeros(3, np.int64)
    for d in range(3):
        if kmesh[d] % 2 == 0:
            xshift[d] = math.floor(float(xmesh[d]) * tau[d])
        else:
            xshift[d] = math.floor(float(xmesh[d]) * (tau[d] - 0.5))
    xshift = xshift * np.array(xshift_override, np.int64)
    tau_new = tau - xshift / np.array(xmesh, np.float64)

    def wrap(i, d):  # i 1-based
        targ = i - xshift[d]
        if targ > xmesh[d]:
            return -xmesh[d]
        if targ < 1:
            return xmesh[d]
        return 0

    nx = xmesh[0] * xmesh[1] * xmesh[2]
    nk = kmesh[0] * kmesh[1] * kmesh[2]
    xoff = np.zeros((nx, 3), np.float64)
    it = 0
    for iz in range(1, xmesh[2] + 1):
        zph = wrap(iz, 2)
        for iy in range(1, xmesh[1] + 1):
            yph = wrap(iy, 1)
            for ix in range(1, xmesh[0] + 1):
                xph = wrap(ix, 0)
                xoff[it] = (xph + ix - 1, yph + iy - 1, zph + iz - 1)
                it += 1
    kidx = np.array([(a, b, c) for a in range(kmesh[0]) for b in range(kmesh[1])
                     for c in range(kmesh[2])], np.float64)
    phsx = 2.0 * np.pi * np.outer(kidx[:, 0], xoff[:, 0]) / (xmesh[0] * kmesh[0])
    phsy = 2.0 * np.pi * np.outer(kidx[:, 1], xoff[:, 1]) / (xmesh[1] * kmesh[1])
    phsz = 2.0 * np.pi * np.outer(kidx[:, 2], xoff[:, 2]) / (xmesh[2] * kmesh[2])
    ph = phsx + phsy + phsz  # [k, x]
    rot = np.cos(ph) + 1j * np.sin(ph)
    assert u_raw.shape[1] == nk and u_raw.shape[2] == nx
    u = np.asfortranarray(u_raw * rot[None, :, :, None])
    return u, tau_new, xshift


def _w_geometry(avec, kmesh, xmesh, tau, xshift, pbc=None):
    """r(k,x) = |A (R_k + f_x - tau)| and the pbc prefactor.  long_range.f90:1744-1817."""
    avec = np.asarray(avec, np.float64)
    amet = avec.T @ avec  # amet(i,j) = avec(:,i).avec(:,j)
    kmesh = tuple(int(v) for v in kmesh)
    xmesh = tuple(int(v) for v in xmesh)
    if pbc is None:
        pbc = kmesh
    else:
        pbc = tuple(kmesh[i] if pbc[i] != 0 else 0 for i in range(3))
    fr = []
    for iz in range(1, xmesh[2] + 1):
        zt = iz - xshift[2]
        zt = zt - xmesh[2] if zt > xmesh[2] else (zt + xmesh[2] if zt < 1 else zt)
        for iy in range(1, xmesh[1] + 1):
            yt = iy - xshift[1]
            yt = yt - xmesh[1] if yt > xmesh[1] else (yt + xmesh[1] if yt < 1 else yt)
            for ix in range(1, xmesh[0] + 1):
                xt = ix - xshift[0]
                xt = xt - xmesh[0] if xt > xmesh[0] else (xt + xmesh[0] if xt < 1 else xt)
                fr.append(((xt - 1) / xmesh[0], (yt - 1) / xmesh[1], (zt - 1) / xmesh[2]))
    fr = np.array(fr)  # [x, 3]

    def kk_of(n):
        kk = np.arange(n)
        kk = np.where(kk >= (n + 1) // 2, kk - n, kk)
        if n == 1:
            kk[:] = 0
        return kk

    kk1, kk2, kk3 = (kk_of(kmesh[d]) for d in range(3))
    xk = np.array([(a, b, c) for a in kk1 for b in kk2 for c in kk3], np.float64)  # [k,3]
    p1 = np.array([a <= pbc[0] for a in kk1 for b in kk2 for c in kk3])
    p2 = np.array([(b <= pbc[1]) for a in kk1 for b in kk2 for c in kk3]) & p1
    p3 = np.array([(c <= pbc[2]) for a in kk1 for b in kk2 for c in kk3]) & p2
    alf = xk[:, None, :] + fr[None, :, :] - np.asarray(tau, np.float64)[None, None, :]
    r = np.sqrt(np.einsum("kxi,ij,kxj->kx", alf, amet, alf))
    cart = np.einsum("ij,kxj->kxi", avec, alf)
    return r, p3.astype(np.float64), cart


def populate_W(avec, kmesh, xmesh, tau, xshift, rtab, ptab, epsilon0,
               interaction_scale=1.0, isolated=False, iso_cut=0.0, pbc=None,
               epsilon3d=None):
    """W(k, x) for the screened core-hole potential.

    long_range.f90:1659-1853 (lr_populate_W2); interpolation by intval 'cap','cap'
    (:1834), 1/(eps r) tail beyond rtab(n) (:1819-1832), isolated cutoff (:1818),
    x-shift wrap (:1749-1772), kk >= (N+1)/2 wrap (:1783, :1795, :1806).
    """
    epsi = 1.0 / epsilon0
    rtab = np.asarray(rtab, np.float64)
    ptab = np.asarray(ptab, np.float64)
    r, pref, cart = _w_geometry(avec, kmesh, xmesh, tau, xshift, pbc)
    potn = intval(rtab, ptab, r, "cap", "cap")
    far = r > rtab[-1]
    rr = np.where(r > 0, r, 1.0)
    if epsilon3d is not None:
        d = cart / rr[..., None]
        e3 = np.asarray(epsilon3d, np.float64)
        pt = (d[..., 0] ** 2 / e3[0] + d[..., 1] ** 2 / e3[1] + d[..., 2] ** 2 / e3[2]) / rr
        blend = r < rtab[-1] + 5.0
        pt = np.where(blend, pt * (r - rtab[-1]) / 5.0 + (rtab[-1] + 5.0 - r) * epsi / (5.0 * rr), pt)
        potn = np.where(far, pt, potn)
    else:
        potn = np.where(far, epsi / rr, potn)
    if isolated:
        potn = np.where(r > iso_cut, 0.0, potn)
    W = potn * pref[:, None] * interaction_scale
    return np.asfortranarray(W)


def populate_W_weights(avec, kmesh, xmesh, tau, xshift, rtab, interaction_scale=1.0,
                       isolated=False, iso_cut=0.0, pbc=None):
    """dW(k,x)/dptab as two hat weights per (k,x).

    W is linear in ptab through intval (CNBSE/zcmsnmult/zf90/intval.f90:47-53) for
    r <= rtab(n), capped at ptab(1) below rtab(1); the 1/(eps r) tail and the
    isolated cutoff carry no ptab dependence (long_range.f90:1818-1840).
    Returns (i0, w0, w1): dW = w0 dptab[i0] + w1 dptab[i0 + 1].
    """
    rtab = np.asarray(rtab, np.float64)
    n = rtab.shape[0]
    r, pref, _ = _w_geometry(avec, kmesh, xmesh, tau, xshift, pbc)
    idx = np.searchsorted(rtab, r, side="left")
    i0 = np.clip(idx - 1, 0, n - 2)
    rat = (r - rtab[i0]) / (rtab[i0 + 1] - rtab[i0])
    w0 = 1.0 - rat
    w1 = rat
    below = r < rtab[0]
    w0 = np.where(below, 1.0, w0)
    w1 = np.where(below, 0.0, w1)
    dead = r > rtab[-1]
    if isolated:
        dead = dead | (r > iso_cut)
    s = pref[:, None] * interaction_scale
    w0 = np.where(dead, 0.0, w0) * s
    w1 = np.where(dead, 0.0, w1) * s
    return i0, w0, w1


# ----------------------------------------------------------------------------
# Multiplet setup restatements
# ----------------------------------------------------------------------------
def getprefs(lmax):
    """Ylm normalisation prefactors.  CNBSE/zcmsnmult/zf90/newgetprefs.f90:8-40."""
    prefs = np.zeros(1001)
    for l in range(lmax + 1):
        prefs[l] = (2 * l + 1) / (4.0 * np.pi)
        lamold = l
        for m in range(1, l + 1):
            lam = 10 * m + l
            prefs[lam] = prefs[lamold] / ((l - m + 1) * (l + m))
            lamold = lam
    for l in range(lmax + 1):
        for m in range(l + 1):
            lam = 10 * m + l
            prefs[lam] = math.sqrt(prefs[lam])
    return prefs


_YLM_POLY = {  # newgetylm.f90 case table: f(u) for lam = 10*|m| + l
    0: lambda u: 1.0 + 0 * u,
    11: lambda u: -1.0 + 0 * u, 1: lambda u: u,
    22: lambda u: 3.0 + 0 * u, 12: lambda u: -3 * u, 2: lambda u: (3 * u ** 2 - 1) / 2,
    33: lambda u: -15.0 + 0 * u, 23: lambda u: 15 * u, 13: lambda u: -(15 * u ** 2 - 3) / 2,
    3: lambda u: (5 * u ** 3 - 3 * u) / 2,
    44: lambda u: 105.0 + 0 * u, 34: lambda u: -105 * u, 24: lambda u: (105 * u ** 2 - 15) / 2,
    14: lambda u: -(35 * u ** 3 - 15 * u) / 2, 4: lambda u: (35 * u ** 4 - 30 * u ** 2 + 3) / 8,
    55: lambda u: -945.0 + 0 * u, 45: lambda u: 945 * u, 35: lambda u: -(945 * u ** 2 - 105) / 2,
    25: lambda u: (315 * u ** 3 - 105 * u) / 2, 15: lambda u: -(315 * u ** 4 - 210 * u ** 2 + 15) / 8,
    5: lambda u: (63 * u ** 5 - 70 * u ** 3 + 15 * u) / 8,
}


def getylm(l, m, x, y, z, prefs):
    """Complex Ylm, Condon-Shortley phase.  CNBSE/zcmsnmult/zf90/newgetylm.f90."""
    if l > 5:
        raise ValueError("l > 5 not yet allowed")
    r = np.sqrt(x * x + y * y + z * z)
    r = np.where(r == 0, 1.0, r)
    xr, yr, zr = x / r, y / r, z / r
    mm = abs(m)
    lam = 10 * mm + l
    f = _YLM_POLY.get(lam, lambda u: 0 * u)(zr)
    ylm = prefs[lam] * f + 0j
    if m > 0:
        for _ in range(m):
            ylm = ylm * (xr + 1j * yr)
    if m < 0:
        for _ in range(mm):
            ylm = -ylm * (xr - 1j * yr)
    return ylm


def threey(ll, ml, lmid, mmid, lr, mr, conjug, x, w, prefs):
    """Gaunt integral int conj(Y_ll,ml) Y_lmid,mmid(^*) Y_lr,mr by quadrature.

    CNBSE/zcmsnmult/zf90/newthreey.f90:8-52: Gauss-Legendre in cos(theta), phi
    integral done analytically (2 pi) once m conservation holds.
    """
    mdo = (ml == mr - mmid) if conjug else (ml == mr + mmid)
    ldo = False
    j = ll + lr
    while j >= max(abs(mmid), abs(ll - lr)):
        if j == lmid:
            ldo = True
        j -= 2
    if not (ldo and mdo):
        return 0j
    ct = np.asarray(x, np.float64)
    st = np.sqrt(1.0 - ct ** 2)
    rx, ry, rz = st, 0.0 * st, ct
    yl = getylm(ll, ml, rx, ry, rz, prefs)
    ym = getylm(lmid, mmid, rx, ry, rz, prefs)
    if conjug:
        ym = np.conj(ym)
    yr = getylm(lr, mr, rx, ry, rz, prefs)
    return complex(np.sum(w * np.conj(yl) * ym * yr) * 2 * np.pi)


def wigner3j(j1, j2, j3, m1, m2, m3):
    """Racah formula (closed form used to pin threey)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    pre = (-1) ** (j1 - j2 - m3) * math.sqrt(
        f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3) / f(j1 + j2 + j3 + 1)
        * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3))
    s = 0.0
    for k in range(0, j1 + j2 + j3 + 2):
        d = [k, j1 + j2 - j3 - k, j1 - m1 - k, j2 + m2 - k, j3 - j2 + m1 + k, j3 - j1 - m2 + k]
        if min(d) < 0:
            continue
        s += (-1) ** k / (f(d[0]) * f(d[1]) * f(d[2]) * f(d[3]) * f(d[4]) * f(d[5]))
    return pre * s


def gaunt_closed(ll, ml, lmid, mmid, lr, mr, conjug):
    """Closed form of threey via 3j symbols (Condon-Shortley Ylm)."""
    # int conj(Y_l1m1) Y_l2m2 Y_l3m3 = (-1)^m1 sqrt((2l1+1)(2l2+1)(2l3+1)/4pi)
    #                                  (l1 l2 l3;0 0 0)(l1 l2 l3;-m1 m2 m3)
    if conjug:  # conj(Y_l,m) = (-1)^m Y_l,-m
        sgn = (-1) ** mmid
        mm = -mmid
    else:
        sgn = 1
        mm = mmid
    val = ((-1) ** ml * math.sqrt((2 * ll + 1) * (2 * lmid + 1) * (2 * lr + 1) / (4 * math.pi))
           * wigner3j(ll, lmid, lr, 0, 0, 0) * wigner3j(ll, lmid, lr, -ml, mm, mr))
    return sgn * val


def gauss_legendre(npt=16):
    """The Pquadrature rule (weights renormalised to sum 2, OCEAN_multiplet.f90:2055-2063)."""
    x, w = np.polynomial.legendre.leggauss(npt)
    return x, w * 2.0 / w.sum()


def mham_size(lc, lv, np_l):
    """mham(l) = 4 (2 lc + 1)(2 l + 1) nproj(l).  OCEAN_multiplet.f90:325."""
    return 4 * (2 * lc + 1) * (2 * lv + 1) * np_l


def hamiltonian_labels(lc, lv, np_l):
    """(cms, cml, vms, vml, vnu) per exchange index.  OCEAN_multiplet.f90:358-373."""
    out = []
    for (cms, cml, vms) in alpha_labels(lc):
        for ivml in range(-lv, lv + 1):
            for nu in range(1, np_l + 1):
                out.append((cms, cml, vms, float(ivml), nu))
    return out


def nbsemhsetup2(lc, lv, np_l, fk, scfk, gk, scgk, scfac=1.0, tdlda=False, quad=None,
                 use_closed_form=False):
    """Exchange/direct multiplet matrix mh_l(i1, i2) (eV) for one valence l.

    OCEAN_multiplet.f90:2007-2181.  fk[k] / gk[k] are (np, np) Slater matrices and
    scfk[k] / scgk[k] their scale factors, keyed by k.  Returns complex (mham, mham)
    with element [i1, i2] as stored in mhr_l/mhi_l.
    """
    mham = mham_size(lc, lv, np_l)
    labels = hamiltonian_labels(lc, lv, np_l)
    if quad is None:
        quad = gauss_legendre(16)
    x, w = quad
    maxll = min(max(2 * lc, 2 * lv), 5)
    prefs = getprefs(maxll)
    kfh = min(2 * lc, 2 * lv)
    kfl = 2
    if tdlda:
        kfl = kfh = 0
    kgh = lc + lv
    kgl = abs(lc - lv)
    ty = (lambda *a: gaunt_closed(*a)) if use_closed_form else \
        (lambda l1, m1, k, mk, l3, m3, cj: threey(l1, m1, k, mk, l3, m3, cj, x, w, prefs))
    mh = np.zeros((mham, mham), np.complex128, order="F")
    for i1 in range(mham):
        nu1 = labels[i1][4]
        for i2 in range(mham):
            nu2 = labels[i2][4]
            if kfh >= kfl:
                l1, m1, s1 = lc, int(round(labels[i2][1])), int(round(2 * labels[i2][0]))
                l2, m2, s2 = lv, int(round(labels[i1][3])), int(round(2 * labels[i1][2]))
                l3, m3, s3 = lc, int(round(labels[i1][1])), int(round(2 * labels[i1][0]))
                l4, m4, s4 = lv, int(round(labels[i2][3])), int(round(2 * labels[i2][2]))
                if s1 == s3 and s2 == s4:
                    mk = m1 - m3
                    if m1 + m2 == m3 + m4:
                        for k in range(kfl, kfh + 1, 2):
                            if abs(mk) <= k:
                                ffk = scfk[k] * fk[k][nu1 - 1, nu2 - 1] * scfac
                                f1 = ty(l1, m1, k, mk, l3, m3, False)
                                f2 = ty(l2, m2, k, mk, l4, m4, True)
                                mh[i1, i2] += -ffk * f1 * f2 * (4 * np.pi / (2 * k + 1))
            if kgh >= kgl:
                l1, m1, s1 = lc, int(round(labels[i2][1])), int(round(2 * labels[i2][0]))
                l2, m2, s2 = lv, int(round(labels[i1][3])), int(round(2 * labels[i1][2]))
                l3, m3, s3 = lv, int(round(labels[i2][3])), int(round(2 * labels[i2][2]))
                l4, m4, s4 = lc, int(round(labels[i1][1])), int(round(2 * labels[i1][0]))
                if s1 == s3 and s2 == s4:
                    mk = m1 - m3
                    if m1 + m2 == m3 + m4:
                        for k in range(kgl, kgh + 1, 2):
                            if abs(mk) <= k:
                                ggk = scgk[k] * gk[k][nu1 - 1, nu2 - 1] * scfac
                                f1 = ty(l1, m1, k, mk, l3, m3, False)
                                f2 = ty(l2, m2, k, mk, l4, m4, True)
                                mh[i1, i2] += ggk * f1 * f2 * (4 * np.pi / (2 * k + 1))
    return mh


def jimel(j, m, mp, i):
    """<j m | J_i | j mp> (i = x, y, z).  CNBSE/zcmsnmult/zf90/jimel.f90:10-38."""
    x = abs(j * (j + 1) - m * mp)
    mdiff = int(round(m - mp))
    c = 0j
    if i == 1:
        if mdiff in (1, -1):
            c = 0.5 * math.sqrt(x)
    elif i == 2:
        if mdiff == 1:
            c = 0.5 * math.sqrt(x) / 1j
        elif mdiff == -1:
            c = -0.5 * math.sqrt(x) / 1j
    elif i == 3:
        if mdiff == 0:
            c = m
    return complex(c)


def soprep(lc, xi_eV, life_time_eV=None):
    """Spin-orbit alpha-mixing matrix S = somelr + i someli (Hartree), [ic, jc].

    OCEAN_multiplet.f90:572-690 (OCEAN_soprep).  xi is read in eV from bse.in
    line 2 (:298-303).  life_time (core_broaden.ipt) enables the imaginary part.
    Returns (somelr, someli, have_spin_orbit).
    """
    xi = xi_eV * EV2HARTREE
    labels = alpha_labels(lc)
    na = len(labels)
    have_so = xi >= 1.0e-12
    if life_time_eV is not None:
        have_so = True
        lt = np.asarray(life_time_eV, np.float64) * EV2HARTREE
        l_alpha = 2.0 / (2 * lc + 1) * (lt[0] - lt[1])
        l_beta = (lc + 1) / (2 * lc + 1) * lt[0] + lc / (2 * lc + 1) * lt[1]
    else:
        l_alpha = 0.0
        l_beta = 0.0
    sr = np.zeros((na, na), order="F")
    si = np.zeros((na, na), order="F")
    for ic in range(na):
        for jc in range(na):
            if labels[ic][2] == labels[jc][2]:
                ct = 0j
                for i in (1, 2, 3):
                    ct += jimel(float(lc), labels[jc][1], labels[ic][1], i) * \
                        jimel(0.5, labels[jc][0], labels[ic][0], i)
                sr[ic, jc] = -xi * ct.real
                si[ic, jc] = -ct.real * l_alpha
                if ic == jc:
                    si[ic, jc] -= l_beta
    return sr, si, have_so


def nbsemkcmel(v, rv, rmax, rad, phi):
    """Trim the core-hole potential and integrate cmel for one l.

    OCEAN2/nbsemkcmel.f90:8-209.  v, rv: rpot table; rad: log radial grid;
    phi[i, nu]: projector radial functions on rad.  Returns (vtrim, cmel_eV, nmel).
    """
    v = np.asarray(v, np.float64)
    rv = np.asarray(rv, np.float64)
    nv = v.shape[0]
    # binary search for the first rv >= rmax (nbsemkcmel.f90:98-110), 1-based j
    j, k = 1, nv
    i = (j + k) // 2
    while j < k:
        i = (j + k) // 2
        if rv[i - 1] < rmax:
            j = i + 1
        elif rv[i - 1] > rmax:
            k = i - 1
        else:
            break
    if rv[j - 1] < rmax:
        j += 1
    vrun = v[j - 1]
    vtrim = v.copy()
    vdiff = np.zeros(nv)
    vtrim[: j - 1] = vrun
    vdiff[: j - 1] = v[: j - 1] - vrun
    bwgt = np.array([28.0, 64.0, 24.0, 64.0]) / 45.0
    nr = rad.shape[0]
    dl = math.log(rad[nr - 1] / rad[0]) / (nr - 1)
    xx = 1 + round(math.log(rmax / rad[0]) / dl)
    ix = int(xx)
    dr = np.zeros(nr)
    dr[0] = 0.5 * dl * rad[0]
    dr[1: ix - 1] = dl * rad[1: ix - 1]
    dr[ix - 1] = 0.5 * dl * rad[ix - 1]
    val = intval(rv, vdiff, rad[:ix], "err", "err")
    nnu = phi.shape[1]
    cmel = np.zeros((nnu, nnu))
    nmel = np.zeros((nnu, nnu))
    for nu in range(nnu):
        for mu in range(nnu):
            su = 0.0
            su1 = 0.0
            for ii0 in range(ix, 4, -4):
                for jj in range(4):
                    q = ii0 - jj - 1
                    t = phi[q, mu] * phi[q, nu] * val[q]
                    su += bwgt[jj] * dr[q] * rad[q] ** 2 * t
                    su1 += bwgt[jj] * dr[q] * rad[q] ** 2 * phi[q, mu] * phi[q, nu]
            cmel[mu, nu] = su * HARTREE2EV
            nmel[mu, nu] = su1
    return vtrim, cmel, nmel


# ----------------------------------------------------------------------------
# Problem container
# ----------------------------------------------------------------------------
@dataclass
class Multiplet:
    """Module state of OCEAN_multiplet after OCEAN_create_central (Hartree)."""
    lmin: int
    lmax: int
    nproj: list            # nproj[l - lmin]
    lvl: int
    lvh: int
    c: np.ndarray          # complex (nb, nk, npmax, 2lmax+1, nl, nspn), F-order
    mpm: np.ndarray        # real (npmax, npmax, nl)
    X: list                # complex (mham_l, mham_l) per l in lvl..lvh, [i1, i2] storage
    somelr: np.ndarray     # (na, na) [ic, jc]
    someli: np.ndarray
    have_spin_orbit: bool = True

    @property
    def npmax(self):
        return max(self.nproj)

    def packed(self):
        """mhr/mhi flat arrays and jbeg/mham (OCEAN_multiplet.f90:318-327, :2125-2175)."""
        mhr, mhi, jbeg, mham = [], [], [], []
        j = 0
        for Xl in self.X:
            jbeg.append(j)
            mham.append(Xl.shape[0])
            flat = np.asfortranarray(Xl).ravel(order="F")
            mhr.append(flat.real)
            mhi.append(flat.imag)
            j += flat.size
        return (np.concatenate(mhr), np.concatenate(mhi), np.array(jbeg), np.array(mham))


@dataclass
class Problem:
    """Everything the core-level action reads (o_system fields + module state)."""
    nb: int
    kmesh: tuple
    xmesh: tuple
    lc: int
    nspn: int
    celvol: float
    inter_scale: float
    u: np.ndarray            # complex (nb, nk, nx, nspn), phased
    W: np.ndarray            # real (nk, nx)
    energy: np.ndarray       # complex (nb, nk, na): E + i Sigma''
    allow: np.ndarray        # real (nb, nk, na)
    mult: Multiplet | None
    e0: bool = True
    do_mult: bool = True
    long_range: bool = True
    extras: dict = field(default_factory=dict)

    @property
    def nk(self):
        return int(np.prod(self.kmesh))

    @property
    def nx(self):
        return int(np.prod(self.xmesh))

    @property
    def nalpha(self):
        return 4 * (2 * self.lc + 1)


# ----------------------------------------------------------------------------
# The action
# ----------------------------------------------------------------------------
def lr_act(P: Problem, psi):
    """Screened direct (long-range) term, vectorised over x.

    long_range.f90:758-870 (lr_act_traditional_x) and :873-963 (lr_kernel):
      phi_k = sum_n u(n,k,x) psi(n,k,a)            (:897-906, no conjugation)
      phi'  = (1/N_k) F_- [ W(:,x) . F_+ phi ]     (:928-939)
      hpsi(n,k,a) -= conj(u(n,k,x)) phi'_k         (:943-961)
    """
    vs = val_spin_of_alpha(P.lc, P.nspn)
    hp = np.zeros(psi.shape, np.complex128, order="F")
    nk = P.nk
    for a in range(psi.shape[2]):
        us = P.u[:, :, :, vs[a]]
        phi = np.einsum("nkx,nk->kx", us, psi[:, :, a], optimize=True)
        t = fft_backward(phi, P.kmesh, axis=0)
        t = t * P.W
        t = fft_forward(t, P.kmesh, axis=0)
        t = t * (1.0 / nk)
        hp[:, :, a] -= np.einsum("nkx,kx->nk", np.conj(us), t, optimize=True)
    return hp


def lr_kernel_loops(P: Problem, psi):
    """Loop-literal restatement of lr_kernel for tiny sizes (pins lr_act)."""
    vs = val_spin_of_alpha(P.lc, P.nspn)
    nb, nk, na = psi.shape
    hpr = np.zeros((nb, nk, na))
    hpi = np.zeros((nb, nk, na))
    for a in range(na):
        for x in range(P.nx):
            ur = P.u[:, :, x, vs[a]].real
            ui = P.u[:, :, x, vs[a]].imag
            xr = np.zeros(nk)
            xi = np.zeros(nk)
            for k in range(nk):
                xr[k] = np.dot(psi[:, k, a].real, ur[:, k]) - np.dot(psi[:, k, a].imag, ui[:, k])
                xi[k] = np.dot(psi[:, k, a].real, ui[:, k]) + np.dot(psi[:, k, a].imag, ur[:, k])
            s = xr + 1j * xi
            # explicit DFTs (no library FFT) with the FFTW sign conventions
            F = dft_matrix_backward(P.kmesh)
            s = F @ s
            s = s * P.W[:, x]
            s = np.conj(F) @ s
            xr = s.real / nk
            xi = s.imag / nk
            for k in range(nk):
                hpr[:, k, a] += -ur[:, k] * xr[k] - ui[:, k] * xi[k]
                hpi[:, k, a] += -ur[:, k] * xi[k] + ui[:, k] * xr[k]
    return hpr + 1j * hpi


# ----------------------------------------------------------------------------
# Orbital-basis-function (OBF) form of the long-range term
# ----------------------------------------------------------------------------
def obf_phase(xmesh, kmesh, tau):
    """OCEAN_obf_make_phase (OCEAN_obf.f90:267-395).

    xshift(d) = floor(X tau_d) for an even K_d and floor(X (tau_d - 1/2)) for an odd
    one, X = xmesh(d), except that for an odd kmesh(1) the reference takes
    xmesh(2) (:300-301, kept as written); tau_c = tau - xshift / xmesh (:316-318);
    the x offsets wrap by -+X_d when i_d - xshift_d leaves 1..X_d (:321-358);
    obf_phs(q, x) = exp(2 pi i sum_d (wrap_d + i_d - 1)(q_d - 1) / (X_d K_d))
    (:360-372).  Returns (phs complex (nk, nx), xoff (nx, 3), tau_c, xshift).
    """
    xm = tuple(int(v) for v in xmesh)
    km = tuple(int(v) for v in kmesh)
    tau = np.array(tau, np.float64)
    xshift = np.zeros(3, np.int64)
    for d in range(3):
        if km[d] % 2 == 0:
            xshift[d] = math.floor(float(xm[d]) * tau[d])
        else:
            X = xm[1] if d == 0 else xm[d]
            xshift[d] = math.floor(float(X) * (tau[d] - 0.5))
    tau_c = tau - xshift / np.array(xm, np.float64)
    nx = xm[0] * xm[1] * xm[2]
    nk = km[0] * km[1] * km[2]
    xoff = np.zeros((nx, 3), np.int64)
    it = 0
    for iz in range(1, xm[2] + 1):
        for iy in range(1, xm[1] + 1):
            for ix in range(1, xm[0] + 1):
                for d, i in enumerate((ix, iy, iz)):
                    targ = i - xshift[d]
                    ph = -xm[d] if targ > xm[d] else (xm[d] if targ < 1 else 0)
                    xoff[it, d] = ph + i - 1
                it += 1
    q = np.array([(a, b, c) for a in range(km[0]) for b in range(km[1]) for c in range(km[2])],
                 np.int64)
    ang = np.zeros((nk, nx))
    for d in range(3):
        ang = ang + 2.0 * np.pi * np.outer(q[:, d], xoff[:, d]).astype(np.float64) / float(xm[d] * km[d])
    phs = np.asfortranarray(np.cos(ang) + 1j * np.sin(ang))
    return phs, xoff, tau_c, xshift


def obf_lrload_bloch(obf, obf2u):
    """Loop-literal OCEAN_obf_lrLOAD before the phase (OCEAN_obf.f90:86-118): the
    band sums and their norm accumulator su, updated inside the o loop (:88-97),
    then u = sum / sqrt(su) (:110-118).  Tiny sizes only.  Returns (nb, nk, nx)."""
    nb, no, nk = obf2u.shape
    nx = obf.shape[1]
    rbs = np.zeros((nb, nk, nx))
    ibs = np.zeros((nb, nk, nx))
    su = np.zeros((nb, nk))
    for x in range(nx):
        for k in range(nk):
            for o in range(no):
                rbs[:, k, x] += obf2u[:, o, k].real * obf[o, x].real - obf2u[:, o, k].imag * obf[o, x].imag
                ibs[:, k, x] += obf2u[:, o, k].imag * obf[o, x].real + obf2u[:, o, k].real * obf[o, x].imag
                for n in range(nb):
                    su[n, k] += rbs[n, k, x] ** 2 + ibs[n, k, x] ** 2
    su = 1.0 / np.sqrt(su)
    return (rbs + 1j * ibs) * su[:, :, None]


def lr_act_obf2(P: Problem, psi, obf, obf2u, phs):
    """Long-range term in the OBF basis.  long_range.f90:262-455 (lr_act_obf2).

    obf (num_obf, nx) = re/im_obf, obf2u (nb, num_obf, nk) = re/im_obf2u, phs (nk, nx):
      beta(o,k,a)  = sum_n obf2u(n,o,k) psi(n,k,a)            (:357-360, DDOT pairs)
      phi(k,x)     = sum_o beta(o,k,a) obf(o,x)                (:373-382, DGEMM 'T','N')
      X            = cfft(-1)[ W(:,x) cfft(+1)[ phi phs ] ]    (:388-400), unnormalised;
                     cfft(+1) is exp(+2 pi i ...), Modules/jlmfft.f:247-262
      beta2(o,k)   = sum_x conj(obf(o,x)) X(k,x) conj(phs(k,x))  (:403-421)
      hp(n,k,a)    = sum_o conj(obf2u(n,o,k)) beta2(o,k)       (:445-452)
    hp is zeroed before it is accumulated (:445-446): the result replaces, rather
    than adds to, whatever the caller's vector held.  Spin-independent (nspn = 1).
    """
    nb, nk, na = psi.shape
    hp = np.zeros(psi.shape, np.complex128, order="F")
    beta = np.einsum("nok,nka->oka", obf2u, psi)
    for a in range(na):
        phi = beta[:, :, a].T @ obf
        t = fft_backward(phi * phs, P.kmesh, axis=0)
        t = t * P.W
        t = fft_forward(t, P.kmesh, axis=0)
        beta2 = np.conj(obf) @ (t * np.conj(phs)).T
        hp[:, :, a] = np.einsum("nok,ok->nk", np.conj(obf2u), beta2)
    return hp


def xact_obf(P: Problem, psi, obf, obf2u, phs, backwards=False, hflag=(1, 0, 0, 0, 1, 1),
             inter=None):
    """H.psi with bloch_control's use_obf (long_range.f90:81-82 selects lr_act_obf2).

    As xact, except that the OBF long-range term replaces the interaction vector
    (lr_act_obf2 zeroes hp, long_range.f90:445-446), so the multiplet term that
    OCEAN_xact accumulated into it first (OCEAN_action.f90:344-351) is dropped
    whenever the long-range term runs; the spin-orbit term (:368-372) and the
    energies (:596-598) are added after it as usual.
    """
    if inter is None:
        inter = P.inter_scale
    psi = np.asarray(psi, np.complex128)
    psi_o = P.allow * psi
    h = np.zeros(psi.shape, np.complex128, order="F")
    if P.do_mult and P.mult is not None and hflag[4] == 1:
        h += fg_combo(P, psi_o, inter)
    if P.long_range and hflag[0] == 1:
        h = lr_act_obf2(P, psi_o, obf, obf2u, phs)
    h = P.allow * h
    if hflag[5] == 1 and P.do_mult and P.mult is not None:
        h += new_soact(P, psi, backwards)
    out = energies_act(P, psi, backwards) if P.e0 else np.zeros_like(h)
    return np.asfortranarray(out + h)


def fg_combo(P: Problem, psi, inter):
    """Core-site multiplet: projection, direct -M_l a, exchange X_l a, back-projection.

    OCEAN_multiplet.f90:1617-2005.  amp = sum_{n,k} c psi (no conjugation,
    :1746-1753); direct h = -mpm a (:1819-1832); exchange over the (nu, m, alpha)
    block with mh(jj, ii) (:1867-1936); both x mul = s/(N_k Omega) (:1641);
    back-projection hpsi += conj(c) h (:1943-1998).
    """
    M = P.mult
    mul = inter / (P.nk * P.celvol)
    nb, nk, na = psi.shape
    nspn = P.nspn
    out = np.zeros(psi.shape, np.complex128, order="F")
    lmax = M.lmax
    # amp[(l, m, nu), a]
    amp = {}
    for li, l in enumerate(range(M.lmin, M.lmax + 1)):
        for m in range(-l, l + 1):
            for nu in range(M.nproj[li]):
                for a in range(na):
                    ispn = (0 if (a % 2 == 0) else 1) if nspn == 2 else 0
                    cc = M.c[:, :, nu, m + lmax, li, ispn]
                    amp[(l, m, nu, a)] = np.sum(psi[:, :, a] * cc)
    hamp = {}
    for li, l in enumerate(range(M.lmin, M.lmax + 1)):
        npl = M.nproj[li]
        for m in range(-l, l + 1):
            for a in range(na):
                for mu in range(npl):
                    s = 0j
                    for nu in range(npl):
                        s -= M.mpm[mu, nu, li] * amp[(l, m, nu, a)]
                    hamp[(l, m, mu, a)] = s * mul
    # exchange (requires lvl <= lmin and lvh >= lmax as the reference loop does)
    for li, l in enumerate(range(M.lmin, M.lmax + 1)):
        npl = M.nproj[li]
        Xl = M.X[l - M.lvl]
        mh = Xl.shape[0]
        pw = np.zeros(mh, np.complex128)
        for a in range(na):
            for m in range(-l, l + 1):
                for nu in range(npl):
                    ii = nu + (m + l) * npl + a * (2 * l + 1) * npl
                    pw[ii] = amp[(l, m, nu, a)]
        hpw = (Xl.T @ pw) * mul  # hpw(ii) = sum_jj mh(jj, ii) pw(jj)
        for a in range(na):
            for m in range(-l, l + 1):
                for nu in range(npl):
                    ii = nu + (m + l) * npl + a * (2 * l + 1) * npl
                    hamp[(l, m, nu, a)] += hpw[ii]
    for a in range(na):
        ispn = (0 if (a % 2 == 0) else 1) if nspn == 2 else 0
        acc = np.zeros((nb, nk), np.complex128)
        for li, l in enumerate(range(M.lmin, M.lmax + 1)):
            for m in range(-l, l + 1):
                for nu in range(M.nproj[li]):
                    cc = M.c[:, :, nu, m + lmax, li, ispn]
                    acc += np.conj(cc) * hamp[(l, m, nu, a)]
        out[:, :, a] += acc
    return out


def new_soact(P: Problem, psi, backwards=False):
    """hpsi(:,:,ic) += sum_jc (Sr +/- i Si)(ic, jc) psi(:,:,jc).

    OCEAN_multiplet.f90:1543-1612; backwards conjugates Si.
    """
    M = P.mult
    out = np.zeros(psi.shape, np.complex128, order="F")
    if not M.have_spin_orbit:
        return out
    S = M.somelr - 1j * M.someli if backwards else M.somelr + 1j * M.someli
    out[...] = np.einsum("ij,nkj->nki", S, psi)
    return out


def energies_act(P: Problem, psi, backwards=False):
    """hpsi = (E + i Sigma'') psi; conjugate if backwards.

    OCEAN_energies.f90:890-906 via OCEAN_psi_f2m_3element_mult (OCEAN_psi.f90:3213-3266).
    """
    e = np.conj(P.energy) if backwards else P.energy
    return e * psi


def xact(P: Problem, psi, backwards=False, hflag=(1, 0, 0, 0, 1, 1), inter=None):
    """One H.psi, core branch, in the reference's order.

    OCEAN_action.f90:286-618: psi_o = A.psi (:337-340); fg with SO off (:344-351);
    lr if hflag(1) (:353-359); allow on the result (:361); SO on the unmasked psi
    (:368-372); energies into min storage (:596-598) plus the interaction part (:608).
    """
    if inter is None:
        inter = P.inter_scale
    psi = np.asarray(psi, np.complex128)
    psi_o = P.allow * psi
    h = np.zeros(psi.shape, np.complex128, order="F")
    if P.do_mult and P.mult is not None and hflag[4] == 1:
        h += fg_combo(P, psi_o, inter)
    if P.long_range and hflag[0] == 1:
        h += lr_act(P, psi_o)
    h = P.allow * h
    if hflag[5] == 1 and P.do_mult and P.mult is not None:
        h += new_soact(P, psi, backwards)
    if P.e0:
        out = energies_act(P, psi, backwards)
    else:
        out = np.zeros_like(h)
    return np.asfortranarray(out + h)


def dense_H(P: Problem, **kw):
    """Brute-force dense matrix of xact (the OCEAN_exact idea; tiny sizes only)."""
    shape = (P.nb, P.nk, P.nalpha)
    N = int(np.prod(shape))
    H = np.zeros((N, N), np.complex128)
    for j in range(N):
        e = np.zeros(N, np.complex128)
        e[j] = 1.0
        H[:, j] = xact(P, e.reshape(shape, order="F"), **kw).ravel(order="F")
    return H


# ----------------------------------------------------------------------------
# Vector layer (OCEAN_psi.f90)
# ----------------------------------------------------------------------------
def psi_dot(p, q):
    """(rval, ival) = (sum p_r q_r + p_i q_i, sum p_r q_i - p_i q_r).  OCEAN_psi.f90:1930-2126."""
    d = np.vdot(p.ravel(order="F"), q.ravel(order="F"))
    return d.real, d.imag


def psi_axmz(x, z, rval=1.0, ival=0.0):
    """y = (rval + i ival) x - z in the reference's operation order.

    OCEAN_psi.f90:1571-1670 (axmz: y = r x - z - i x_i, r x_i - z_i + i x_r) and
    :1468-1562 (axmy: the buffer (r x - y) then -/+ i x, the same arithmetic with z = y).
    """
    xr, xi, zr, zi = x.real, x.imag, z.real, z.imag
    yr = (rval * xr - zr) - ival * xi
    yi = (rval * xi - zi) + ival * xr
    return np.asfortranarray(yr + 1j * yi)


def psi_divide(x, rval=None, ival=None):
    """x / (rval + i ival) with the branches of OCEAN_psi_divide (OCEAN_psi.f90:1764-1874)."""
    xr, xi = x.real, x.imag
    if rval is not None and ival is not None:
        den = rval * rval + ival * ival
        rn, inn = rval / den, ival / den
        outr = rn * xr + inn * xi
        outi = -inn * xr + rn * xi
    elif rval is not None:
        rn = 1.0 / rval
        outr, outi = rn * xr, rn * xi
    else:
        inn = 1.0 / ival
        outr, outi = xi * inn, -inn * xr
    return np.asfortranarray(outr + 1j * outi)


def psi_nrm(p):
    """Squared norm.  OCEAN_psi.f90:1673-1760."""
    return float(np.sum(p.real ** 2 + p.imag ** 2))


def core_pnorm(psi, nk, celvol):
    """Normalise psi0 and return kpref = 4 pi |psi|^2 / (N_k Omega^2).  OCEAN_psi.f90:5589-5638."""
    val = 0.0
    for a in range(psi.shape[2]):
        val += np.sum(psi[:, :, a].real ** 2 + psi[:, :, a].imag ** 2)
    val = math.sqrt(val)
    kpref = 4.0 * math.pi * val ** 2 / (nk * celvol ** 2)
    return np.asfortranarray(psi * (1.0 / val)), kpref


def energies_core_xas(E, efermi, shift, lc, nspn):
    """p_energy and allow of a core XAS run with the default occupation.

    OCEAN_energies.f90:303-368 (load), :421-487 (core_allow_fixed without a
    temperature: allow 1 for E >= E_F, 0 below), :829-887 (stubby: alpha takes the
    band energies of its valence spin, shifted by the core-level shift).  E is
    (nb, nk, nspn) in Hartree.  Returns (energy complex (nb, nk, na), allow).
    """
    nb, nk = E.shape[0], E.shape[1]
    na = 4 * (2 * lc + 1)
    energy = np.zeros((nb, nk, na), np.complex128, order="F")
    allow = np.zeros((nb, nk, na), order="F")
    ialpha = 0
    for _icms in range(2):
        for _icml in range(-lc, lc + 1):
            for ivms in range(2):
                vs = 0 if nspn == 1 else ivms
                for k in range(nk):
                    for n in range(nb):
                        energy[n, k, ialpha] = E[n, k, vs] + shift
                        allow[n, k, ialpha] = 1.0 if E[n, k, vs] >= efermi else 0.0
                ialpha += 1
    return energy, allow


def psi_dotter(mels, cks, lc, nb, nk, nspn, bands_in_file=None):
    """psi0(n,k,a) from core->projector dipole elements and projector overlaps.

    OCEAN_psi.f90:5887-6031.  mels: complex (nptot, 2lc+1) [p, m_c]; cks: complex
    (nptot, ntot, nspn) with ntot = bands_in_file * nk.  Nonzero only where core
    spin equals valence spin (:5985); value = conj-combination (:5990-5995).
    """
    if bands_in_file is None:
        bands_in_file = nb
    na = 4 * (2 * lc + 1)
    psi = np.zeros((nb, nk, na), np.complex128, order="F")
    mer, mei = mels.real, mels.imag
    ia = 0
    for icms in (-1, 1):
        for icml in range(-lc, lc + 1):
            for ivms in (-1, 1):
                if icms == ivms:
                    sp = 0 if nspn == 1 else (0 if ivms == -1 else 1)
                    pcr = cks[:, :, sp].real
                    pci = cks[:, :, sp].imag
                    it = 0
                    for k in range(nk):
                        sl = slice(it, it + nb)
                        rr = mer[:, icml + lc] @ pcr[:, sl]
                        ri = mer[:, icml + lc] @ pci[:, sl]
                        ir = mei[:, icml + lc] @ pcr[:, sl]
                        ii = mei[:, icml + lc] @ pci[:, sl]
                        psi[:, k, ia] = (rr - ii) + 1j * (-ri - ir)
                        it += bands_in_file
                ia += 1
    return psi


# ----------------------------------------------------------------------------
# Haydock
# ----------------------------------------------------------------------------
def haydock(P: Problem, psi0, niter, act=None):
    """Hermitian two-term Haydock recursion.

    OCEAN_haydock.f90:190-320 (loop) and :824-938 (OCEAN_hay_ab_twoterm):
    h -= b(i-1) psi_old; a(i-1) = Re<h|psi>; h -= a(i-1) psi; b(i) = sqrt(<h|h>);
    psi_old <- psi; psi <- h / b(i).  b(0) = 0, psi_old = 0 (:2223-2239).
    Returns (real_a, real_b, imag_a, ierr); ierr = -1 on NaN in a (:912-920).
    """
    if act is None:
        act = lambda v: xact(P, v)
    psi = np.asfortranarray(np.array(psi0, np.complex128))
    old = np.zeros_like(psi)
    ra = np.zeros(niter + 1)
    rb = np.zeros(niter + 1)
    ia = np.zeros(niter + 1)
    for it in range(1, niter + 1):
        h = act(psi)
        h = h + (-rb[it - 1]) * old
        ar, ai = psi_dot(h, psi)
        ra[it - 1] = ar
        h = h + (-ar) * psi
        bb = psi_nrm(h)
        old = psi
        rb[it] = math.sqrt(bb)
        psi = h * (1.0 / rb[it])
        ia[it - 1] = ai
        if ar != ar:
            return ra, rb, ia, -1
    return ra, rb, ia, 0


def psi_axpy(rval, x, y, ival=0.0):
    """y + (rval + i ival) x in the reference's order: the real part of the scalar
    first, then the imaginary one.  OCEAN_psi.f90:1386-1458 (OCEAN_psi_axpy)."""
    yr = y.real + rval * x.real
    yi = y.imag + rval * x.imag
    yr = yr + (-ival) * x.imag
    yi = yi + ival * x.real
    return np.asfortranarray(yr + 1j * yi)


def haydock_abc(P: Problem, psi0, niter, vectors=False):
    """Non-Hermitian (bi-Lanczos) Haydock recursion.

    Driver OCEAN_haydock.f90:60-160 (OCEAN_haydock_nonHerm_do): psi = back_psi = psi0,
    old vectors zero; per iteration the allow mask on psi and back_psi, H psi and the
    Hermitian conjugate's action on back_psi (OCEAN_xact with backwards), then
    haydock_abc_1 (:1250-1433):
      mask h, bh; alpha = (bpsi, h) [= <bpsi|h>];
      h  += -b(i-1) old;          bh += -conj(c(i-1)) bold;
      h  += -alpha psi;           bh += -conj(alpha) bpsi;
      mask h, bh; r = <bh|h>;
      c(i) = sqrt((|r| + Re r)/2) + i sign(sqrt((|r| - Re r)/2), Im r);
      b(i) = conj(r / c(i));
      bh /= conj(b(i)); h /= c(i); old <- psi, psi <- h (and the back set).
    b(0) = c(0) = 0.  Returns (ra, ia, rb, ib, rc, ic, ierr); ierr = -1 when a is NaN
    (:1414-1422).  With vectors, also the lists of masked psi and back_psi of
    iterations 1 .. niter + 1 (for the three-term relations of the tests).
    """
    n1 = niter + 1
    ra, ia, rb, ib, rc, ic = (np.zeros(n1) for _ in range(6))
    psi = np.asfortranarray(np.array(psi0, np.complex128))
    bpsi = psi.copy(order="F")
    old = np.zeros_like(psi)
    bold = np.zeros_like(psi)
    V, W = [], []
    for it in range(1, n1):
        psi = P.allow * psi
        bpsi = P.allow * bpsi
        V.append(psi)
        W.append(bpsi)
        h = P.allow * xact(P, psi)
        bh = P.allow * xact(P, bpsi, backwards=True)
        ar, ai = psi_dot(bpsi, h)
        h = psi_axpy(-rb[it - 1], old, h, -ib[it - 1])
        bh = psi_axpy(-rc[it - 1], bold, bh, ic[it - 1])
        h = psi_axpy(-ar, psi, h, -ai)
        bh = psi_axpy(-ar, bpsi, bh, ai)
        ra[it - 1], ia[it - 1] = ar, ai
        h = P.allow * h
        bh = P.allow * bh
        rr, ri = psi_dot(bh, h)
        old, bold = psi, bpsi
        mag = math.sqrt(rr ** 2 + ri ** 2)
        rc[it] = math.sqrt((mag + rr) / 2.0)
        ic[it] = math.copysign(math.sqrt((mag - rr) / 2.0), ri)
        t = complex(rr, ri) / complex(rc[it], ic[it])
        rb[it], ib[it] = t.real, -t.imag
        bpsi = psi_divide(bh, rb[it], -ib[it])
        psi = psi_divide(h, rc[it], ic[it])
        if ar != ar:
            return (ra, ia, rb, ib, rc, ic, -1) + ((V, W) if vectors else ())
    V.append(P.allow * psi)
    W.append(P.allow * bpsi)
    return (ra, ia, rb, ib, rc, ic, 0) + ((V, W) if vectors else ())


def write_core(a, b, niter, kpref, gam0, el, eh, ne, imag_a=None, literal_cmplx=True,
               hartree2ev=HARTREE2EV, ebase=0.0):
    """Continued fraction with square-root terminator (Lorentzian, gamGauss off).

    OCEAN_haydock.f90:1801-1894.  With literal_cmplx the three kind-less
    ``cmplx(...)`` sites (:1846-1847, :1870-1871) are rounded to default
    (single) precision, as OCEAN2/Makefile builds them without default-real
    promotion.  Returns columns (E_eV, Spect, Spect(0), SPKK, SPKK(0)).
    """
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ia = np.zeros_like(a) if imag_a is None else np.asarray(imag_a, np.float64)
    n = niter
    ie = np.arange(0, 2 * ne + 1, 2, dtype=np.float64)
    e = el + (eh - el) * ie / (2 * ne)
    gam = gam0
    f32 = (lambda v: np.float64(np.float32(v))) if literal_cmplx else (lambda v: v)

    def bc(j):  # cmplx(b) * cmplx(c), c = b, imag parts 0
        if literal_cmplx:
            return np.float64(np.float32(np.float32(b[j]) * np.float32(b[j])))
        return b[j] * b[j]

    ctmp = (e - a[n - 1]) + 1j * (gam + ia[n - 1])
    if literal_cmplx:
        four_bc = np.float64(np.float32(np.float32(4.0 * np.float32(b[n])) * np.float32(b[n])))
    else:
        four_bc = 4.0 * b[n] * b[n]
    disc = np.sqrt(ctmp ** 2 - four_bc)
    delta = np.where(disc.imag > 0.0, (ctmp + disc) / 2.0, (ctmp - disc) / 2.0)
    for jj in range(n - 1, -1, -1):
        t = bc(jj + 1)
        if literal_cmplx:
            z = np.float32(e - a[jj]).astype(np.float64) + 1j * np.float64(np.float32(gam + ia[jj]))
        else:
            z = (e - a[jj]) + 1j * (gam + ia[jj])
        delta = z - t / delta
    dr = delta.real
    di = np.abs(delta.imag)
    spct = kpref * di / (dr ** 2 + di ** 2)
    spkk = kpref * dr / (dr ** 2 + di ** 2)
    ener = ebase + hartree2ev * e
    del f32
    return np.stack([ener, spct, spct, spkk, spkk], axis=1)


def write_core_abc(ra, ia, rb, ib, rc, ic, niter, kpref, gam0, el, eh, ne, literal_cmplx=True,
                   hartree2ev=HARTREE2EV):
    """write_core's continued fraction with complex b and c (the bi-Lanczos
    coefficients of haydock_abc), OCEAN_haydock.f90:1836-1880:
      ctmp = cmplx(e - a(n-1), gam + ia(n-1), DP);
      disc = sqrt(ctmp^2 - 4 cmplx(b(n)) cmplx(c(n))), delta = (ctmp +/- disc)/2 by the
        sign of Im disc;
      delta = cmplx(e - a(jj), gam + ia(jj)) - cmplx(b(jj+1)) cmplx(c(jj+1)) / delta.
    With literal_cmplx the kind-less cmplx() operands and their products are single
    precision (complex(4) arithmetic, every product and sum rounded on its own); the
    division by delta and the rest are FP64.  Returns (E_eV, Spect, Spect(0), SPKK,
    SPKK(0)) columns."""
    ra, ia, rb, ib, rc, ic = (np.asarray(v, np.float64) for v in (ra, ia, rb, ib, rc, ic))
    n = niter
    e = el + (eh - el) * np.arange(0, 2 * ne + 1, 2, dtype=np.float64) / (2 * ne)
    gam = gam0
    f = np.float32

    def bc(j, scale):
        if literal_cmplx:
            x, y = f(scale) * f(rb[j]), f(scale) * f(ib[j])
            u, v = f(rc[j]), f(ic[j])
            return complex(np.float64(f(x * u) - f(y * v)), np.float64(f(x * v) + f(y * u)))
        return scale * complex(rb[j], ib[j]) * complex(rc[j], ic[j])

    ctmp = (e - ra[n - 1]) + 1j * (gam + ia[n - 1])
    disc = np.sqrt(ctmp ** 2 - bc(n, 4.0))
    delta = np.where(disc.imag > 0.0, (ctmp + disc) / 2.0, (ctmp - disc) / 2.0)
    for jj in range(n - 1, -1, -1):
        t = bc(jj + 1, 1.0)
        if literal_cmplx:
            z = f(e - ra[jj]).astype(np.float64) + 1j * np.float64(f(gam + ia[jj]))
        else:
            z = (e - ra[jj]) + 1j * (gam + ia[jj])
        delta = z - t / delta
    dr = delta.real
    di = np.abs(delta.imag)
    spct = kpref * di / (dr ** 2 + di ** 2)
    spkk = kpref * dr / (dr ** 2 + di ** 2)
    return np.stack([hartree2ev * e, spct, spct, spkk, spkk], axis=1)


# OCEAN_constants.f90:17-29 (2022 CODATA)
ALPHA = 0.0072973525643
BOHR = 0.529177210544
AU2SEC = 2.4188843265864e-17
PI_DP = 3.1415926535897932384626433832795028841971693993751


def gamfcn(e, n, e0):
    """Plasmon-pole lifetime broadening of an excited electron at energy e (Hartree)
    in a gas of density n (bohr^-3) with dielectric constant e0.

    CNBSE/zcmsnmult/zf90/gamfcn.f90:8-42 (q stepped by kf/200 until the plasmon
    threshold test fails).
    """
    pi = 3.14159265358979
    kf = (3.0 * pi * pi * n) ** (1.0 / 3.0)
    wp = math.sqrt(4.0 * pi * n)
    k = kf ** 2 / 2.0 + e
    if k < kf ** 2 / 2.0:
        return 0.0
    k = math.sqrt(2.0 * k)
    wf = kf ** 2 / 2.0
    lam = math.sqrt(wp ** 2 / (wf ** 2 * (e0 - 1.0)))
    dq = 0.005 * kf
    q = 0.5 * dq
    su = 0.0
    test1 = True
    while test1:
        qq = q / kf
        xp = qq * (2.0 + qq) / lam
        xm = qq * (2.0 - qq) / lam
        term1 = lam / (2.0 * qq ** 3) * (math.atan(xp) + math.atan(xm))
        term2 = (0.125 * lam ** 2 / qq ** 5 + 0.5 / qq ** 3 - 0.125 / qq) * math.log((1.0 + xp ** 2) / (1.0 + xm ** 2))
        epsq = 1.0 + 2.0 / (pi * kf) * (1.0 / qq ** 2 - term1 + term2)
        wq = wp / math.sqrt(1.0 - 1.0 / epsq)
        test1 = k ** 2 - kf ** 2 - 2.0 * wq > 0.0
        test2 = k * q - wq - q ** 2 / 2.0 > 0.0
        if test1 and test2:
            su = su + dq / (q * wq)
        q = q + dq
    return wp * wp * su / (2.0 * k)


def _cf_delta(a, b, ia, n, e, gam, literal):
    """delta(e) of write_core's fraction (one energy array, one gamma array)."""
    ctmp = (e - a[n - 1]) + 1j * (gam + ia[n - 1])
    if literal:
        four_bc = np.float64(np.float32(np.float32(4.0 * np.float32(b[n])) * np.float32(b[n])))
    else:
        four_bc = 4.0 * b[n] * b[n]
    disc = np.sqrt(ctmp ** 2 - four_bc)
    delta = np.where(disc.imag > 0.0, (ctmp + disc) / 2.0, (ctmp - disc) / 2.0)
    for jj in range(n - 1, -1, -1):
        if literal:
            t = np.float64(np.float32(np.float32(b[jj + 1]) * np.float32(b[jj + 1])))
            z = np.float32(e - a[jj]).astype(np.float64) + 1j * np.float32(gam + ia[jj]).astype(np.float64)
        else:
            t = b[jj + 1] * b[jj + 1]
            z = (e - a[jj]) + 1j * (gam + ia[jj])
        delta = z - t / delta
    return delta


def write_core_full(a, b, niter, kpref, gam0, el, eh, ne, ebase_eV=0.0, old_broaden=False, nval=0.0,
                    eps=1.0, ucvol=1.0, gam_gauss=0.0, imag_a=None, literal_cmplx=True,
                    hartree2ev=HARTREE2EV):
    """write_core as OCEAN 3.2.1 writes it (OCEAN_haydock.f90:1801-1894): the nine
    columns (E, Spect, Spect(0), SPKK, SPKK(0), sigma a.u.^2, sigma cm^2, two-photon
    a.u.^5 and cm^4 s), Spect/SPKK with gamfcn broadening when oldXASbroaden, and the
    Voigt variant write_core_Voigt (:1896-1989) when gamGauss > gam0/10 (five
    columns, its fine-grid reuse and printed-energy convention included).
    """
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ia = np.zeros_like(a) if imag_a is None else np.asarray(imag_a, np.float64)
    n = niter
    if gam_gauss > gam0 / 10.0:
        return _write_core_voigt(a, b, ia, n, kpref, gam0, el, eh, ne, ebase_eV, nval, eps, gam_gauss,
                                 literal_cmplx, hartree2ev)
    e = el + (eh - el) * np.arange(0, 2 * ne + 1, 2, dtype=np.float64) / (2 * ne)
    d0 = _cf_delta(a, b, ia, n, e, np.full_like(e, gam0), literal_cmplx)
    spct0 = kpref * np.abs(d0.imag) / (d0.real ** 2 + d0.imag ** 2)
    spkk0 = kpref * d0.real / (d0.real ** 2 + d0.imag ** 2)
    if old_broaden:
        g1 = gam0 + np.array([gamfcn(v, nval, eps) for v in e])
        d1 = _cf_delta(a, b, ia, n, e, g1, literal_cmplx)
        spct1 = kpref * np.abs(d1.imag) / (d1.real ** 2 + d1.imag ** 2)
        spkk1 = kpref * d1.real / (d1.real ** 2 + d1.imag ** 2)
    else:
        spct1, spkk1 = spct0, spkk0
    ener = ebase_eV + hartree2ev * e
    a_by_v = EV2HARTREE * ALPHA * ucvol
    bohr_sq = BOHR * BOHR * 1e-16
    tpa = (ALPHA * ALPHA * 2.0 * PI_DP) * ucvol
    tpacm4 = tpa * bohr_sq * bohr_sq * AU2SEC
    sigma = a_by_v * ener * spct0
    return np.stack([ener, spct1, spct0, spkk1, spkk0, sigma, sigma * bohr_sq, spct0 * tpa, spct0 * tpacm4],
                    axis=1)


def _write_core_voigt(a, b, ia, n, kpref, gam0, el, eh, ne, ebase_eV, nval, eps, gam_gauss, literal,
                      hartree2ev=HARTREE2EV):
    """write_core_Voigt line by line: the fine grid of step estep around each output
    energy, reused (shifted by nratio) from one output point to the next."""
    de = (eh - el) / float(ne)
    estep = min(gam0, gam_gauss) / 2.0
    nratio = int((de - estep) / estep) + 1
    estep = de / float(nratio)
    nstep = int((2.99999 * gam_gauss) / estep) + 1
    prefactor = estep / (gam_gauss * 2.506628274631)
    inv2s2 = 1.0 / (2.0 * gam_gauss * gam_gauss)
    sa = np.zeros((2 * nstep + 1, 2))
    sk = np.zeros((2 * nstep + 1, 2))
    rows = []
    e = 0.0
    for ie in range(0, 2 * ne + 1, 2):
        ee = el + (eh - el) * float(ie) / float(2 * ne)
        if ie == 0:
            istart = -nstep
        else:
            istart = nstep - nratio
            sa[0:2 * nstep + 1 - nratio] = sa[nratio:2 * nstep + 1].copy()
            sk[0:2 * nstep + 1 - nratio] = sk[nratio:2 * nstep + 1].copy()
        for ii in range(istart, nstep + 1):
            e = ee + float(ii) * estep
            for jd in (0, 1):
                g = gam0 + gamfcn(e, nval, eps) * jd
                d = _cf_delta(a, b, ia, n, np.array([e]), np.array([g]), literal)[0]
                dr, di = d.real, abs(d.imag)
                sa[ii + nstep, jd] = prefactor * kpref * di / (dr ** 2 + di ** 2)
                sk[ii + nstep, jd] = prefactor * kpref * dr / (dr ** 2 + di ** 2)
        s = np.zeros(2)
        k = np.zeros(2)
        for ii in range(-nstep, nstep + 1):
            pref = math.exp(-float(ii ** 2) * estep ** 2 * inv2s2)
            s += pref * sa[ii + nstep]
            k += pref * sk[ii + nstep]
        rows.append([ebase_eV + hartree2ev * e, s[1], s[0], k[1], k[0]])
    return np.array(rows)


ALPHAINV = 137.035999177  # OCEAN_constants.f90:19


def o_spect_core(a, b, it, kpref, ne, el, eh, gam0, ebase, nval, eps0):
    """o_spect's core branch (OCEAN2/o_spect.f90:228-266) for one abslanc run: el, eh,
    gam0, ebase in Hartree (spect.in values times eV2Hartree, :30-33); it levels.

    Full FP64 complex arithmetic (rm1 is complex(dp)); with ebase > 0.01 Hartree the
    anti-resonant fraction mdelta (poles at -(a + 2 ebase)) is subtracted; both SPKK
    columns print -Re(eps) of the gamfcn-broadened (jdamp = 1) pass, as the whole-array
    assignment at :257 leaves them.  Rows (E_eV, Spect, Spect(0), SPKK, SPKK(0))."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    rows = []
    for ie in range(0, 2 * ne + 1, 2):
        e = el + (eh - el) * float(ie) / float(2 * ne)
        spct = [0.0, 0.0]
        for jd in (0, 1):
            gam = gam0 + gamfcn(e, nval, eps0) * jd
            ct = complex(e - a[it - 1], gam)
            disc = np.sqrt(ct * ct - 4.0 * b[it] ** 2 + 0j)
            delta = (ct + disc) / 2 if disc.imag > 0.0 else (ct - disc) / 2
            md = complex(delta.real - e - 2.0 * ebase, delta.imag - gam)
            for jj in range(it - 1, -1, -1):
                delta = complex(e - a[jj], gam) - b[jj + 1] ** 2 / delta
                md = complex(-e - a[jj] - 2.0 * ebase, -gam) - b[jj + 1] ** 2 / md
            eps = -kpref / delta
            if ebase > 0.01:
                eps = eps - kpref / md
            spct[jd] = eps.imag
        spkk = -eps.real
        rows.append([(e + ebase) * HARTREE2EV, spct[1], spct[0], spkk, spkk])
    return np.array(rows)


def o_spect_val(a, b, it, kpref, ucvol, ne, el, eh, gam0):
    """o_spect's valence branch (OCEAN2/o_spect.f90:166-222): eps(w) = 1 - f/al - f/be
    with f = 2 kpref ucvol, al/be the resonant and anti-resonant fractions at the ne
    midpoints w = el + (eh - el)(2i - 1)/(2 ne), then n, kappa, mu (cm^-1), R and the
    loss function.  Rows (w_eV, eps1, eps2, n, kappa, mu, R, loss)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    fact = kpref * ucvol * 2.0
    rows = []
    for ie in range(1, 2 * ne, 2):
        w = el + (eh - el) * float(ie) / float(2 * ne)
        ct = complex(w, gam0)
        arg = np.sqrt(complex((w - a[it - 1]) ** 2 - 4.0 * b[it] ** 2, 0.0))
        rp = 0.5 * (w - a[it - 1] + arg)
        rm = 0.5 * (w - a[it - 1] - arg)
        rrr = rp if rp.imag < 0.0 else rm
        al = ct - a[it - 1] - rrr
        be = -ct - a[it - 1] - rrr
        for i in range(it - 1, -1, -1):
            al = ct - a[i] - b[i + 1] ** 2 / al
            be = -ct - a[i] - b[i + 1] ** 2 / be
        eps = 1.0 - fact / al - fact / be
        re, im = eps.real, eps.imag
        loss = im / (re ** 2 + im ** 2)
        rad = math.sqrt(re ** 2 + im ** 2)
        th = math.acos(re / rad) / 2
        nref, kap = math.sqrt(rad) * math.cos(th), math.sqrt(rad) * math.sin(th)
        N = np.sqrt(eps)
        r = (N - 1.0) / (N + 1.0)
        refl = (r * np.conj(r)).real
        mu = 2.0 * w * HARTREE2EV * N.imag / (BOHR * ALPHAINV * 1000)
        rows.append([w * HARTREE2EV, re, im, nref, kap, mu, refl, loss])
    return np.array(rows)


def calc_spect_core(a, b, n, kpref, gam0, el, eh, ne, imag_a=None):
    """calc_spect_core (OCEAN_haydock.f90:1991-2027): the n-level fraction at the ne
    midpoints e = el + (eh - el)(2 ie - 1)/(2 ne); rows (E_eV, SPKK, Spect).  Its
    kind-less cmplx sites round like write_core's."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ia = np.zeros_like(a) if imag_a is None else np.asarray(imag_a, np.float64)
    e = el + (eh - el) * (2.0 * np.arange(1, ne + 1) - 1.0) / (2.0 * ne)
    ctmp = (e - a[n - 1]) + 1j * (gam0 + ia[n - 1])
    four_bc = np.float64(np.float32(np.float32(4.0 * np.float32(b[n])) * np.float32(b[n])))
    disc = np.sqrt(ctmp ** 2 - four_bc)
    delta = np.where(disc.imag > 0.0, (ctmp + disc) / 2.0, (ctmp - disc) / 2.0)
    for jj in range(n - 1, -1, -1):
        t = np.float64(np.float32(np.float32(b[jj + 1]) * np.float32(b[jj + 1])))
        z = np.float32(e - a[jj]).astype(np.float64) + 1j * np.float64(np.float32(gam0 + ia[jj]))
        delta = z - t / delta
    dr = delta.real
    di = np.abs(delta.imag)
    return np.stack([HARTREE2EV * e, kpref * dr / (dr ** 2 + di ** 2), kpref * di / (dr ** 2 + di ** 2)])


def haydock_early_exit(a, b, niter, spacing, threshold, kpref, gam0, el, eh, ne, imag_a=None):
    """The stopping iteration of OCEAN_haydock_Herm_do's early exit (OCEAN_haydock.f90:272-291,
    check_convergence :1608-1661) replayed over a finished recursion's coefficients."""
    prev = False
    for it in range(1, niter + 1):
        if it <= spacing:
            continue
        s1 = calc_spect_core(a, b, it, kpref, gam0, el, eh, ne, imag_a)[2]
        s2 = calc_spect_core(a, b, it - spacing, kpref, gam0, el, eh, ne, imag_a)[2]
        rel = 2.0 * np.sum(np.abs(s1 - s2)) / (np.sum(np.abs(s1)) + np.sum(np.abs(s2)))
        if rel < threshold:
            if prev:
                return it
            prev = True
        else:
            prev = False
    return niter


def redtrid_eigs(a, b, n):
    """Eigenvalues of the (n+1) x (n+1) tridiagonal Lanczos matrix.

    OCEAN_haydock.f90:2273-2316 (redtrid, via elsch; Hermitian case).
    """
    T = np.zeros((n + 1, n + 1))
    for i in range(n + 1):
        T[i, i] = a[i]
        if i < n:
            T[i + 1, i] = b[i + 1]
            T[i, i + 1] = b[i + 1]
    return np.linalg.eigvalsh(T)


# ----------------------------------------------------------------------------
# GMRES (recycled, preconditioned)
# ----------------------------------------------------------------------------
def min_set_prec(energy, gprc, h1):
    """pcdiv = (E - (h1_r + i(h1_i + g))) / ((E - h1)^2 + g^2).  OCEAN_psi.f90:2907-2994."""
    ct = (energy - h1) ** 2 + gprc * gprc
    num = energy - (h1.real + 1j * (h1.imag + gprc))
    return np.asfortranarray(num / ct)


def gmres_do_recycle(P: Problem, hay_vec, energies_eV, gres_eV, gprc_eV, ffff, depth,
                     kpref, act=None, nhflag=(1, 0, 0, 0, 1, 1), return_x=False):
    """Recycled, preconditioned GMRES over a list of energies.

    OCEAN_gmres.f90:508-908 (OCEAN_gmres_do_recycle), :241-344 (update_gmres:
    normal-equation least squares on the au basis, zhptrf/zhptrs), :207-239
    (make_recycle), :1213-1283 (set_initial_vector; force_reuse_x = .true.),
    OCEAN_psi.f90:2907-2994 (preconditioner).  Output per energy:
    (E_eV, 1 - Re<b|x> f, -Im<b|x> f, rel_error, complete_iter, global_iter).
    """
    if act is None:
        act = lambda v: xact(P, v, hflag=nhflag)
    gres = gres_eV * EV2HARTREE
    gprc = gprc_eV * EV2HARTREE
    rhs = np.asfortranarray(P.allow * np.asarray(hay_vec, np.complex128))  # sfact_copy
    one = np.ones(rhs.shape, np.complex128, order="F")
    h1 = act(one)
    x = np.zeros_like(rhs)
    ax = np.zeros_like(rhs)
    # u_matrix / au_matrix (rows 1..depth, F-ordered vectors): the Gram column, the
    # right-hand side <au_i|g> and the x / ax updates are matrix-vector products
    shape, nvec = rhs.shape, rhs.size
    U = np.zeros((depth + 1, nvec), np.complex128)
    AU = np.zeros((depth + 1, nvec), np.complex128)
    flat = lambda v: np.ravel(v, order="F")  # noqa: E731
    unflat = lambda v: np.asfortranarray(v.reshape(shape, order="F"))  # noqa: E731
    C = np.zeros((depth, depth), np.complex128)  # upper triangle of the packed Gram matrix
    rows = []
    xs = []
    global_iter = 0
    prev_iter = 0
    prev_energy = -np.inf
    pcdiv = None
    gval = 0.0
    rval = ival = 0.0

    def update(cur, g):
        # C(i, cur) = conj(<au_cur|au_i>) = <au_i|au_cur> for i < cur; C(cur,cur) = |au_cur|^2
        nonlocal x, ax
        if cur > 1:
            C[:cur - 1, cur - 1] = np.conj(AU[1:cur]) @ AU[cur]
        C[cur - 1, cur - 1] = np.vdot(AU[cur], AU[cur]).real
        rvec = np.conj(AU[1:cur + 1]) @ flat(g)
        A = np.triu(C[:cur, :cur])
        A = A + np.conj(np.triu(A, 1)).T
        coeff = np.linalg.solve(A, -rvec)
        x = x + unflat(coeff @ U[1:cur + 1])
        ax = ax + unflat(coeff @ AU[1:cur + 1])

    energies = list(energies_eV)
    for step, E_eV in enumerate(energies, start=1):
        complete_iter = 0
        ener = E_eV * EV2HARTREE
        if prev_iter == 0:
            prev_energy = ener
            pcdiv = min_set_prec(ener, gprc, h1)
        if prev_iter == 0:
            # set_initial_vector, 'keep' for step > 1 (force_reuse_x)
            if step > 1:
                ax = ener * x + 1j * gres * x - act(x)
            else:
                x = np.zeros_like(rhs)
                ax = np.zeros_like(rhs)
            g = ax - rhs
        else:
            delta = ener - energies[step - 2] * EV2HARTREE
            AU[1:prev_iter + 1] += delta * U[1:prev_iter + 1]
            ax = ener * x + 1j * gres * x - act(x)
            g = ax - rhs
            for it in range(1, prev_iter + 1):
                update(it, g)
                g = ax - rhs
            gval = psi_nrm(g)
            rval, ival = psi_dot(rhs, x)
            converged = False
            if gval < ffff:
                rel = -gval / ival
                if rel < 0.001 or gval < ffff / 10.0:
                    converged = True
            if converged:
                rows.append((E_eV, 1.0 - rval * kpref, -ival * kpref, -gval / ival, complete_iter,
                             global_iter))
                xs.append(x.copy())
                continue
            if abs(prev_energy - ener) > gprc / 4.0:
                prev_iter = 0
                prev_energy = ener
                pcdiv = min_set_prec(ener, gprc, h1)
        done = False
        while not done:
            for it in range(prev_iter + 1, depth + 1):
                complete_iter += 1
                global_iter += 1
                pg = g * pcdiv
                U[it] = flat(pg)
                apg = ener * pg + 1j * gres * pg - act(pg)
                AU[it] = flat(apg)
                update(it, g)
                if it == depth:
                    ax = ener * x + 1j * gres * x - act(x)
                    if ener != prev_energy:
                        pcdiv = min_set_prec(ener, gprc, h1)
                        prev_energy = ener
                    prev_iter = 0
                g = ax - rhs
                gval = psi_nrm(g)
                rval, ival = psi_dot(rhs, x)
                if gval < ffff:
                    rel = -gval / ival
                    if rel < 0.001 or gval < ffff / 10.0:
                        prev_iter = it
                        if it == depth:
                            prev_iter = 0
                        done = True
                        break
            else:
                prev_iter = 0
                continue
        rows.append((E_eV, 1.0 - rval * kpref, -ival * kpref, -gval / ival, complete_iter, global_iter))
        xs.append(x.copy())
    out = np.array(rows)
    return (out, xs) if return_x else out


# ----------------------------------------------------------------------------
# Differentiation form (no reference counterpart: OCEAN2 computes no gradients).
# The spectrum is the GMRES one (OCEAN_gmres.f90:761, :864-869):
#   sigma(E) = -kpref Im<b|x_E>,  ((E + i gam) - H) x_E = b,  b = allow . hay
# The restatement solves densely (tiny sizes); its analytic gradient is pinned
# by central finite differences of spectral_sigma in tests/test_oracle.py.
# ----------------------------------------------------------------------------
def spectral_sigma(P: Problem, hay_vec, energies_eV, gam_eV, kpref, H=None):
    """sigma(E) by dense solves of the resolvent system."""
    if H is None:
        H = dense_H(P)
    b = (P.allow * np.asarray(hay_vec, np.complex128)).ravel(order="F")
    N = b.size
    gam = gam_eV * EV2HARTREE
    out = np.zeros(len(energies_eV))
    for i, E_eV in enumerate(energies_eV):
        x = np.linalg.solve((E_eV * EV2HARTREE + 1j * gam) * np.eye(N) - H, b)
        out[i] = -kpref * np.vdot(b, x).imag
    return out


def spectral_grad(P: Problem, hay_vec, energies_eV, gam_eV, kpref, target, wweights=None, nptab=None):
    """Loss sum_E (sigma - target)^2 and its gradient by adjoint (conjugate-energy) solves.

    With M = (E + i gam) - H, x = M^-1 b and y = M^-H b:
      d sigma/d gam   =  kpref Re<y|x>          (dM/dgam = i)
      d sigma/d Delta = -kpref Im<y|x>          (H -> H + Delta on the energy term)
      d sigma/d s     = -kpref Im<y|(H_lr + H_mult) x> / s   (W and mul scale with s)
      d sigma/d W(k,x) = (kpref / N_k) Im sum_a conj(P_y) P_x,  P_v = F_+ U (A v)
    (H_lr = -sum_x A U_x^+ N_k^-1 F_- W_x F_+ U_x A, long_range.f90:897-961).  Gamma and
    Delta derivatives are returned per eV; wweights = (i0, w0, w1) from
    populate_W_weights gives dL/dptab.
    """
    H = dense_H(P)
    import copy
    P2 = copy.copy(P)
    P2.e0 = False
    H_int = dense_H(P2, hflag=(1, 0, 0, 0, 1, 0))  # H_lr + H_mult
    shape = (P.nb, P.nk, P.nalpha)
    b = (P.allow * np.asarray(hay_vec, np.complex128)).ravel(order="F")
    N = b.size
    gam = gam_eV * EV2HARTREE
    vs = val_spin_of_alpha(P.lc, P.nspn)
    sig = np.zeros(len(energies_eV))
    gG = gD = gS = 0.0
    Q = np.zeros((P.nk, P.nx))
    for i, E_eV in enumerate(energies_eV):
        M = (E_eV * EV2HARTREE + 1j * gam) * np.eye(N) - H
        x = np.linalg.solve(M, b)
        y = np.linalg.solve(M.conj().T, b)
        sig[i] = -kpref * np.vdot(b, x).imag
        c = 2.0 * (sig[i] - target[i]) * kpref
        z = np.vdot(y, x)
        gG += c * z.real
        gD -= c * z.imag
        gS -= c * np.vdot(y, H_int @ x).imag / P.inter_scale
        xa = P.allow * x.reshape(shape, order="F")
        ya = P.allow * y.reshape(shape, order="F")
        for a in range(P.nalpha):
            us = P.u[:, :, :, vs[a]]
            px = fft_backward(np.einsum("nkx,nk->kx", us, xa[:, :, a]), P.kmesh, axis=0)
            py = fft_backward(np.einsum("nkx,nk->kx", us, ya[:, :, a]), P.kmesh, axis=0)
            Q += c / P.nk * (np.conj(py) * px).imag
    loss = float(np.sum((sig - np.asarray(target)) ** 2))
    out = dict(loss=loss, sigma=sig, g_gamma_eV=gG * EV2HARTREE, g_scissor_eV=gD * EV2HARTREE,
               g_scale=gS, dL_dW=Q)
    if wweights is not None:
        i0, w0, w1 = wweights
        g = np.zeros(nptab)
        np.add.at(g, i0.ravel(), (Q * w0).ravel())
        np.add.at(g, i0.ravel() + 1, (Q * w1).ravel())
        out["g_ptab"] = g
    return out
