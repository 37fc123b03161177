"""Host-side setup for the core-level BSE action (product code, numpy).

These build the module state that OCEAN2's Fortran host builds before the hot
loop and hands to the action: the phased Bloch functions, W(k,x), and the
multiplet matrices.  In a real OCEAN2 run the Fortran host keeps doing this and
passes the arrays through the C-ABI; here it feeds the synthetic workloads and
the device-side builders are checked against it.

Reference semantics followed (OCEAN 3.2.1 tree):
  bloch_phase      OCEAN2/OCEAN_bloch.f90:54-219 (OCEAN_bloch_lrLOAD)
  w_geometry/build_w  OCEAN2/long_range.f90:1659-1853 (lr_populate_W2),
                   CNBSE/zcmsnmult/zf90/intval.f90:8-57
  exchange_matrix  OCEAN2/OCEAN_multiplet.f90:2007-2181 (nbsemhsetup2), with the
                   angular integrals of CNBSE/zcmsnmult/zf90/newthreey.f90 in
                   closed form (Gaunt coefficients from 3j symbols)
  spin_orbit       OCEAN2/OCEAN_multiplet.f90:572-690 (OCEAN_soprep) with
                   <l m|L_i|l m'><s|S_i|s'> from ladder-operator closed forms
  normalise_psi0   OCEAN2/OCEAN_psi.f90:5589-5638 (core_pnorm)
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

EV2HARTREE = 0.036749322175666   # OCEAN2/OCEAN_constants.f90:15
HARTREE2EV = 27.211386245980     # OCEAN2/OCEAN_constants.f90:16


def alpha_table(lc):
    """(2*cms, cml, 2*vms) per alpha; core spin outer, valence spin fastest."""
    return [(s, m, v) for s in (-1, 1) for m in range(-lc, lc + 1) for v in (-1, 1)]


def x_offsets(xmesh, kmesh, tau, xshift_override=(1, 1, 1)):
    """The site centring and wrapped x offsets of OCEAN_bloch_lrLOAD.

    Returns (xoff int32 [x, 3] with ix fastest over x, tau_centred, xshift); xoff
    is the layout ocean_gpu_load_bloch_raw takes (xoff(3, nx) in Fortran order).
    """
    xm = np.array(xmesh, np.int64)
    km = np.array(kmesh, np.int64)
    tau = np.asarray(tau, np.float64)
    half = np.where(km % 2 == 0, 0.0, 0.5)
    xshift = np.floor(xm * (tau - half)).astype(np.int64) * np.asarray(xshift_override, np.int64)
    tau_c = tau - xshift / xm
    # x grid, ix fastest
    iz, iy, ix = np.meshgrid(np.arange(1, xm[2] + 1), np.arange(1, xm[1] + 1),
                             np.arange(1, xm[0] + 1), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)  # [x, 3], 1-based
    targ = idx - xshift[None, :]
    wrap = np.where(targ > xm[None, :], -xm[None, :], np.where(targ < 1, xm[None, :], 0))
    return np.ascontiguousarray(wrap + idx - 1, dtype=np.int32), tau_c, xshift


def obf_x_offsets(xmesh, kmesh, tau):
    """The site centring and wrapped x offsets of OCEAN_obf_make_phase.

    As x_offsets, except that for an odd kmesh(1) the shift along the first axis is
    floor(xmesh(2) (tau_1 - 1/2)), as OCEAN2/OCEAN_obf.f90:300-301 has it (the Bloch
    path, and the OBF path for an even kmesh(1), use xmesh(1)).  Returns
    (xoff int32 [x, 3], tau_centred, xshift); xoff is what ocean_gpu_load_obf takes.
    """
    xm = np.array(xmesh, np.int64)
    km = np.array(kmesh, np.int64)
    tau = np.asarray(tau, np.float64)
    half = np.where(km % 2 == 0, 0.0, 0.5)
    xs_mesh = xm.copy()
    if km[0] % 2 == 1:
        xs_mesh[0] = xm[1]
    xshift = np.floor(xs_mesh * (tau - half)).astype(np.int64)
    tau_c = tau - xshift / xm
    iz, iy, ix = np.meshgrid(np.arange(1, xm[2] + 1), np.arange(1, xm[1] + 1),
                             np.arange(1, xm[0] + 1), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    targ = idx - xshift[None, :]
    wrap = np.where(targ > xm[None, :], -xm[None, :], np.where(targ < 1, xm[None, :], 0))
    return np.ascontiguousarray(wrap + idx - 1, dtype=np.int32), tau_c, xshift


def obf_fake_bloch(obf, obf2u):
    """Bloch functions built from the OBF form as OCEAN_obf_lrLOAD does for
    bloch_control's use_fake_obf (OCEAN2/OCEAN_obf.f90:86-118), before the phase:
    u(n,k,x) = sum_o obf2u(n,o,k) obf(o,x) / sqrt(su(n,k)), where su adds
    |partial sum over o|^2 after every basis function (the accumulation sits inside
    the o loop, :88-97), not only the square of the finished sum.

    obf [o, x], obf2u [n, o, k]; returns u_raw (nb, nk, nx, 1) for load_bloch_raw
    (the phase and centring of that path are the Bloch form's, :121-262)."""
    obf = np.asarray(obf, np.complex128)
    c = np.asarray(obf2u, np.complex128)
    nb, no, nk = c.shape
    nx = obf.shape[1]
    u = np.zeros((nb, nk, nx, 1), np.complex128, order="F")
    for k in range(nk):
        part = np.cumsum(c[:, :, k][:, :, None] * obf[None, :, :], axis=1)  # (nb, o, nx)
        su = np.sum(part.real ** 2 + part.imag ** 2, axis=(1, 2))
        u[:, k, :, 0] = part[:, -1, :] / np.sqrt(su)[:, None]
    return u


def bloch_phase(u_raw, xmesh, kmesh, tau, xshift_override=(1, 1, 1)):
    """Phase u(n,k,x) by exp(2 pi i sum_d (x_d + wrap_d) k_d / (Nx_d Nk_d)).

    Returns (u_phased, tau_centred, xshift).  Vectorised over (k, x).
    """
    xm = np.array(xmesh, np.int64)
    km = np.array(kmesh, np.int64)
    xo, tau_c, xshift = x_offsets(xmesh, kmesh, tau, xshift_override)
    xoff = xo.astype(np.float64)  # [x, 3]
    k1, k2, k3 = np.meshgrid(np.arange(km[0]), np.arange(km[1]), np.arange(km[2]), indexing="ij")
    kid = np.stack([k1.ravel(), k2.ravel(), k3.ravel()], axis=1).astype(np.float64)  # [k, 3]
    ph = np.zeros((kid.shape[0], xoff.shape[0]))
    for d in range(3):
        ph = ph + 2.0 * np.pi * np.outer(kid[:, d], xoff[:, d]) / float(xm[d] * km[d])
    rot = np.cos(ph) + 1j * np.sin(ph)
    u = np.asfortranarray(u_raw * rot[None, :, :, None])
    return u, tau_c, xshift


def w_geometry(avec, kmesh, xmesh, tau_c, xshift):
    """r(k, x) = |A (R_k + f_x - tau)| with the reference's k and x wraps.

    Returns (r[k, x], cart[k, x, 3]).
    """
    avec = np.asarray(avec, np.float64)
    km = np.array(kmesh, np.int64)
    xm = np.array(xmesh, np.int64)
    iz, iy, ix = np.meshgrid(np.arange(1, xm[2] + 1), np.arange(1, xm[1] + 1),
                             np.arange(1, xm[0] + 1), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    targ = idx - np.asarray(xshift, np.int64)[None, :]
    targ = np.where(targ > xm[None, :], targ - xm[None, :], np.where(targ < 1, targ + xm[None, :], targ))
    fr = (targ - 1) / xm[None, :]  # [x, 3]

    def wrapk(n):
        kk = np.arange(n)
        kk = np.where(kk >= (n + 1) // 2, kk - n, kk)
        return np.zeros(n, np.int64) if n == 1 else kk

    a, b, c = np.meshgrid(wrapk(km[0]), wrapk(km[1]), wrapk(km[2]), indexing="ij")
    R = np.stack([a.ravel(), b.ravel(), c.ravel()], axis=1).astype(np.float64)  # [k, 3]
    frac = R[:, None, :] + fr[None, :, :] - np.asarray(tau_c, np.float64)[None, None, :]
    cart = frac @ avec.T  # cart_i = sum_j avec(i, j) frac_j
    r = np.sqrt(np.sum(cart * cart, axis=-1))
    return r, cart


def interp_cap(rtab, ptab, r):
    """Piecewise-linear table lookup, clamped at both ends (intval 'cap','cap')."""
    rtab = np.asarray(rtab, np.float64)
    ptab = np.asarray(ptab, np.float64)
    j = np.clip(np.searchsorted(rtab, r, side="left") - 1, 0, rtab.size - 2)
    t = (r - rtab[j]) / (rtab[j + 1] - rtab[j])
    v = ptab[j] + t * (ptab[j + 1] - ptab[j])
    return np.where(r < rtab[0], ptab[0], np.where(r > rtab[-1], ptab[-1], v))


def build_w(avec, kmesh, xmesh, tau_c, xshift, rtab, ptab, epsilon0, inter_scale=1.0):
    """W(k, x) = s * { table(r) for r <= rtab_n ; 1/(eps r) beyond }  (periodic, bulk)."""
    r, _ = w_geometry(avec, kmesh, xmesh, tau_c, xshift)
    rr = np.where(r > 0, r, 1.0)
    pot = np.where(r > rtab[-1], (1.0 / epsilon0) / rr, interp_cap(rtab, ptab, r))
    return np.asfortranarray(pot * inter_scale)


# --------------------------------------------------------------------------
# Angular algebra in closed form
# --------------------------------------------------------------------------
@lru_cache(maxsize=None)
def _fact(n):
    return math.factorial(n)


def three_j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol for integer arguments (Racah's formula)."""
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    tri = _fact(j1 + j2 - j3) * _fact(j1 - j2 + j3) * _fact(-j1 + j2 + j3) / _fact(j1 + j2 + j3 + 1)
    norm = math.sqrt(tri * _fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2)
                     * _fact(j3 + m3) * _fact(j3 - m3))
    tot = 0.0
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    for k in range(kmin, kmax + 1):
        tot += (-1) ** k / (_fact(k) * _fact(j1 + j2 - j3 - k) * _fact(j1 - m1 - k)
                            * _fact(j2 + m2 - k) * _fact(j3 - j2 + m1 + k) * _fact(j3 - j1 - m2 + k))
    return (-1) ** (j1 - j2 - m3) * norm * tot


def gaunt(l1, m1, k, q, l3, m3, conj_mid):
    """int conj(Y_l1m1) Y_kq(^*) Y_l3m3 dOmega with Condon-Shortley Ylm."""
    sign = 1.0
    if conj_mid:
        sign = (-1.0) ** q
        q = -q
    if m1 != q + m3:
        return 0.0
    return sign * ((-1.0) ** m1) * math.sqrt((2 * l1 + 1) * (2 * k + 1) * (2 * l3 + 1) / (4 * math.pi)) \
        * three_j(l1, k, l3, 0, 0, 0) * three_j(l1, k, l3, -m1, q, m3)


def exchange_matrix(lc, lv, nproj, fk, gk, scfac=1.0, tdlda=False):
    """Multiplet block mh_l (eV) for valence l, element [i1, i2] (nbsemhsetup2 layout).

    fk[k], gk[k]: (nproj, nproj) Slater integrals already multiplied by their scale
    factors.  Index order inside the block: nu fastest, then valence m, then alpha.
    tdlda: the direct-like sum runs over k = 0 only (fk[0] from the kk file).
    """
    labels = [(s, m, v, vm, nu) for (s, m, v) in alpha_table(lc)
              for vm in range(-lv, lv + 1) for nu in range(nproj)]
    n = len(labels)
    mh = np.zeros((n, n), np.complex128, order="F")
    kf = [0] if tdlda else list(range(2, min(2 * lc, 2 * lv) + 1, 2))
    kg = list(range(abs(lc - lv), lc + lv + 1, 2))
    for i1, (cs1, cm1, vs1, vm1, nu1) in enumerate(labels):
        for i2, (cs2, cm2, vs2, vm2, nu2) in enumerate(labels):
            acc = 0.0
            # direct-like F^k: core (i2 -> i1), valence (i1 <- i2)
            if cs2 == cs1 and vs1 == vs2 and cm2 + vm1 == cm1 + vm2:
                q = cm2 - cm1
                for k in kf:
                    if abs(q) <= k:
                        acc -= fk[k][nu1, nu2] * scfac * gaunt(lc, cm2, k, q, lc, cm1, False) \
                            * gaunt(lv, vm1, k, q, lv, vm2, True) * (4 * math.pi / (2 * k + 1))
            # exchange G^k: core(i2) <-> valence(i2), valence(i1) <-> core(i1)
            if cs2 == vs2 and vs1 == cs1 and cm2 + vm1 == vm2 + cm1:
                q = cm2 - vm2
                for k in kg:
                    if abs(q) <= k:
                        acc += gk[k][nu1, nu2] * scfac * gaunt(lc, cm2, k, q, lv, vm2, False) \
                            * gaunt(lv, vm1, k, q, lc, cm1, True) * (4 * math.pi / (2 * k + 1))
            mh[i1, i2] = acc
    return mh


def _lm_matrix(l):
    """<l m | L_i | l m'> for i = x, y, z as (2l+1)^2 matrices, m = -l..l."""
    ms = np.arange(-l, l + 1)
    n = ms.size
    Lp = np.zeros((n, n))
    for j, m in enumerate(ms[:-1]):
        Lp[j + 1, j] = math.sqrt(l * (l + 1) - m * (m + 1))
    Lm = Lp.T.copy()
    Lx = 0.5 * (Lp + Lm)
    Ly = -0.5j * (Lp - Lm)
    Lz = np.diag(ms.astype(float))
    return Lx + 0j, Ly, Lz + 0j


def spin_orbit(lc, xi_eV, life_time_eV=None):
    """S(ic, jc) = -xi <core ic| L.S |core jc> where valence spins match (Hartree).

    Returns (somelr, someli, have_spin_orbit).  life_time_eV (the two values of
    core_broaden.ipt) adds i(-<L.S> l_alpha - l_beta delta) with l_alpha and l_beta
    chosen so that the j = lc + 1/2 and lc - 1/2 levels get -life_time(1) and
    -life_time(2) (OCEAN2/OCEAN_multiplet.f90:613-627, :664-668); without it the
    imaginary part is zero.
    """
    xi = xi_eV * EV2HARTREE
    lab = alpha_table(lc)
    na = len(lab)
    L = _lm_matrix(lc)
    S = _lm_matrix_half()
    ls = np.zeros((na, na))
    for ic, (s1, m1, v1) in enumerate(lab):
        for jc, (s2, m2, v2) in enumerate(lab):
            if v1 != v2:
                continue
            # jimel(l, cml(jc), cml(ic), i) * jimel(1/2, cms(jc), cms(ic), i):
            # <m_jc | L_i | m_ic> <s_jc | S_i | s_ic>
            val = 0j
            for i in range(3):
                val += L[i][m2 + lc, m1 + lc] * S[i][(s2 + 1) // 2, (s1 + 1) // 2]
            ls[ic, jc] = val.real
    sr = np.asfortranarray(-xi * ls)
    si = np.zeros((na, na), order="F")
    if life_time_eV is None:
        return sr, si, xi >= 1e-12
    t1, t2 = (float(v) * EV2HARTREE for v in life_time_eV[:2])
    l_alpha = 2.0 / (2 * lc + 1) * (t1 - t2)
    l_beta = (lc + 1) / (2 * lc + 1) * t1 + lc / (2 * lc + 1) * t2
    for ic in range(na):
        for jc in range(na):
            if lab[ic][2] == lab[jc][2]:
                si[ic, jc] = -ls[ic, jc] * l_alpha - (l_beta if ic == jc else 0.0)
    return sr, si, True


def _lm_matrix_half():
    sx = 0.5 * np.array([[0, 1], [1, 0]], complex)
    sy = 0.5 * np.array([[0, 1j], [-1j, 0]], complex)  # rows/cols ordered m = -1/2, +1/2
    sz = 0.5 * np.array([[-1, 0], [0, 1]], complex)
    return sx, sy, sz


def trim_core_hole_potential(v, rv, rmax):
    """Split the screened core-hole potential at the projector radius, as
    nbsemkcmel does before the multiplet integrals (OCEAN2/nbsemkcmel.f90:98-127):
    j is where its bisection over rv for rmax ends (1-based, moved past rv(j) <
    rmax); inside, the long-range table keeps the flat value v(j) and the
    difference v - v(j) goes to the projector matrix.  Returns (vtrim, vdiff)."""
    v = np.asarray(v, np.float64)
    rv = np.asarray(rv, np.float64)
    n = v.size
    lo, hi = 1, n
    while lo < hi:
        mid = (lo + hi) // 2
        if rv[mid - 1] < rmax:
            lo = mid + 1
        elif rv[mid - 1] > rmax:
            hi = mid - 1
        else:
            break
    if rv[lo - 1] < rmax:
        lo += 1
    if lo > n:
        raise ValueError("the core-hole potential table ends inside the projector radius")
    inside = np.arange(n) < lo - 1
    vtrim = np.where(inside, v[lo - 1], v)
    vdiff = np.where(inside, v - v[lo - 1], 0.0)
    return vtrim, vdiff


def projector_cmel(rad, phi, rv, vdiff, rmax):
    """The direct core-hole matrix of the projectors and their overlap,
    cmel(mu, nu) = int phi_mu phi_nu vdiff r^2 dr (eV) and nmel (no potential), by
    nbsemkcmel's Bode rule on the logarithmic grid up to rmax
    (OCEAN2/nbsemkcmel.f90:145-204): four-point panels ending at rmax with weights
    (28, 64, 24, 64)/45, half end steps of the log measure, vdiff interpolated
    linearly (intval with 'err': the grid must lie inside the table).
    rad[nr], phi[nr, nnu]; returns (cmel_eV, nmel) of shape (nnu, nnu)."""
    rad = np.asarray(rad, np.float64)
    phi = np.asarray(phi, np.float64).reshape(rad.size, -1)
    nr = rad.size
    dl = math.log(rad[-1] / rad[0]) / (nr - 1)
    ix = 1 + int(math.floor(math.log(rmax / rad[0]) / dl + 0.5))  # nint of a positive value
    if ix < 5 or ix > nr:
        raise ValueError("rmax does not lie on the radial grid")
    dr = dl * rad[:ix]
    dr[0] *= 0.5
    dr[ix - 1] *= 0.5
    rv = np.asarray(rv, np.float64)
    r = rad[:ix]
    if r.min() < rv[0] or r.max() > rv[-1]:
        raise ValueError("the radial grid leaves the core-hole potential table (intval 'err')")
    val = interp_cap(rv, vdiff, r)
    # panel weights: panel ending at i (1-based, i = ix, ix-4, .., >= 5) gives point
    # i - j the weight b(j)
    w = np.zeros(ix)
    bw = np.array([28.0, 64.0, 24.0, 64.0]) / 45.0
    for i in range(ix, 4, -4):
        w[i - 4:i] += bw[::-1]
    q = w * dr * r * r
    cmel = (phi[:ix].T * (q * val)) @ phi[:ix] * HARTREE2EV
    nmel = (phi[:ix].T * q) @ phi[:ix]
    return cmel, nmel


def radial_cmel(rgrid, phis, vdiff):
    """Direct core-projector matrix int phi_mu phi_nu vdiff r^2 dr (eV), trapezoid rule."""
    f = phis[:, :, None] * phis[:, None, :] * (vdiff * rgrid ** 2)[:, None, None]
    return np.trapezoid(f, rgrid, axis=0) * HARTREE2EV


def normalise_psi0(psi, nk, celvol):
    """Return (psi / |psi|, kpref = 4 pi |psi|^2 / (N_k Omega^2))."""
    nrm2 = float(np.sum(psi.real ** 2 + psi.imag ** 2))
    kpref = 4.0 * math.pi * nrm2 / (nk * celvol ** 2)
    return np.asfortranarray(psi / math.sqrt(nrm2)), kpref
