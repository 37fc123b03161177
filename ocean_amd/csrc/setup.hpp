// Device side of the multiplet set-up of OCEAN_create_central: the exchange/direct
// blocks of nbsemhsetup2 (OCEAN2/OCEAN_multiplet.f90:2007-2181) and the core
// spin-orbit matrix of soprep (:572-690).  Angular factors in closed form: the
// Gaunt integral of three Condon-Shortley harmonics through Racah's 3j formula.
#pragma once
#include "common.hpp"

namespace og {

__device__ __forceinline__ double fact_d(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= (double)i;
  return f;
}

// Wigner 3j symbol for integer arguments (Racah's formula).
__device__ inline double three_j_d(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0 || j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
  if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
  const double tri = fact_d(j1 + j2 - j3) * fact_d(j1 - j2 + j3) * fact_d(-j1 + j2 + j3) / fact_d(j1 + j2 + j3 + 1);
  const double norm = sqrt(tri * fact_d(j1 + m1) * fact_d(j1 - m1) * fact_d(j2 + m2) * fact_d(j2 - m2) *
                           fact_d(j3 + m3) * fact_d(j3 - m3));
  const int kmin = max(0, max(j2 - j3 - m1, j1 - j3 + m2));
  const int kmax = min(j1 + j2 - j3, min(j1 - m1, j2 + m2));
  double tot = 0.0;
  for (int k = kmin; k <= kmax; ++k) {
    const double den = fact_d(k) * fact_d(j1 + j2 - j3 - k) * fact_d(j1 - m1 - k) * fact_d(j2 + m2 - k) *
                       fact_d(j3 - j2 + m1 + k) * fact_d(j3 - j1 - m2 + k);
    tot += ((k & 1) ? -1.0 : 1.0) / den;
  }
  return (((j1 - j2 - m3) & 1) ? -1.0 : 1.0) * norm * tot;
}

// int conj(Y_l1m1) Y_kq(^*) Y_l3m3 dOmega (conj_mid: the middle harmonic conjugated),
// the quantity threey integrates (CNBSE/zcmsnmult/zf90/threey.f90).
__device__ inline double gaunt_d(int l1, int m1, int k, int q, int l3, int m3, bool conj_mid) {
  double sign = 1.0;
  if (conj_mid) {
    sign = (q & 1) ? -1.0 : 1.0;
    q = -q;
  }
  if (m1 != q + m3) return 0.0;
  const double fourpi = 12.566370614359172953850573533118;
  return sign * ((m1 & 1) ? -1.0 : 1.0) * sqrt((double)((2 * l1 + 1) * (2 * k + 1) * (2 * l3 + 1)) / fourpi) *
         three_j_d(l1, k, l3, 0, 0, 0) * three_j_d(l1, k, l3, -m1, q, m3);
}

// alpha = ((core spin) (2 lc + 1) + core m + lc) 2 + valence spin, spins -1/+1 as 0/1
// (OCEAN_multiplet.f90:343-355).
__device__ __forceinline__ void alpha_labels_d(int a, int lc, int& cs, int& cm, int& vs) {
  vs = (a & 1) ? 1 : -1;
  const int r = a >> 1;
  cm = r % (2 * lc + 1) - lc;
  cs = (r / (2 * lc + 1)) ? 1 : -1;
}

// One block mh_l (nbsemhsetup2), element (i1, i2) stored at i1 + i2 mham; the
// index inside the block is nu fastest, then valence m, then alpha
// (OCEAN_multiplet.f90:358-373).  slater: the F^k matrices for k = 2, 4, ..,
// min(2 lc, 2 lv), then the G^k for k = |lc - lv|, .., lc + lv, each npl x npl
// column-major (nu1, nu2) and already multiplied by its own scale factor.
//   direct  -F^k gaunt(lc, cm2, k, q, lc, cm1) gaunt*(lv, vm1, k, q, lv, vm2) 4pi/(2k+1)
//   exchange G^k gaunt(lc, cm2, k, q, lv, vm2) gaunt*(lv, vm1, k, q, lc, cm1) 4pi/(2k+1)
// (:2083-2160), all times scfac; the finished element is multiplied by post (the
// module's eV -> Hartree, OCEAN_multiplet.f90:388-389).  tdlda: the direct-like sum
// runs over k = 0 only, with the kk-file matrix in the F slot (:2082-2094).
__global__ void k_mham(int lc, int lv, int npl, const double* __restrict__ slater, double scfac, double post,
                       int tdlda, double* __restrict__ mhr, double* __restrict__ mhi) {
  const int nm = 2 * lv + 1;
  const int mham = 4 * (2 * lc + 1) * nm * npl;
  const int nF = tdlda ? 1 : (min(2 * lc, 2 * lv) >= 2 ? min(2 * lc, 2 * lv) / 2 : 0);
  const double fourpi = 12.566370614359172953850573533118;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < mham * mham; e += gridDim.x * blockDim.x) {
    const int i1 = e % mham, i2 = e / mham;
    const int nu1 = i1 % npl, vm1 = (i1 / npl) % nm - lv, a1 = i1 / (npl * nm);
    const int nu2 = i2 % npl, vm2 = (i2 / npl) % nm - lv, a2 = i2 / (npl * nm);
    int cs1, cm1, vs1, cs2, cm2, vs2;
    alpha_labels_d(a1, lc, cs1, cm1, vs1);
    alpha_labels_d(a2, lc, cs2, cm2, vs2);
    double acc = 0.0;
    if (cs2 == cs1 && vs1 == vs2 && cm2 + vm1 == cm1 + vm2) {
      const int q = cm2 - cm1;
      for (int j = 0; j < nF; ++j) {
        const int k = tdlda ? 0 : 2 * (j + 1);
        if (abs(q) > k) continue;
        const double f = slater[(size_t)j * npl * npl + nu1 + (size_t)nu2 * npl];
        acc -= f * scfac * gaunt_d(lc, cm2, k, q, lc, cm1, false) * gaunt_d(lv, vm1, k, q, lv, vm2, true) *
               (fourpi / (2 * k + 1));
      }
    }
    if (cs2 == vs2 && vs1 == cs1 && cm2 + vm1 == vm2 + cm1) {
      const int q = cm2 - vm2;
      int j = nF;
      for (int k = abs(lc - lv); k <= lc + lv; k += 2, ++j) {
        if (abs(q) > k) continue;
        const double g = slater[(size_t)j * npl * npl + nu1 + (size_t)nu2 * npl];
        acc += g * scfac * gaunt_d(lc, cm2, k, q, lv, vm2, false) * gaunt_d(lv, vm1, k, q, lc, cm1, true) *
               (fourpi / (2 * k + 1));
      }
    }
    mhr[e] = acc * post;
    mhi[e] = 0.0;
  }
}

// S(ic, jc) = -xi <cm_jc, cs_jc | L.S | cm_ic, cs_ic> where the valence spins agree
// (soprep through jimel, OCEAN_multiplet.f90:642-671; CNBSE/zcmsnmult/zf90/jimel.f90),
// with L.S = Lz Sz + (L+ S- + L- S+) / 2; the spin-orbit-dependent lifetimes add
// i (-<L.S> l_alpha - l_beta on the diagonal) (:664-668).  somelr/someli(ic, jc)
// column-major.
__global__ void k_soprep(int lc, double xi, double l_alpha, double l_beta, double* __restrict__ sr,
                         double* __restrict__ si) {
  const int na = 4 * (2 * lc + 1);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < na * na; e += gridDim.x * blockDim.x) {
    const int ic = e % na, jc = e / na;
    int s1, m1, v1, s2, m2, v2;
    alpha_labels_d(ic, lc, s1, m1, v1);
    alpha_labels_d(jc, lc, s2, m2, v2);
    double v = 0.0;
    if (v1 == v2) {
      const double ll = (double)(lc * (lc + 1));
      if (m2 == m1 && s2 == s1) v += (double)m1 * 0.5 * (double)s1;
      if (m2 == m1 + 1 && s1 == 1 && s2 == -1) v += 0.5 * sqrt(ll - (double)(m1 * (m1 + 1)));  // L+ S-
      if (m2 == m1 - 1 && s1 == -1 && s2 == 1) v += 0.5 * sqrt(ll - (double)(m1 * (m1 - 1)));  // L- S+
    }
    sr[e] = -xi * v;
    si[e] = -v * l_alpha - (ic == jc && v1 == v2 ? l_beta : 0.0);
  }
}

}  // namespace og
