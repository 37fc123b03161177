// Orbital-basis-function (OBF) form of the long-range term, OCEAN2/long_range.f90:262-455
// (lr_act_obf2, selected by bloch_control's use_obf, :81-82 and :128-146).
//
// The Bloch functions are u(n,k,x) = sum_o obf2u(n,o,k) obf(o,x) phs(k,x).  The band
// side (beta = obf2u . psi and hpsi = conj(obf2u) . b) has the shape of the Bloch
// projection and adjoint with the x index replaced by o, so it runs on k_proj and
// k_adj over [k][n][o] and [k][o][n] copies of obf2u.  This file holds the x side:
// the two products with obf and the set-up kernels; the phase rides on the
// transform (k_fft_w with PHS).
#pragma once
#include "common.hpp"

namespace og {

// phi(k, x, a) = sum_o beta(k, o, a) obf(o, x)              (long_range.f90:373-382)
// beta [k][o][a] (k stride ldb), obf [o][x], phi [k][x][a] (k stride ldphi).
// Lanes run over x, so obf rows are read coalesced; beta depends only on the
// workgroup's k tile and the loop index, so it is read through the scalar cache.
template <int NA, int KT>
__global__ __launch_bounds__(256) void k_obf_fwd(const double2* __restrict__ beta, int ldb,
                                                 const double2* __restrict__ obf,
                                                 double2* __restrict__ phi, int ldphi, int nk,
                                                 int nx, int no) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int xc = min(x, nx - 1);
  const int k0 = blockIdx.y * KT;
  double2 acc[KT][NA];
#pragma unroll
  for (int t = 0; t < KT; ++t)
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[t][a] = make_double2(0.0, 0.0);
  const double2* __restrict__ fp = obf + xc;
#pragma unroll 2
  for (int o = 0; o < no; ++o) {
    const double2 f = fp[(size_t)o * nx];
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      const double2* __restrict__ bp = beta + (size_t)min(k0 + t, nk - 1) * ldb + (size_t)o * NA;
#pragma unroll
      for (int a = 0; a < NA; ++a) cfma(acc[t][a], bp[a], f);
    }
  }
  if (x >= nx) return;
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    if (k0 + t >= nk) break;
    double2* __restrict__ out = phi + (size_t)(k0 + t) * ldphi + (size_t)x * NA;
#pragma unroll
    for (int a = 0; a < NA; ++a) out[a] = acc[t][a];
  }
}

// b(k, o, a) = scale sum_x conj(obf(o, x)) Y(k, x, a)        (long_range.f90:412-421)
// Y [k][x][a] is the transform output (already multiplied by conj(phs)); obfT [x][o];
// b [k][o][a].  Lanes run over o (coalesced obfT rows); the four waves of a
// workgroup take consecutive quarters of the x range and their partial sums are
// added in LDS in a fixed order, so the result is reproducible.  Y depends only on
// the wave's x index and the k tile and is read through the scalar cache.
template <int NA, int KT>
__global__ __launch_bounds__(256) void k_obf_bwd(const double2* __restrict__ Y, int ldphi,
                                                 const double2* __restrict__ obfT,
                                                 double2* __restrict__ b, int ldb, int nk, int nx,
                                                 int no, double scale) {
  constexpr int XS = 4;
  __shared__ double2 part[XS - 1][KT * NA][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int o = blockIdx.x * 64 + lane;
  const int oc = min(o, no - 1);
  const int k0 = blockIdx.y * KT;
  const int xa = (nx * w) / XS, xb = (nx * (w + 1)) / XS;
  double2 acc[KT][NA];
#pragma unroll
  for (int t = 0; t < KT; ++t)
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[t][a] = make_double2(0.0, 0.0);
  const double2* __restrict__ fp = obfT + oc;
#pragma unroll 2
  for (int x = xa; x < xb; ++x) {
    const double2 f = fp[(size_t)x * no];
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      const double2* __restrict__ yp = Y + (size_t)min(k0 + t, nk - 1) * ldphi + (size_t)x * NA;
#pragma unroll
      for (int a = 0; a < NA; ++a) cfmac(acc[t][a], f, yp[a]);
    }
  }
  if (w > 0) {
#pragma unroll
    for (int t = 0; t < KT; ++t)
#pragma unroll
      for (int a = 0; a < NA; ++a) part[w - 1][t * NA + a][lane] = acc[t][a];
  }
  __syncthreads();
  if (w != 0 || o >= no) return;
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    if (k0 + t >= nk) break;
    double2* __restrict__ out = b + (size_t)(k0 + t) * ldb + (size_t)o * NA;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      double2 v = acc[t][a];
#pragma unroll
      for (int j = 0; j < XS - 1; ++j) {
        const double2 p = part[j][t * NA + a][lane];
        v.x += p.x;
        v.y += p.y;
      }
      out[a] = make_double2(v.x * scale, v.y * scale);
    }
  }
}

// obf_phs(k, x) = exp(i (phsx + phsy + phsz)),
// phs_d = 2 pi xoff_d (q_d - 1) / (X_d K_d)                  (OCEAN_obf.f90:360-372)
// from the wrapped offsets xoff(3, nx); stored [x][k].
__global__ void k_obf_phase(const int* __restrict__ xoff, double2* __restrict__ phs, int nk, int nx,
                            int K1, int K2, int K3, int X1, int X2, int X3) {
  const double twopi = 6.283185307179586476925286766559;
  const size_t tot = (size_t)nk * nx;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(e / nk), k = (int)(e % nk);
    const int q1 = k / (K2 * K3), q2 = (k / K3) % K2, q3 = k % K3;
    const double px = twopi * (double)(xoff[3 * x + 0] * q1) / (double)(X1 * K1);
    const double py = twopi * (double)(xoff[3 * x + 1] * q2) / (double)(X2 * K2);
    const double pz = twopi * (double)(xoff[3 * x + 2] * q3) / (double)(X3 * K3);
    double sp, cp;
    sincos(px + py + pz, &sp, &cp);
    phs[e] = make_double2(cp, sp);
  }
}

// obf2u planes (nb, no, nk) (re/im_obf2u, OCEAN_obf.f90:704-715) -> the projection
// copy u[k][n][o] (row stride ldu) and the adjoint copy uT[k][o][n] (k stride ldT).
__global__ void k_obf_load2u(const double* __restrict__ re, const double* __restrict__ im,
                             double2* __restrict__ u, double2* __restrict__ uT, int nb, int no, int nk,
                             int ldu, int ldT) {
  const size_t tot = (size_t)nb * no * nk;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int n = (int)(e % nb);
    const size_t r = e / nb;
    const int o = (int)(r % no), k = (int)(r / no);
    const double2 v = make_double2(re[e], im[e]);
    uT[(size_t)k * ldT + (size_t)o * nb + n] = v;
    u[((size_t)k * nb + n) * ldu + o] = v;
  }
}

// obf planes (no, nx) (re/im_obf, OCEAN_obf.f90:551-600) -> obf[o][x] and obfT[x][o].
__global__ void k_obf_loadf(const double* __restrict__ re, const double* __restrict__ im,
                            double2* __restrict__ f, double2* __restrict__ fT, int no, int nx) {
  const size_t tot = (size_t)no * nx;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int o = (int)(e % no), x = (int)(e / no);
    const double2 v = make_double2(re[e], im[e]);
    fT[e] = v;
    f[(size_t)o * nx + x] = v;
  }
}

}  // namespace og
