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
#include "gemm64.hpp"

namespace og {

// The two x-side products are GEMMs (M = x or o, N = the nk nalpha columns, K = o
// or x) on the FP64 matrix cores; the fragment layout and the complex unrolling
// are described in gemm64.hpp.
// phi(k, x, a) = sum_o beta(k, o, a) obf(o, x)              (long_range.f90:373-382)
// beta [k][o][a] (k stride ldb), obf [o][x], phi [k][x][a] (k stride ldphi).
// Workgroup: 64 x points (four M tiles) by four N tiles, one per wave; the waves
// share the A panel through L1.  Loads of the next group of four o are in flight
// while the current group's eight instructions run.
template <int NA, int NT>
__global__ __launch_bounds__(256) void k_obf_fwd(const double2* __restrict__ beta, int ldb,
                                                 const double2* __restrict__ obf,
                                                 double2* __restrict__ phi, int ldphi, int nk,
                                                 int nx, int no) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, kk = lane >> 4;
  const int x0 = blockIdx.x * 64;
  MfCol col[NT];
  const double2* bp[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    col[n] = mf_col(((blockIdx.y * 4 + w) * NT + n) * 16 + i, NA, nk);
    bp[n] = beta + (size_t)col[n].k * ldb + col[n].a;
  }
  int xr[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) xr[m] = min(x0 + 16 * m + i, nx - 1);
  og_d4 acc[NT][4];
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[n][m] = og_d4{0.0, 0.0, 0.0, 0.0};
  // two register buffers, alternating: the loads of one group of four o are in
  // flight while the other group's matrix instructions run
  double2 bA[NT], aA[4], bB[NT], aB[4];
  auto load = [&](int o0, double2 (&bn)[NT], double2 (&an)[4]) {
    const int o = min(o0 + kk, no - 1);
#pragma unroll
    for (int n = 0; n < NT; ++n) bn[n] = bp[n][(size_t)o * NA];
#pragma unroll
    for (int m = 0; m < 4; ++m) an[m] = obf[(size_t)o * nx + xr[m]];
  };
  // C = A B: (ar, ai) x (br, bi) -> re: ar br - ai bi, im: ar bi + ai br
  auto step = [&](int o0, const double2 (&bv)[NT], const double2 (&an)[4]) {
    const bool ok = o0 + kk < no;
    double2 av[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) av[m] = ok ? an[m] : make_double2(0.0, 0.0);
    double b0[NT], b1[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      b0[n] = col[n].c ? bv[n].y : bv[n].x;
      b1[n] = col[n].c ? bv[n].x : -bv[n].y;
    }
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int m = 0; m < 4; ++m) acc[n][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m].x, b0[n], acc[n][m], 0, 0, 0);
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int m = 0; m < 4; ++m) acc[n][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m].y, b1[n], acc[n][m], 0, 0, 0);
  };
  load(0, bA, aA);
  for (int o0 = 0; o0 < no; o0 += 8) {
    load(o0 + 4, bB, aB);
    asm volatile("" ::: "memory");
    step(o0, bA, aA);
    load(o0 + 8, bA, aA);
    asm volatile("" ::: "memory");
    step(o0 + 4, bB, aB);
  }
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if (!col[n].ok) continue;
    double* __restrict__ out = reinterpret_cast<double*>(phi + (size_t)col[n].k * ldphi + col[n].a) + col[n].c;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = x0 + 16 * m + kk + 4 * r;
        if (x < nx) out[(size_t)x * NA * 2] = acc[n][m][r];
      }
  }
}

// b(k, o, a) = scale sum_x conj(obf(o, x)) Y(k, x, a)        (long_range.f90:412-421)
// Y [k][x][a] is the transform output (already multiplied by conj(phs)); obfT [x][o];
// b [k][o][a].  Workgroup: 64 basis functions (four M tiles) by one N tile; its four
// waves take consecutive quarters of the x range and their partial tiles are
// added in LDS in a fixed order, so the result is reproducible.
template <int NA, int NT>
__global__ __launch_bounds__(256) void k_obf_bwd(const double2* __restrict__ Y, int ldphi,
                                                 const double2* __restrict__ obfT,
                                                 double2* __restrict__ b, int ldb, int nk, int nx,
                                                 int no, double scale) {
  __shared__ double part[4][NT * 16][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform x quarter
  const int i = lane & 15, kk = lane >> 4;
  const int o0 = blockIdx.x * 64;
  MfCol col[NT];
  const double2* yp[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    col[n] = mf_col((blockIdx.y * NT + n) * 16 + i, NA, nk);
    yp[n] = Y + (size_t)col[n].k * ldphi + col[n].a;
  }
  int orow[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) orow[m] = min(o0 + 16 * m + i, no - 1);
  const int xa = (nx * w) >> 2, xb = (nx * (w + 1)) >> 2;
  og_d4 acc[NT][4];
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[n][m] = og_d4{0.0, 0.0, 0.0, 0.0};
  {
    // two register buffers, alternating, as in k_obf_fwd (an empty quarter, nx < 4,
    // loads row 0 and runs no step)
    double2 yA[NT], aA[4], yB[NT], aB[4];
    auto load = [&](int xs, double2 (&yn)[NT], double2 (&an)[4]) {
      const int x = max(min(xs + kk, xb - 1), 0);
#pragma unroll
      for (int n = 0; n < NT; ++n) yn[n] = yp[n][(size_t)x * NA];
#pragma unroll
      for (int m = 0; m < 4; ++m) an[m] = obfT[(size_t)x * no + orow[m]];
    };
    // C = conj(A) Y: re: ar yr + ai yi, im: ar yi - ai yr
    auto step = [&](int xs, const double2 (&yv)[NT], const double2 (&av)[4]) {
      const bool ok = xs + kk < xb;  // rows past the quarter contribute zero
      double b0[NT], b1[NT];
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const double2 y = ok ? yv[n] : make_double2(0.0, 0.0);
        b0[n] = col[n].c ? y.y : y.x;
        b1[n] = col[n].c ? -y.x : y.y;
      }
#pragma unroll
      for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[n][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m].x, b0[n], acc[n][m], 0, 0, 0);
#pragma unroll
      for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[n][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m].y, b1[n], acc[n][m], 0, 0, 0);
    };
    load(xa, yA, aA);
    for (int xs = xa; xs < xb; xs += 8) {
      load(xs + 4, yB, aB);
      asm volatile("" ::: "memory");
      step(xs, yA, aA);
      load(xs + 8, yA, aA);
      asm volatile("" ::: "memory");
      step(xs + 4, yB, aB);
    }
  }
  // every wave parks its tiles in LDS; wave 0 adds the four quarters in order
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[w][16 * n + 4 * m + r][lane] = acc[n][m][r];
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if (!col[n].ok) continue;
    double* __restrict__ out = reinterpret_cast<double*>(b + (size_t)col[n].k * ldb + col[n].a) + col[n].c;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = o0 + 16 * m + kk + 4 * r;
        double v = part[0][16 * n + 4 * m + r][lane];
#pragma unroll
        for (int j = 1; j < 4; ++j) v += part[j][16 * n + 4 * m + r][lane];
        if (o < no) out[(size_t)o * NA * 2] = v * scale;
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
