// FP64 matrix-core GEMM pieces shared by the OBF x-side products and the matrix-core
// adjoint of the many-alpha edges.
#pragma once
#include "action.hpp"
#include "common.hpp"

namespace og {

// Complex GEMMs on the FP64 matrix cores: v_mfma_f64_16x16x4f64, where lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives D[l / 16 + 4 r]
// [l % 16], r = 0..3 (measured on gfx950; the FP32 16x16x4 form returns rows
// 4 (l / 16) + r instead).  The complex product is a real GEMM with the complex parts
// unrolled along K: for every group of four K indices (kk = l / 16) the first
// instruction takes the real parts of A and the second the imaginary parts, with
// B[(kk, re)][(col, c)] and B[(kk, im)][(col, c)] formed from the lane's own complex
// B element, so each lane loads one complex A element per M tile and one complex B
// element per group.  Output columns are flattened as G = 2 (k nalpha + a) + c
// (c = real/imaginary part); an N tile is 16 consecutive G.

struct MfCol {
  int k, a, c;
  bool ok;
};
__device__ __forceinline__ MfCol mf_col(int G, int na, int nk) {
  MfCol r;
  r.k = G / (2 * na);
  r.a = (G >> 1) % na;
  r.c = G & 1;
  r.ok = r.k < nk;
  r.k = min(r.k, nk - 1);
  return r;
}


typedef float __attribute__((ext_vector_type(4))) og_f4;

// One 16x16x4 matrix-core step in either precision, and the output row of a lane's
// r-th accumulator: the f64 form returns rows l/16 + 4r, the f32 form 4(l/16) + r
// (both measured on gfx950).
template <typename T> struct Mf;
template <> struct Mf<double> {
  using V = og_d4;
  static __device__ __forceinline__ V step(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kk, int r) { return kk + 4 * r; }
};
template <> struct Mf<float> {
  using V = og_f4;
  static __device__ __forceinline__ V step(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kk, int r) { return 4 * kk + r; }
};

// lr(n, k, a) = sum_x conj(uT(k, x, n)) phi'(k, x, a)          (long_range.f90:943-961)
// for the many-alpha edges, written to lrbuf [a][k][n] (FP64) for the adjoint's
// epilogue.  The VALU adjoint reads each phi'(k, x, :) as nalpha LDS broadcasts per
// band row, which bounds it at twelve alphas; here it is the B operand of a GEMM
// per k (M = bands, N = the 2 nalpha real columns, K = x), in the precision of the
// stored u (FP64, or FP32 for the mixed form, which accumulates in FP32 as its VALU
// adjoint does).  Workgroup: 64 bands (four M tiles) by all N tiles of one k, so
// every uT element is read once; the four waves take quarters of the x range and
// their tiles are added in LDS in a fixed order.  nspn = 1.
template <typename T2, int NA>
__device__ __forceinline__ void adj_lr_body(const T2* __restrict__ uT, int ldT, const T2* __restrict__ phi,
                                            int ldphi, double2* __restrict__ lrbuf, int nb, int nk, int nx, int bx,
                                            int k) {
  using T = typename Scalar<T2>::T;
  using M = Mf<T>;
  using V = typename M::V;
  constexpr int NT = (2 * NA + 15) / 16;
  __shared__ T part[4][NT * 16][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 15, kk = lane >> 4;
  const int n0 = bx * 64;
  int ca[NT], cc[NT];
  bool cok[NT];
  const T2* yp[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int G = t * 16 + i;
    ca[t] = G >> 1;
    cc[t] = G & 1;
    cok[t] = ca[t] < NA;
    yp[t] = phi + (size_t)k * ldphi + min(ca[t], NA - 1);
  }
  const T2* __restrict__ ap = uT + (size_t)k * ldT;
  int nrow[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) nrow[m] = min(n0 + 16 * m + i, nb - 1);
  const int xa = (nx * w) >> 2, xb = (nx * (w + 1)) >> 2;
  V acc[NT][4];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[t][m] = V{0, 0, 0, 0};
  {
    T2 yA[NT], aA[4], yB[NT], aB[4];
    auto load = [&](int xs, T2 (&yn)[NT], T2 (&an)[4]) {
      const int x = max(min(xs + kk, xb - 1), 0);
#pragma unroll
      for (int t = 0; t < NT; ++t) yn[t] = yp[t][(size_t)x * NA];
#pragma unroll
      for (int m = 0; m < 4; ++m) an[m] = ld_stream(ap + (size_t)x * nb + nrow[m]);
    };
    // C = conj(A) Y: re: ar yr + ai yi, im: ar yi - ai yr
    auto step = [&](int xs, const T2 (&yv)[NT], const T2 (&av)[4]) {
      const bool ok = xs + kk < xb;
      T b0[NT], b1[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const T2 y = ok ? yv[t] : cmake<T2>(0, 0);
        b0[t] = cc[t] ? y.y : y.x;
        b1[t] = cc[t] ? -y.x : y.y;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[t][m] = M::step(av[m].x, b0[t], acc[t][m]);
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[t][m] = M::step(av[m].y, b1[t], acc[t][m]);
    };
    load(xa, yA, aA);
    for (int xs = xa; xs < xb; xs += 8) {
      load(xs + 4, yB, aB);
      asm volatile("" ::: "memory");  // keep the next group's loads ahead of this group's products
      step(xs, yA, aA);
      load(xs + 8, yA, aA);
      asm volatile("" ::: "memory");
      step(xs + 4, yB, aB);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[w][16 * t + 4 * m + r][lane] = acc[t][m][r];
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (!cok[t]) continue;
    double* __restrict__ out = reinterpret_cast<double*>(lrbuf + (size_t)ca[t] * nk * nb + (size_t)k * nb) + cc[t];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + 16 * m + M::row(kk, r);
        T v = part[0][16 * t + 4 * m + r][lane];
#pragma unroll
        for (int j = 1; j < 4; ++j) v += part[j][16 * t + 4 * m + r][lane];
        if (n < nb) out[(size_t)n * 2] = (double)v;
      }
  }
}

template <typename T2, int NA>
__global__ __launch_bounds__(256) void k_adj_lr(const T2* __restrict__ uT, int ldT, const T2* __restrict__ phi,
                                                int ldphi, double2* __restrict__ lrbuf, int nb, int nk, int nx,
                                                const int* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  adj_lr_body<T2, NA>(uT, ldT, phi, ldphi, lrbuf, nb, nk, nx, blockIdx.x, blockIdx.y);
}

// Batched form: problem p owns workgroups [off[p], off[p+1]), numbered band block
// fastest, then k; problems without lrbuf own none.
template <typename T2, int NA>
__global__ __launch_bounds__(256) void k_adj_lr_batch(const ActArgs<T2>* __restrict__ As,
                                                      const int* __restrict__ off, int np) {
  const int p = find_problem(off, np, blockIdx.x);
  const ActArgs<T2>& A = As[p];
  if (skipped(A)) return;
  const int nbb = (A.nb + 63) / 64;
  const int loc = blockIdx.x - off[p];
  adj_lr_body<T2, NA>(A.uT, A.ldT, A.phi, A.ldphi, const_cast<double2*>(A.lrbuf), A.nb, A.nk, A.nx, loc % nbb,
                      loc / nbb);
}

}  // namespace og
