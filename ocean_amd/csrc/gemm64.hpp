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
  static __device__ __forceinline__ int kk_of(int row) { return row & 3; }
  static __device__ __forceinline__ int r_of(int row) { return row >> 2; }
};
template <> struct Mf<float> {
  using V = og_f4;
  static __device__ __forceinline__ V step(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kk, int r) { return 4 * kk + r; }
  static __device__ __forceinline__ int kk_of(int row) { return row >> 2; }
  static __device__ __forceinline__ int r_of(int row) { return row & 3; }
};

// The adjoint for the many-alpha edges (one spin), on the matrix cores:
//   lr(n, k, a) = sum_x conj(uT(k, x, n)) phi'(k, x, a)          (long_range.f90:943-961)
// as a GEMM per k (M = bands, N = the 2 nalpha real columns, K = x) in the precision
// of the stored u (FP64, or FP32 accumulating in FP32 for the mixed form), followed
// in the same workgroup by the rest of OCEAN_xact for its 64 bands:
//   fg(n, k, a)  = sum_p conj(c(n, k, p)) h(p, a)                  OCEAN_multiplet.f90:1943-1998
//   out          = E psi + A (fg - lr) + S psi                     OCEAN_action.f90:337-372, :596-608
// (and c psi - H psi for a GMRES step).  The VALU adjoint reads each phi'(k, x, :) as
// nalpha LDS broadcasts per band row, which bounds it at twelve alphas; here
// phi'(k, x, :) is the B operand.  Workgroup: 64 bands (four M tiles) by all N tiles
// of one k, so every uT element is read once; the four waves take quarters of the x
// range and their tiles are added in LDS in a fixed order.  The epilogue then gives
// wave w the alphas w, w + 4, ... of the block's 64 bands (lane = band): rows of c
// and the element's own psi, E and allow are coalesced band rows, h(p, a) is uniform
// across the wave, and the spin-orbit term reads psi(:, k, n) and S staged in the LDS
// area the tiles leave free.
template <typename T2, int NA, bool DEEP = false>
__device__ __forceinline__ void adj_lr_body(const ActArgs<T2>& A, int bx, int k) {
  using T = typename Scalar<T2>::T;
  using M = Mf<T>;
  using V = typename M::V;
  constexpr int NT = (2 * NA + 15) / 16;
  __shared__ __align__(16) T part[4][NT * 16][64];  // also holds double2 staging after the tiles
  const T2* __restrict__ uT = A.uT;
  const T2* __restrict__ phi = A.phi;
  const int nb = A.nb, nk = A.nk, nx = A.nx, ldT = A.ldT, ldphi = A.ldphi;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 15, kk = lane >> 4;
  const int n0 = bx * 64;
  int cc[NT];
  const T2* yp[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int G = t * 16 + i;
    cc[t] = G & 1;
    yp[t] = phi + (size_t)k * ldphi + min(G >> 1, NA - 1);
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
  // without the long-range term (nhflag) only the epilogue runs: lr = 0
  if (A.flags & F_LR) {
    // D groups of four x in flight per wave (a ring of register buffers): the
    // loads of group g + D - 1 are issued before the products of group g.  The deep
    // ring (four groups in FP32, three in FP64) costs one wave of occupancy per SIMD,
    // so it pays only when the whole grid is resident at once anyway.
    constexpr int D = DEEP ? (sizeof(T) == 4 ? 4 : 3) : 2;
    T2 yv[D][NT], av[D][4];
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
#pragma unroll
    for (int d = 0; d < D - 1; ++d) load(xa + 4 * d, yv[d], av[d]);
    for (int xs = xa; xs < xb; xs += 4 * D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        load(xs + 4 * (d + D - 1), yv[(d + D - 1) % D], av[(d + D - 1) % D]);
        asm volatile("" ::: "memory");  // keep the later groups' loads ahead of this group's products
        step(xs + 4 * d, yv[d], av[d]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[w][16 * t + 4 * m + r][lane] = acc[t][m][r];
  __syncthreads();

  // ---- epilogue: wave w, band n0 + lane, alphas w + 4 j ----
  constexpr int APW = (NA + 3) / 4;
  const int nl = lane, n = n0 + nl, nn = min(n, nb - 1);
  // lr of this lane's band: column G = 2 a + c of tile t = G / 16 sits at lane group
  // i = G % 16; band row nl = 16 m + (the instruction's output row of (kk, r))
  const int mrow = nl >> 4, row = nl & 15;
  const int lkk = M::kk_of(row), lr_ = M::r_of(row);
  double2 res[APW];
#pragma unroll
  for (int j = 0; j < APW; ++j) {
    const int a = min(w + 4 * j, NA - 1);
    double v[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int G = 2 * a + c, t = G >> 4, gi = G & 15;
      const int rr = 16 * t + 4 * mrow + lr_, cl = lkk * 16 + gi;
      T s = part[0][rr][cl];
#pragma unroll
      for (int q = 1; q < 4; ++q) s += part[q][rr][cl];
      v[c] = (double)s;
    }
    res[j] = make_double2(-v[0], -v[1]);  // fg - lr starts at -lr
  }
  const int flags = A.flags;
  const bool back = (flags & F_BACK) != 0;
  const bool so = (flags & F_SO) != 0;
  const int na0 = A.na0;
  __syncthreads();  // every lane has its lr: the tile area is free
  double2* Ps = reinterpret_cast<double2*>(&part[0][0][0]);  // psi(a, k, n0 + q) at Ps[a * 64 + q]
  double2* Ss = Ps + NA * 64;                                 // S (conjugated backwards), na0 x na0
  static_assert(sizeof(part) >= sizeof(double2) * (NA * 64 + NA * NA), "spin-orbit staging exceeds the tile area");
  if (so) {
    for (int e = threadIdx.x; e < NA * 64; e += blockDim.x) {
      const int a = e >> 6, q = e & 63;
      Ps[e] = A.psi[((size_t)a * nk + k) * nb + min(n0 + q, nb - 1)];
    }
    for (int e = threadIdx.x; e < na0 * na0; e += blockDim.x) {
      double2 v = A.S[e];
      if (back) v.y = -v.y;
      Ss[e] = v;
    }
  }
  // the element's own operands, and the projector rows (one spin: c[0][k][p][n])
  double2 pv[APW], ev[APW];
  double am[APW];
#pragma unroll
  for (int j = 0; j < APW; ++j) {
    const int a = min(w + 4 * j, NA - 1);
    const size_t off = ((size_t)a * nk + k) * nb + nn;
    pv[j] = A.psi[off];
    ev[j] = A.energy[off];
    am[j] = A.allow[off];
  }
  if (flags & F_FG) {
    const int P = A.P;
    const double2* __restrict__ cb = A.cproj + (size_t)k * P * nb + nn;
    const double2* __restrict__ hm = A.hmult;
    constexpr int kC = 4;
    for (int p0 = 0; p0 < P; p0 += kC) {
      double2 cv[kC];
#pragma unroll
      for (int q = 0; q < kC; ++q) cv[q] = cb[(size_t)min(p0 + q, P - 1) * nb];
#pragma unroll
      for (int q = 0; q < kC; ++q)
        if (p0 + q < P) {
#pragma unroll
          for (int j = 0; j < APW; ++j)
            if (w + 4 * j < NA) cfmac(res[j], cv[q], hm[(w + 4 * j) * P + p0 + q]);
        }
    }
  }
  if (so) __syncthreads();
  if (n >= nb) return;
#pragma unroll
  for (int j = 0; j < APW; ++j) {
    const int a = w + 4 * j;
    if (a >= NA) break;
    double2 inter = res[j];
    inter.x *= am[j];
    inter.y *= am[j];
    if (so) {
      const int ar = a % na0, a0 = a - ar;
      for (int b = 0; b < na0; ++b) cfma(inter, Ss[ar * na0 + b], Ps[(a0 + b) * 64 + nl]);
    }
    double2 r = inter;
    if (flags & F_E0) {
      double2 e = ev[j];
      if (back) e.y = -e.y;
      r.x = e.x * pv[j].x - e.y * pv[j].y + inter.x;
      r.y = e.x * pv[j].y + e.y * pv[j].x + inter.y;
    }
    A.out[((size_t)a * nk + k) * nb + n] = axmy_epilogue(A, pv[j], r);
  }
}

template <typename T2, int NA, bool DEEP>
__global__ __launch_bounds__(256) void k_adj_lr(const ActArgs<T2> A) {
  if (skipped(A)) return;
  // consecutive workgroups take the band blocks of one k, then the next k
  const int nbb = (A.nb + 63) / 64;
  adj_lr_body<T2, NA, DEEP>(A, blockIdx.x % nbb, blockIdx.x / nbb);
}

// Batched form: problem p owns workgroups [off[p], off[p+1]), numbered band block
// fastest, then k.
template <typename T2, int NA>
__global__ __launch_bounds__(256) void k_adj_lr_batch(const ActArgs<T2>* __restrict__ As,
                                                      const int* __restrict__ off, int np) {
  const int p = find_problem(off, np, blockIdx.x);
  const ActArgs<T2> A = As[p];
  if (skipped(A)) return;
  const int nbb = (A.nb + 63) / 64;
  const int loc = blockIdx.x - off[p];
  adj_lr_body<T2, NA>(A, loc % nbb, loc / nbb);
}

}  // namespace og
