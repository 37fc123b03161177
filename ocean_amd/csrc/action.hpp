// Kernels of the core-level BSE action H.psi (OCEAN_xact, OCEAN2/OCEAN_action.f90:286-618).
//
// Device layouts (chosen for gfx950, not the reference's):
//   u   [s][k][n][x]  x fastest: the projection streams rows over x with lanes over x,
//                     the adjoint streams the same rows and reduces over x per wave.
//                     One copy serves both passes, so the second pass can hit the
//                     256 MB Infinity Cache when u fits (FP32 form).
//   phi [k][x][a]     projection output / FFT in-out (columns c = x*na + a)
//   W   [x][k]        = the reference's W(nk, nx) column-major
//   psi [a][k][n]     interleaved complex, = OCEAN_vector %r/%i order
//   c   [s][k][p][n]  compact projector index p = (l, m, nu)
#pragma once
#include "common.hpp"

namespace og {

enum ActFlags : int {
  F_LR = 1, F_FG = 2, F_SO = 4, F_E0 = 8, F_BACK = 16,
};

// GMRES control state on the device: the solve of one energy enqueues several
// iterations ahead and every kernel of them returns at once when `done` is set, so
// the host reads the state back once per block of iterations, not per iteration.
struct GmCtl {
  int done;     // 0 running, 1 converged, 2 least-squares pivot failure
  int conv_it;  // the iteration (1..depth of its cycle) at which it converged
  int info;     // zhptrf-style pivot index of a failure
  int pad1;     // (keeps the doubles 16-byte aligned)
  double gval, rval, ival, pad2;  // |g|^2, Re / Im <rhs|x> of the last residual
  double rxr, rxi, pad3, pad4;    // <rhs|x_b> of the materialised x (x = x_b + sum_i C_i U_i)
};

// The finish of a GMRES residual step: the residual kernel wrote one partial triple
// per workgroup; the finish sums them in a fixed order and, by mode:
//   kResX   (x materialised, C = 0): v = (|g|^2, <rhs|x>); <rhs|x> becomes <rhs|x_b>;
//   kResU   v = (|g|^2, <rhs|U_slot>) (slot < 0: no new vector); <rhs|x> is
//           <rhs|x_b> + sum_{i<m} C_i <rhs|U_i>;
//   kDotU   v = (0, <rhs|U_slot>): the dot only.
// kResX and kResU store |g|^2 and <rhs|x> and, with check, apply the reference's
// stopping test (OCEAN_gmres.f90:845-852): |g|^2 < ffff and (-|g|^2 / Im<rhs|x> < 0.001
// or |g|^2 < ffff / 10).  It runs as its own one-workgroup kernel or, inside a
// restart cycle, in workgroup 0 of the next action's projection (c == nullptr: none).
enum { kResX = 0, kResU = 1, kDotU = 2 };
struct GmFin {
  GmCtl* c = nullptr;
  const double* part = nullptr;
  const double2* C = nullptr;
  double2* rU = nullptr;
  double ffff = 0.0;
  int nblk = 0, it = 0, check = 0, mode = 0, m = 0, slot = -1;
};

// One workgroup of any size (one wave or more).
__device__ __noinline__ void gm_finish(const GmFin F) {
  GmCtl* c = F.c;
  if (c->done) return;
  __shared__ double sv[5];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int task = w; task < 4; task += nw) {
    if (task < 3) {
      double acc = 0.0;
      for (int b = lane; b < F.nblk; b += 64) acc += F.part[3 * b + task];
      acc = wave_sum(acc);
      if (lane == 0) sv[task] = acc;
    } else if (F.mode == kResU) {
      // sum_{i<m} C_i <rhs|U_i>, lane-strided then a butterfly (fixed order)
      double ar = 0.0, ai = 0.0;
      for (int i = lane; i < F.m; i += 64) {
        const double2 ci = F.C[i], u = F.rU[i];
        ar += ci.x * u.x - ci.y * u.y;
        ai += ci.x * u.y + ci.y * u.x;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) { sv[3] = ar; sv[4] = ai; }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (F.slot >= 0) F.rU[F.slot] = make_double2(sv[1], sv[2]);
    if (F.mode != kDotU) {
      double rr, ri;
      if (F.mode == kResX) {
        rr = sv[1]; ri = sv[2];
        c->rxr = rr; c->rxi = ri;
      } else {
        rr = c->rxr + sv[3]; ri = c->rxi + sv[4];
      }
      c->gval = sv[0];
      c->rval = rr;
      c->ival = ri;
      if (F.check && sv[0] < F.ffff) {
        const double rel = -sv[0] / ri;
        if (rel < 0.001 || sv[0] < F.ffff / 10.0) {
          c->done = 1;
          c->conv_it = F.it;
        }
      }
    }
  }
  __syncthreads();
}

// Everything one problem's action kernels read: pointers into its device state,
// its sizes and its launch geometry.  Passed by value to the single-problem
// kernels; the batched kernels read an array of them, one per problem.
template <typename T2>
struct ActArgs {
  using T = typename Scalar<T2>::T;
  const T2* u;                // [s][k][n][x], row stride ldu
  const T2* uT;               // [s][k][x][n], k-slab stride ldT
  T2* phi;                    // [k][x][a], row stride ldphi
  const T* W;                 // [x][k]
  const double2* psi;         // in  [a][k][n]
  double2* out;               // out [a][k][n]
  const double2* energy;
  const double* allow;
  const double2* cproj;       // [s][k][p][n]
  double2* apart;             // [a*P+p][k]
  double2* amp;               // [a*P+p]
  const double2* Kmat;        // (na P)^2
  double2* hmult;             // [a*P+p]
  int* mcount;                // reduction arrival counter (rests at 0)
  const double2* S;           // [a][b] over the na0 alphas of one right-hand side
  double mul;                 // inter_scale / (N_k celvol)
  int nb, nk, nx, na, ns, P, nj;
  int na0;                    // alphas per right-hand side: alpha a of a vector belongs to
                              // right-hand side a / na0 (na = na0 for one right-hand side)
  int ldu, ldT, ldphi;
  int flags;
  int px;                     // projection x-tiles (blockDim.x x-points each)
  int bstage;                 // FP64 matrix-core projection: A psi of the k staged in LDS
  int G, ncol, nfft, nred;    // transform: columns per strip, strips, reduction workgroups
  int K1, K2, K3;
  int XC, JR;                 // adjoint: phi' chunk, alphas per reduction round
  const T2* phs;              // OBF form: supercell phase [x][k] applied around the transform
  const int* skip;            // nullable: when *skip != 0 every kernel of the action returns at once
                              // (a solver that has converged on the device drops the work it
                              // enqueued ahead)
  int axmy;                   // 1: the adjoint writes c psi - H psi instead of H psi
  double2 axmy_c;             // c = E + i gamma of a GMRES step (OCEAN_psi_axmy)
  GmFin fin;                  // fin.c != nullptr: workgroup 0 of the projection first runs
                              // the previous GMRES residual step's finish
};

// out = c psi - r in OCEAN_psi_axmy's operation order (OCEAN2/OCEAN_psi.f90:1468-1562)
template <typename T2>
__device__ __forceinline__ double2 axmy_epilogue(const ActArgs<T2>& A, double2 pv, double2 r) {
  if (!A.axmy) return r;
  const double br = A.axmy_c.x * pv.x - r.x, bi = A.axmy_c.x * pv.y - r.y;
  return make_double2(br - A.axmy_c.y * pv.y, bi + A.axmy_c.y * pv.x);
}

template <typename T2>
__device__ __forceinline__ bool skipped(const ActArgs<T2>& A) {
  return A.skip != nullptr && *A.skip != 0;
}

// Problem p of a batched launch owns workgroups [off[p], off[p+1]).
__device__ __forceinline__ int find_problem(const int* __restrict__ off, int np, int b) {
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void mproj_k(const double2* __restrict__ c, const double2* __restrict__ psi,
                                        const double* __restrict__ allow, double2* __restrict__ apart,
                                        int nb, int nk, int na, int ns, int P, int k, int s, int c0,
                                        int cstep);

// ---------------------------------------------------------------------------
// Projection: phi(k,x,a) = sum_n u(n,k,x) (A psi)(n,k,a)
//   long_range.f90:897-906 (no conjugation of u).  Grid (ceil(nx/256), nk, ns).
// With fg set, the slab workgroups of each k also form its multiplet projection
// partials (a 16-KB read of c beside the 512-KB u slab), so the multiplet term needs
// no launch or stream of its own.
// ---------------------------------------------------------------------------
// FP32 projection on the matrix cores, for many alphas (L2,3 edges), where the FMA
// count per byte of u would otherwise bound the streaming loop.  Per (k, s) it is the
// real GEMM phi[x][c] = sum_{n, r} U[x][(n, r)] B[(n, r)][c] with the complex parts
// unrolled: r = re/im part of u(n, x), columns c < acnt the real parts of phi(x, a),
// c - acnt the imaginary parts, so B[(n, re)] = (psi_r, psi_i), B[(n, im)] =
// (-psi_i, psi_r) of A psi(n, a).  v_mfma_f32_16x16x4f32: lane l supplies
// U[x(l % 16)][k = l / 16] and B[k = l / 16][c = l % 16] and receives D[4 (l / 16) + r]
// [l % 16] in register r.  Each lane loads 16 bytes, u(n, x) and u(n, x + 1), which
// feed two tiles (even and odd x) of 16 x points; one wave covers 32 x points, four
// band rows per step, and all column tiles (two for acnt <= 12, three for 24).
typedef float f32x4 __attribute__((ext_vector_type(4)));
// floats per half B row (the part multiplying re u, then the part multiplying im u)
// and per row: halves of 32 (48 for three column tiles) and 16 floats of pad, so
// the four band rows one instruction reads start 16 or 48 banks apart (no conflicts)
template <int NA> constexpr int proj_bhalf() { return NA > 12 ? 48 : 32; }
template <int NA> constexpr int proj_brow() { return 2 * proj_bhalf<NA>() + 16; }
template <int NA>
__device__ __forceinline__ void proj_mfma_body(const ActArgs<float2>& A, int bx, int k, int s) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int kBRow = proj_brow<NA>(), HW = proj_bhalf<NA>();
  float* bt = reinterpret_cast<float*>(smem);  // [nb4][kBRow]: B rows (n, re) then (n, im)
  const int nb = A.nb, nk = A.nk, na = A.na, ns = A.ns;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  const int nb4 = (nb + 3) & ~3;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xb = (bx * (int)(blockDim.x >> 6) + wave) * 32;
  const bool xon = xb < A.nx;  // waves past nx stay for the multiplet partials after the slab
  const int m16 = lane & 15, kq = lane >> 4;
  // x + 1 <= xb + 31 < nx + 31 < ldu: the pair stays inside the padded row
  const float2* ub = A.u + ((size_t)(s * nk + k) * nb) * A.ldu + min(xb, A.nx - 1) + 2 * m16;
  // batches of kS four-row steps, double-buffered: the loads of batch b + 1 are
  // issued before the products of batch b, so each wave keeps HBM busy while its
  // matrix-core work runs.  The first batch is issued before the B staging below,
  // so the slab stream starts while psi and allow are read.
  // (four at 24 alphas: 84.5 -> 66.8 us at nb 64, 111 instead of 133 VGPRs, so two
  // 512-thread workgroups fit a CU)
  constexpr int kS = NA > 12 ? 4 : 8;
  const int nfull = nb / (4 * kS);
  float4 qa[kS], qb[kS];
  auto load_batch = [&](float4* q, int b) {
#pragma unroll
    for (int i = 0; i < kS; ++i) q[i] = ld_stream4(ub + (size_t)(b * 4 * kS + 4 * i + kq) * A.ldu);
  };
  if (xon && nfull > 0) load_batch(qa, 0);
  // each (n, alpha) element is read once (kF loads in flight per thread) and fills
  // its four B entries; columns past 2 acnt and rows past nb are zero.  (Bands
  // fastest over the threads, which makes the reads contiguous, measured slower:
  // its LDS writes land four banks apart, 52.5 -> 61.4 us at nb 64 and 12 alphas.)
  const int nel = nb4 * acnt;
  constexpr int kF = 8;
  for (int i0 = threadIdx.x; i0 < nel; i0 += kF * blockDim.x) {
    double2 pv[kF];
    double mv[kF];
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int i = i0 + f * blockDim.x;
      const int n = i / acnt, c = i - n * acnt;
      pv[f] = make_double2(0.0, 0.0);
      mv[f] = 0.0;
      if (i < nel && n < nb) {
        const size_t off = ((size_t)(s + c * astep) * nk + k) * nb + n;
        pv[f] = A.psi[off];
        mv[f] = A.allow[off];
      }
    }
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int i = i0 + f * blockDim.x;
      if (i < nel) {
        const int n = i / acnt, c = i - n * acnt;
        float* row = bt + n * kBRow;
        const float pr = (float)(pv[f].x * mv[f]), pim = (float)(pv[f].y * mv[f]);
        row[c] = pr;
        row[HW + c] = -pim;
        row[acnt + c] = pim;
        row[HW + acnt + c] = pr;
      }
    }
  }
  const int nz = HW - 2 * acnt;
  for (int i = threadIdx.x; i < nb4 * nz; i += blockDim.x) {
    const int n = i / nz, c = 2 * acnt + (i - n * nz);
    bt[n * kBRow + c] = 0.f;
    bt[n * kBRow + HW + c] = 0.f;
  }
  __syncthreads();
  if (xon) {
    constexpr int NT = (2 * NA + 15) / 16;  // column tiles (2 acnt <= 16 fits one)
    f32x4 ce[NT], co[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) { ce[t] = f32x4{0, 0, 0, 0}; co[t] = f32x4{0, 0, 0, 0}; }
    // one four-row step: tile t takes re(u) x B(n, re) + im(u) x B(n, im); the four
    // chains are interleaved so that dependent products are NT * 2 issues apart
    auto step = [&](const float4& q, const float* br) {
      float b[NT], c[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) { b[t] = br[16 * t]; c[t] = br[HW + 16 * t]; }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        ce[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(q.x, b[t], ce[t], 0, 0, 0);
        co[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(q.z, b[t], co[t], 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        ce[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(q.y, c[t], ce[t], 0, 0, 0);
        co[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(q.w, c[t], co[t], 0, 0, 0);
      }
    };
    auto use_batch = [&](const float4* q, int b) {
#pragma unroll
      for (int i = 0; i < kS; ++i) step(q[i], bt + (b * 4 * kS + 4 * i + kq) * kBRow + m16);
    };
    for (int b = 0; b < nfull; b += 2) {
      if (b + 1 < nfull) load_batch(qb, b + 1);
      __builtin_amdgcn_sched_barrier(0);
      use_batch(qa, b);
      if (b + 1 < nfull) {
        if (b + 2 < nfull) load_batch(qa, b + 2);
        __builtin_amdgcn_sched_barrier(0);
        use_batch(qb, b + 1);
      }
    }
    // remaining rows, four at a time; rows past nb read zero B rows (nb4 covers them)
    for (int n0 = nfull * 4 * kS; n0 < nb; n0 += 4) {
      const float4 q = ld_stream4(ub + (size_t)min(n0 + kq, nb - 1) * A.ldu);
      step(q, bt + (n0 + kq) * kBRow + m16);
    }
    // D[4 kq + r][m16]: x = xb + 2 (4 kq + r) (+1 for the odd tile), column 16 t + m16
    float* ph = reinterpret_cast<float*>(A.phi + (size_t)k * A.ldphi);
    auto put = [&](const f32x4& d, int t, int odd) {
      const int c = 16 * t + m16;
      if (c >= 2 * acnt) return;
      const int j = c < acnt ? c : c - acnt, part = c < acnt ? 0 : 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = xb + 2 * (4 * kq + r) + odd;
        if (x < A.nx) ph[2 * ((size_t)x * na + s + j * astep) + part] = d[r];
      }
    };
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      put(ce[t], t, 0);
      put(co[t], t, 1);
    }
  }
}

template <typename T2, int NA>
__device__ __forceinline__ void proj_body(const ActArgs<T2>& A, int bx, int k, int s) {
  if constexpr (sizeof(T2) == 8 && NA >= 6) {
    proj_mfma_body<NA>(A, bx, k, s);
    return;
  }
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  T2* ps = reinterpret_cast<T2*>(smem);  // [acnt][nb]
  const int nb = A.nb, nk = A.nk, na = A.na, ns = A.ns;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  // 16 bytes per lane and row: one x point in FP64, two adjacent ones in FP32 (the
  // row stride ldu is even, so the pair is 16-byte aligned)
  constexpr int VX = sizeof(T) == 8 ? 1 : 2;
  const int x = (bx * blockDim.x + threadIdx.x) * VX;
  const bool on = x < A.nx;  // threads past nx stay for the multiplet partials after the slab
  const T2* ub = A.u + ((size_t)(s * nk + k) * nb) * A.ldu + min(x, (A.nx - 1) & ~(VX - 1));
  // Batches of kB band rows: all kB loads are issued before the first use, so
  // each wave keeps kB row segments in flight (8 KB per wave in FP64; FP32 runs
  // half as many threads per workgroup and keeps 16 KB per wave in flight).
  // Full batches carry no per-row predicate: with one, the compiler sinks the FP32
  // loads into per-row branches and waits on each row in turn.
  constexpr int kB = VX == 1 ? 8 : 16;
  auto load_row = [&](int n, T2* dst) {
    const T2* row = ub + (size_t)n * A.ldu;
    if constexpr (VX == 1) {
      dst[0] = ld_stream(row);
    } else {
      const float4 q = ld_stream4(reinterpret_cast<const float2*>(row));
      dst[0] = cmake<T2>(q.x, q.y);
      dst[VX - 1] = cmake<T2>(q.z, q.w);
    }
  };
  // the first batch is in flight while A psi is staged
  T2 uv0[kB][VX];
  const bool pre = on && nb >= kB;
  if (pre) {
#pragma unroll
    for (int i = 0; i < kB; ++i) load_row(i, uv0[i]);
  }
  // A psi of this k in LDS, with up to kF loads in flight per thread
  constexpr int kF = 4;
  for (int i0 = threadIdx.x; i0 < acnt * nb; i0 += kF * blockDim.x) {
    double2 pv[kF];
    double mv[kF];
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int i = min(i0 + f * (int)blockDim.x, acnt * nb - 1);
      const int j = i / nb, n = i - j * nb;
      const size_t off = ((size_t)(s + j * astep) * nk + k) * nb + n;
      pv[f] = A.psi[off];
      mv[f] = A.allow[off];
    }
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int i = i0 + f * blockDim.x;
      if (i < acnt * nb) ps[i] = cmake<T2>((T)(pv[f].x * mv[f]), (T)(pv[f].y * mv[f]));
    }
  }
  __syncthreads();
  if (on) {
    T2 acc[VX][NA];
#pragma unroll
    for (int v = 0; v < VX; ++v)
#pragma unroll
      for (int j = 0; j < NA; ++j) acc[v][j] = cmake<T2>(0, 0);
    auto use_row = [&](int n, const T2* uvr) {
#pragma unroll
      for (int j = 0; j < NA; ++j) {
        const T2 pv = ps[j * nb + n];
#pragma unroll
        for (int v = 0; v < VX; ++v) cfma(acc[v][j], uvr[v], pv);
      }
    };
    int n0 = 0;
    if (pre) {
#pragma unroll
      for (int i = 0; i < kB; ++i) use_row(i, uv0[i]);
      n0 = kB;
    }
    for (; n0 + kB <= nb; n0 += kB) {
      T2 uv[kB][VX];
#pragma unroll
      for (int i = 0; i < kB; ++i) load_row(n0 + i, uv[i]);
      // keep the kB loads together ahead of the arithmetic
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < kB; ++i) use_row(n0 + i, uv[i]);
    }
    for (; n0 < nb; ++n0) {
      T2 uv[VX];
      load_row(n0, uv);
      use_row(n0, uv);
    }
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      if (x + v < A.nx) {
        T2* out = A.phi + (size_t)k * A.ldphi + (size_t)(x + v) * na;
#pragma unroll
        for (int j = 0; j < NA; ++j)
          if (j < acnt) out[s + j * astep] = acc[v][j];
      }
    }
  }
}

// M tiles (16 x points each) per wave of the FP64 matrix-core projection
constexpr int kProjMT64 = 4;

// FP64 projection for many alphas on the matrix cores: per k the GEMM
// phi[x][(a, c)] = sum_n u(n, k, x) (A psi)(n, k, a) with the complex parts unrolled
// along K (v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16], B[l / 16][l % 16]
// and receives D[l / 16 + 4 r][l % 16]).  Each wave covers 16 MT x points (MT M tiles)
// and all 2 nalpha real columns of its spin, so every u element is read once; A psi
// is formed per lane from psi and allow (L1/L2 hits).  Single-buffered fragments
// keep four waves per SIMD, which hide the u stream's latency; double buffering
// needed 166 VGPRs (three waves) and was slower.
template <int NA, int MT>
__device__ __forceinline__ void proj_mf64_body(const ActArgs<double2>& A, int bx, int k, int s) {
  constexpr int NT = (2 * NA + 15) / 16;
  const int nb = A.nb, nk = A.nk, nx = A.nx, na = A.na, ns = A.ns;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, kk = lane >> 4;
  const int x0 = bx * (int)blockDim.x * MT / 4 + w * 16 * MT;
  int ja[NT], jc[NT], jl[NT];
  bool jok[NT];
  size_t boff[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int G = t * 16 + i;
    jc[t] = G & 1;
    jok[t] = (G >> 1) < acnt;
    jl[t] = min(G >> 1, acnt - 1);
    ja[t] = s + jl[t] * astep;
    boff[t] = ((size_t)ja[t] * nk + k) * nb;
  }
  // A psi of this k staged once per workgroup (A.bstage; the host sets it when nb acnt
  // complex values fit the LDS budget): read with bands fastest over the threads, so
  // every load is a contiguous band row of one alpha, where each wave otherwise loads
  // 2 NT scattered (alpha, band) pairs per lane and step from L2
  extern __shared__ __align__(16) unsigned char smem[];
  double2* Bs = reinterpret_cast<double2*>(smem);  // A psi(n, j) at Bs[n * acnt + j]
  const bool staged = NA <= 12 && A.bstage != 0;  // not built at 24 alphas (slower there)
  if (staged) {
    const int nel = nb * acnt;
    constexpr int kF = 8;
    for (int i0 = threadIdx.x; i0 < nel; i0 += kF * blockDim.x) {
      double2 pv[kF];
      double mv[kF];
#pragma unroll
      for (int f = 0; f < kF; ++f) {
        const int e = min(i0 + f * (int)blockDim.x, nel - 1);
        const int j = e / nb, n = e - j * nb;
        const size_t off = ((size_t)(s + j * astep) * nk + k) * nb + n;
        pv[f] = A.psi[off];
        mv[f] = A.allow[off];
      }
#pragma unroll
      for (int f = 0; f < kF; ++f) {
        const int e = i0 + f * (int)blockDim.x;
        if (e < nel) {
          const int j = e / nb, n = e - j * nb;
          Bs[n * acnt + j] = make_double2(pv[f].x * mv[f], pv[f].y * mv[f]);
        }
      }
    }
    __syncthreads();
  }
  const double2* __restrict__ up = A.u + ((size_t)s * nk + k) * nb * A.ldu;
  int xr[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) xr[m] = min(x0 + 16 * m + i, nx - 1);
  og_d4 acc[NT][MT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[t][m] = og_d4{0.0, 0.0, 0.0, 0.0};
  double2 bA[NT], aA[MT];
  auto load = [&](int n0, double2 (&bn)[NT], double2 (&an)[MT]) {
    const int n = min(n0 + kk, nb - 1);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (staged) {
        bn[t] = Bs[n * acnt + jl[t]];
      } else {
        double2 pv = A.psi[boff[t] + n];
        const double m = A.allow[boff[t] + n];
        bn[t] = make_double2(pv.x * m, pv.y * m);
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) an[m] = ld_stream(up + (size_t)n * A.ldu + xr[m]);
  };
  // C = A B: re: ar br - ai bi, im: ar bi + ai br
  auto step = [&](int n0, const double2 (&bv)[NT], const double2 (&an)[MT]) {
    const bool ok = n0 + kk < nb;
    double2 av[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) av[m] = ok ? an[m] : make_double2(0.0, 0.0);
    double b0[NT], b1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      b0[t] = jc[t] ? bv[t].y : bv[t].x;
      b1[t] = jc[t] ? bv[t].x : -bv[t].y;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[t][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m].x, b0[t], acc[t][m], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[t][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m].y, b1[t], acc[t][m], 0, 0, 0);
  };
  for (int n0 = 0; n0 < nb; n0 += 4) {
    load(n0, bA, aA);
    step(n0, bA, aA);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (!jok[t]) continue;
    double* __restrict__ out = reinterpret_cast<double*>(A.phi + (size_t)k * A.ldphi + ja[t]) + jc[t];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = x0 + 16 * m + kk + 4 * r;
        if (x < nx) out[(size_t)x * na * 2] = acc[t][m][r];
      }
  }
}

// Workgroup b of a projection launch: slab tile bx of (k, s).  With the multiplet
// on, each tile then forms its share (chunks bx, bx + px, ...) of that k's multiplet
// partials (16 KB of c per k), after its slab so that the stream starts at once.
template <typename T2, int NA>
__device__ __forceinline__ void proj_dispatch(const ActArgs<T2>& A, int b) {
  const int r = b / A.px, bx = b - r * A.px, k = r % A.nk, s = r / A.nk;
  if constexpr (sizeof(T2) == 16 && NA >= 6)
    proj_mf64_body<NA, kProjMT64>(A, bx, k, s);
  else
    proj_body<T2, NA>(A, bx, k, s);
  if (A.flags & F_FG) mproj_k(A.cproj, A.psi, A.allow, A.apart, A.nb, A.nk, A.na, A.ns, A.P, k, s, bx, A.px);
}

// One problem: a 1-D grid of px * nk * ns slab tiles.  Workgroups are dispatched
// round-robin over the 8 XCDs; the remap gives each XCD a contiguous range of tiles,
// so the px tiles of one k (which stage the same A psi) share an L2.
// The FP64 matrix-core form at twelve alphas keeps four waves per SIMD (at most 128
// VGPRs); at 24 alphas that cap spilled 96 bytes per lane, so it runs at three.
template <typename T2, int NA>
constexpr int kProjWaves = (sizeof(T2) == 16 && NA >= 6 && NA <= 12) ? 4 : 1;

template <typename T2, int NA>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kProjWaves<T2, NA>))) void k_proj(
    const ActArgs<T2> A) {
  if (skipped(A)) return;
  if (A.fin.c != nullptr && blockIdx.x == 0) gm_finish(A.fin);
  int b = blockIdx.x;
  const int n = gridDim.x;
  if ((n & 7) == 0) b = (b & 7) * (n >> 3) + (b >> 3);
  proj_dispatch<T2, NA>(A, b);
}

// Many problems: workgroup b belongs to the problem whose range holds it.
template <typename T2, int NA>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kProjWaves<T2, NA>))) void k_proj_batch(
    const ActArgs<T2>* __restrict__ As, const int* __restrict__ off, int np) {
  const int p = find_problem(off, np, blockIdx.x);
  const ActArgs<T2> A = As[p];
  if (skipped(A)) return;
  const int local = blockIdx.x - off[p];
  proj_dispatch<T2, NA>(A, local);
}

// ---------------------------------------------------------------------------
// Batched 3-D DFT over the k-mesh, staged in LDS, fused with the W multiply:
//   phi' = (1/N_k) F_- [ W(:,x) . F_+ phi ]    long_range.f90:920-939
// F_+ = FFTW_BACKWARD (exp +2 pi i), F_- = FFTW_FORWARD, k3 fastest
// (Modules/FFT_wrapper.f90:67-72, :185-215).  One pencil per thread; edge classes
// LMAX = 8, 16 and 24 (radix 2 for powers of two up to 16, direct sums otherwise).
// ---------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 1; b < L; b <<= 1) { r = (r << 1) | (i & 1); i >>= 1; }
  return r;
}

// LDS index of tile element e with one pad slot every 8 elements: pencils along
// the fastest k dimension start 8G elements apart, which without the pad would
// land on the same LDS banks.
__device__ __forceinline__ int lidx(int e) { return e + (e >> 3); }

// One pass of length-L pencils along the k dimension of stride S, over the G
// columns of the strip: out[m] = sum_j v[j] w^(j m), w = exp(sign 2 pi i / L),
// tw[m] = w^m from LDS.  Power-of-two L: radix-2 decimation in time on data
// loaded in bit-reversed order; otherwise the direct sum.  ld(k, g) / st(k, g, v)
// move one element, so the first pass reads HBM straight into the pencil
// registers and the last writes HBM from them.
// cos and sin of 2 pi m / 16, correctly rounded (c1 = cos(pi/8), s1 = sin(pi/8),
// r = sqrt(2)/2): the power-of-two twiddles
#define OG_C1 0.92387953251128674
#define OG_S1 0.38268343236508978
#define OG_R2 0.70710678118654752
__device__ constexpr double kCos16[16] = {1.0, OG_C1, OG_R2, OG_S1, 0.0, -OG_S1, -OG_R2, -OG_C1,
                                          -1.0, -OG_C1, -OG_R2, -OG_S1, 0.0, OG_S1, OG_R2, OG_C1};
__device__ constexpr double kSin16[16] = {0.0, OG_S1, OG_R2, OG_C1, 1.0, OG_C1, OG_R2, OG_S1,
                                          0.0, -OG_S1, -OG_R2, -OG_C1, -1.0, -OG_C1, -OG_R2, -OG_S1};
#undef OG_C1
#undef OG_S1
#undef OG_R2

// v * w^m, w = exp(SIGN 2 pi i / L), L a power of two <= 16.  After unrolling m is
// a constant, so the multiplications by 1, -1 and +-i fold away and the others
// use literal twiddles.
template <int SIGN, int L, typename T2>
__device__ __forceinline__ T2 tw_mul(int m, T2 v) {
  using T = typename Scalar<T2>::T;
  const int idx = (m * (16 / L)) & 15;
  if (idx == 0) return v;
  if (idx == 8) return cmake<T2>(-v.x, -v.y);
  if (idx == 4) return SIGN > 0 ? cmake<T2>(-v.y, v.x) : cmake<T2>(v.y, -v.x);
  if (idx == 12) return SIGN > 0 ? cmake<T2>(v.y, -v.x) : cmake<T2>(-v.y, v.x);
  const T c = (T)kCos16[idx], sn = (T)(SIGN * kSin16[idx]);
  return cmake<T2>(c * v.x - sn * v.y, c * v.y + sn * v.x);
}

template <typename T2, int L, int SIGN, typename Ld, typename St>
__device__ __forceinline__ void dft_pass(const T2* __restrict__ tw, int G, int nk, int S, Ld ld, St st) {
  const int npen = (G * nk) / L;
  for (int q = threadIdx.x; q < npen; q += blockDim.x) {
    const int g = q % G, r = q / G;
    const int base = (r / S) * (S * L) + (r % S);
    T2 v[L];
    if constexpr ((L & (L - 1)) == 0) {
#pragma unroll
      for (int j = 0; j < L; ++j) v[bitrev<L>(j)] = ld(base + j * S, g);
#pragma unroll
      for (int len = 2; len <= L; len <<= 1) {
        const int half = len / 2, step = L / len;
#pragma unroll
        for (int i = 0; i < L; i += len) {
#pragma unroll
          for (int j = 0; j < half; ++j) {
            const T2 t = tw_mul<SIGN, L>(j * step, v[i + j + half]);
            const T2 u = v[i + j];
            v[i + j] = cmake<T2>(u.x + t.x, u.y + t.y);
            v[i + j + half] = cmake<T2>(u.x - t.x, u.y - t.y);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < L; ++m) st(base + m * S, g, v[m]);
    } else {
#pragma unroll
      for (int j = 0; j < L; ++j) v[j] = ld(base + j * S, g);
#pragma unroll
      for (int m = 0; m < L; ++m) {
        T2 acc = cmake<T2>(0, 0);
#pragma unroll
        for (int j = 0; j < L; ++j) cfma(acc, v[j], tw[(j * m) % L]);
        st(base + m * S, g, acc);
      }
    }
  }
}

// In-register radix-2 DFT of L (a power of two) values given in bit-reversed order;
// the result is in natural order.
template <int SIGN, int L, typename T2>
__device__ __forceinline__ void fft_regs(T2 (&v)[L]) {
#pragma unroll
  for (int len = 2; len <= L; len <<= 1) {
    const int half = len / 2, step = L / len;
#pragma unroll
    for (int i = 0; i < L; i += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        const T2 t = tw_mul<SIGN, L>(j * step, v[i + j + half]);
        const T2 u = v[i + j];
        v[i + j] = cmake<T2>(u.x + t.x, u.y + t.y);
        v[i + j + half] = cmake<T2>(u.x - t.x, u.y - t.y);
      }
    }
  }
}

// The last forward pass, the W product and the first backward pass fused along one
// k dimension: each thread keeps its pencil in registers through F_+, W and F_-, so
// the W product costs no LDS round trip (the 3-D transform is separable: the
// backward passes may run in the reverse dimension order).
template <typename T2, int L, typename Ld, typename St, typename Wf>
__device__ __forceinline__ void dft_pass_w(const T2* __restrict__ twp, const T2* __restrict__ twm, int G, int nk,
                                           int S, Ld ld, St st, Wf wf) {
  using T = typename Scalar<T2>::T;
  const int npen = (G * nk) / L;
  for (int q = threadIdx.x; q < npen; q += blockDim.x) {
    const int g = q % G, r = q / G;
    const int base = (r / S) * (S * L) + (r % S);
    T2 v[L], u[L];
    if constexpr ((L & (L - 1)) == 0) {
#pragma unroll
      for (int j = 0; j < L; ++j) v[bitrev<L>(j)] = ld(base + j * S, g);
      fft_regs<1, L>(v);
#pragma unroll
      for (int m = 0; m < L; ++m) {
        const T w = wf(base + m * S, g);
        u[bitrev<L>(m)] = cmake<T2>(v[m].x * w, v[m].y * w);
      }
      fft_regs<-1, L>(u);
#pragma unroll
      for (int m = 0; m < L; ++m) st(base + m * S, g, u[m]);
    } else {
#pragma unroll
      for (int j = 0; j < L; ++j) v[j] = ld(base + j * S, g);
#pragma unroll
      for (int m = 0; m < L; ++m) {
        T2 acc = cmake<T2>(0, 0);
#pragma unroll
        for (int j = 0; j < L; ++j) cfma(acc, v[j], twp[(j * m) % L]);
        const T w = wf(base + m * S, g);
        u[m] = cmake<T2>(acc.x * w, acc.y * w);
      }
#pragma unroll
      for (int m = 0; m < L; ++m) {
        T2 acc = cmake<T2>(0, 0);
#pragma unroll
        for (int j = 0; j < L; ++j) cfma(acc, u[j], twm[(j * m) % L]);
        st(base + m * S, g, acc);
      }
    }
  }
}

template <typename T2, int LMAX, typename Ld, typename St, typename Wf>
__device__ __forceinline__ void dft_run_w(const T2* twp, const T2* twm, int G, int nk, int L, int S, Ld ld, St st,
                                          Wf wf) {
  switch (L) {
    case 1: dft_pass_w<T2, 1>(twp, twm, G, nk, S, ld, st, wf); break;
    case 2: dft_pass_w<T2, 2>(twp, twm, G, nk, S, ld, st, wf); break;
    case 3: dft_pass_w<T2, 3>(twp, twm, G, nk, S, ld, st, wf); break;
    case 4: dft_pass_w<T2, 4>(twp, twm, G, nk, S, ld, st, wf); break;
    case 5: dft_pass_w<T2, 5>(twp, twm, G, nk, S, ld, st, wf); break;
    case 6: dft_pass_w<T2, 6>(twp, twm, G, nk, S, ld, st, wf); break;
    case 7: dft_pass_w<T2, 7>(twp, twm, G, nk, S, ld, st, wf); break;
    case 8: dft_pass_w<T2, 8>(twp, twm, G, nk, S, ld, st, wf); break;
    default:
      if constexpr (LMAX > 8) {
        switch (L) {
          case 9: dft_pass_w<T2, 9>(twp, twm, G, nk, S, ld, st, wf); break;
          case 10: dft_pass_w<T2, 10>(twp, twm, G, nk, S, ld, st, wf); break;
          case 12: dft_pass_w<T2, 12>(twp, twm, G, nk, S, ld, st, wf); break;
          case 16: dft_pass_w<T2, 16>(twp, twm, G, nk, S, ld, st, wf); break;
          default:
            if constexpr (LMAX > 16) {
              switch (L) {
                case 15: dft_pass_w<T2, 15>(twp, twm, G, nk, S, ld, st, wf); break;
                case 18: dft_pass_w<T2, 18>(twp, twm, G, nk, S, ld, st, wf); break;
                case 20: dft_pass_w<T2, 20>(twp, twm, G, nk, S, ld, st, wf); break;
                case 24: dft_pass_w<T2, 24>(twp, twm, G, nk, S, ld, st, wf); break;
                default: break;  // rejected on the host
              }
            }
            break;
        }
      }
      break;
  }
}

template <typename T2, int LMAX, int SIGN, typename Ld, typename St>
__device__ __forceinline__ void dft_run(const T2* tw, int G, int nk, int L, int S, Ld ld, St st) {
  switch (L) {
    case 1: dft_pass<T2, 1, SIGN>(tw, G, nk, S, ld, st); break;
    case 2: dft_pass<T2, 2, SIGN>(tw, G, nk, S, ld, st); break;
    case 3: dft_pass<T2, 3, SIGN>(tw, G, nk, S, ld, st); break;
    case 4: dft_pass<T2, 4, SIGN>(tw, G, nk, S, ld, st); break;
    case 5: dft_pass<T2, 5, SIGN>(tw, G, nk, S, ld, st); break;
    case 6: dft_pass<T2, 6, SIGN>(tw, G, nk, S, ld, st); break;
    case 7: dft_pass<T2, 7, SIGN>(tw, G, nk, S, ld, st); break;
    case 8: dft_pass<T2, 8, SIGN>(tw, G, nk, S, ld, st); break;
    default:
      if constexpr (LMAX > 8) {
        switch (L) {
          case 9: dft_pass<T2, 9, SIGN>(tw, G, nk, S, ld, st); break;
          case 10: dft_pass<T2, 10, SIGN>(tw, G, nk, S, ld, st); break;
          case 12: dft_pass<T2, 12, SIGN>(tw, G, nk, S, ld, st); break;
          case 16: dft_pass<T2, 16, SIGN>(tw, G, nk, S, ld, st); break;
          default:
            if constexpr (LMAX > 16) {
              switch (L) {
                case 15: dft_pass<T2, 15, SIGN>(tw, G, nk, S, ld, st); break;
                case 18: dft_pass<T2, 18, SIGN>(tw, G, nk, S, ld, st); break;
                case 20: dft_pass<T2, 20, SIGN>(tw, G, nk, S, ld, st); break;
                case 24: dft_pass<T2, 24, SIGN>(tw, G, nk, S, ld, st); break;
                default: break;  // rejected on the host
              }
            }
            break;
        }
      }
      break;
  }
}

__device__ void mred_tail(const double2* __restrict__ apart, double2* __restrict__ amp,
                          const double2* __restrict__ K, double2* __restrict__ hmult, int* counter,
                          int nk, int nj, int nj0, double mul, int r, int nred, double2* lds);

// Strip of G consecutive phi columns (c = x*na + a) per workgroup, all nk rows:
//   pass 1 (K3, F_+): HBM -> registers -> LDS, with W(x, :) staged beside it
//   pass 2 (K2, F_+)
//   pass 3 (K1): F_+, the W product and F_- on each pencil in registers
//   passes 4, 5 (K2, K3, F_-), the last one scaling by 1/N_k into HBM.
// nred > 0 prepends that many workgroups running the multiplet reduction, which
// is independent of phi and overlaps the transform.
// PHS (the OBF form, long_range.f90:388-406): phi is multiplied by phs(k, x) as it
// is read and the result by conj(phs(k, x)) as it is written.
template <typename T2, int LMAX, bool PHS = false>
__device__ __forceinline__ void fft_body(const ActArgs<T2>& A, int blk) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  // the reduction workgroups come first so that they are dispatched first
  const int nred = A.nred, nfft = A.nfft;
  if (blk < nred) {
    mred_tail(A.apart, A.amp, A.Kmat, A.hmult, A.mcount, A.nk, A.nj, A.na0 * A.P, A.mul, blk, nred,
              reinterpret_cast<double2*>(smem));
    return;
  }
  const int nk = A.nk, ncol = A.ncol, na = A.na, G = A.G, ldphi = A.ldphi;
  const int K1 = A.K1, K2 = A.K2, K3 = A.K3;
  T2* phi = A.phi;
  const T* W = A.W;
  // twiddles: [dim][sign][m], dims ordered (K3, K2, K1), 2 * LMAX per dim
  T2* twl = reinterpret_cast<T2*>(smem);
  T2* lds = twl + 6 * LMAX;                               // [k][g], padded
  T* wl = reinterpret_cast<T*>(lds + lidx(nk * G - 1) + 1);  // W(x, :) of the strip
  const int Ks[3] = {K3, K2, K1};
  // the LDS twiddle table serves the direct sums of non-power-of-two edges only
  const bool direct = ((K1 & (K1 - 1)) | (K2 & (K2 - 1)) | (K3 & (K3 - 1))) != 0;
  for (int t = threadIdx.x; direct && t < 6 * LMAX; t += blockDim.x) {
    const int d = t / (2 * LMAX), rem = t - d * 2 * LMAX, sg = rem / LMAX, m = rem - sg * LMAX;
    const int L = Ks[d];
    T sn = 0, cs = 1;
    if (m < L) {
      if constexpr (sizeof(T) == 8) sincospi((double)(2 * m) / (double)L, &sn, &cs);
      else sincospif((float)(2 * m) / (float)L, &sn, &cs);
    }
    twl[t] = cmake<T2>(cs, sg == 0 ? sn : -sn);  // sg 0: +1 (backward), 1: -1 (forward)
  }
  // Workgroups are dispatched round-robin over the 8 XCDs; remap so that
  // consecutive column strips (which share 128-B lines) run on the same XCD and
  // therefore hit the same L2.
  int b = blk - nred;
  if ((nfft & 7) == 0) b = (b & 7) * (nfft >> 3) + (b >> 3);
  const int c0 = b * G;
  const int gcnt = min(G, ncol - c0);
  const int xlo = c0 / na, xhi = (c0 + gcnt - 1) / na;
  const int nwl = (xhi - xlo + 1) * nk;
  const T* wsrc = W + (size_t)xlo * nk;
  // W loads issued ahead of the first pass, written to LDS after it
  constexpr int kWR = 4;
  T wr[kWR];
#pragma unroll
  for (int i = 0; i < kWR; ++i) wr[i] = wsrc[min((int)threadIdx.x + i * (int)blockDim.x, nwl - 1)];
  __syncthreads();  // twiddles
  T2* __restrict__ gphi = phi + c0;
  auto ld_hbm = [&](int k, int g) -> T2 {
    const int gc = min(g, gcnt - 1);
    T2 v = gphi[(size_t)k * ldphi + gc];
    if constexpr (PHS) v = cmul(v, A.phs[(size_t)((c0 + gc) / na) * nk + k]);
    return g < gcnt ? v : cmake<T2>(0, 0);
  };
  auto ld_lds = [&](int k, int g) -> T2 { return lds[lidx(k * G + g)]; };
  auto st_lds = [&](int k, int g, T2 v) { lds[lidx(k * G + g)] = v; };
  auto w_at = [&](int k, int g) -> T { return wl[((c0 + min(g, gcnt - 1)) / na - xlo) * nk + k]; };
  const T inv = (T)1 / (T)nk;
  auto st_hbm = [&](int k, int g, T2 v) {
    if (g < gcnt) {
      if constexpr (PHS) {
        const T2 p = A.phs[(size_t)((c0 + g) / na) * nk + k];
        v = cmake<T2>(v.x * p.x + v.y * p.y, v.y * p.x - v.x * p.y);
      }
      v.x *= inv; v.y *= inv;
      gphi[(size_t)k * ldphi + g] = v;
    }
  };
  dft_run<T2, LMAX, 1>(twl + 0 * LMAX, G, nk, K3, 1, ld_hbm, st_lds);
#pragma unroll
  for (int i = 0; i < kWR; ++i) {
    const int idx = threadIdx.x + i * blockDim.x;
    if (idx < nwl) wl[idx] = wr[i];
  }
  for (int idx = threadIdx.x + kWR * blockDim.x; idx < nwl; idx += blockDim.x) wl[idx] = wsrc[idx];
  __syncthreads();
  dft_run<T2, LMAX, 1>(twl + 2 * LMAX, G, nk, K2, K3, ld_lds, st_lds);
  __syncthreads();
  // K1: F_+, W, F_- in registers
  dft_run_w<T2, LMAX>(twl + 4 * LMAX, twl + 5 * LMAX, G, nk, K1, K2 * K3, ld_lds, st_lds, w_at);
  __syncthreads();
  dft_run<T2, LMAX, -1>(twl + 3 * LMAX, G, nk, K2, K3, ld_lds, st_lds);
  __syncthreads();
  dft_run<T2, LMAX, -1>(twl + 1 * LMAX, G, nk, K3, 1, ld_lds, st_hbm);
}

template <typename T2, int LMAX, bool PHS = false>
__global__ __launch_bounds__(256) void k_fft_w(const ActArgs<T2> A) {
  if (skipped(A)) return;
  fft_body<T2, LMAX, PHS>(A, blockIdx.x);
}

template <typename T2, int LMAX>
__global__ __launch_bounds__(256) void k_fft_w_batch(const ActArgs<T2>* __restrict__ As,
                                                     const int* __restrict__ off, int np) {
  const int p = find_problem(off, np, blockIdx.x);
  const ActArgs<T2> A = As[p];
  if (skipped(A)) return;
  fft_body<T2, LMAX>(A, blockIdx.x - off[p]);
}

// ---------------------------------------------------------------------------
// Differentiation: the forward half of the transform, P = F_+ (U A v), with no W
// and no scaling (in place on phi), and the contraction of two such fields into
// the W-derivative of <y|H_lr|x>:
//   d<y|H_lr|x>/dW(k,x) = -(1/N_k) sum_a conj(P_y(k,x,a)) P_x(k,x,a)
// (H_lr = -sum_x A U_x^+ (N_k^-1 F_- W_x F_+) U_x A with F_- = F_+^+).
// ---------------------------------------------------------------------------
template <typename T2, int LMAX>
__global__ __launch_bounds__(256) void k_fft_fwd(const ActArgs<T2> A) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  T2* twl = reinterpret_cast<T2*>(smem);  // [dim][m], exp(+2 pi i m / L), dims (K3, K2, K1)
  T2* lds = twl + 3 * LMAX;
  const int nk = A.nk, ncol = A.ncol, G = A.G, ldphi = A.ldphi;
  const int Ks[3] = {A.K3, A.K2, A.K1};
  const bool direct = ((A.K1 & (A.K1 - 1)) | (A.K2 & (A.K2 - 1)) | (A.K3 & (A.K3 - 1))) != 0;
  for (int t = threadIdx.x; direct && t < 3 * LMAX; t += blockDim.x) {
    const int d = t / LMAX, m = t - d * LMAX;
    const int L = Ks[d];
    T sn = 0, cs = 1;
    if (m < L) {
      if constexpr (sizeof(T) == 8) sincospi((double)(2 * m) / (double)L, &sn, &cs);
      else sincospif((float)(2 * m) / (float)L, &sn, &cs);
    }
    twl[t] = cmake<T2>(cs, sn);
  }
  __syncthreads();
  int b = blockIdx.x;
  const int nfft = gridDim.x;
  if ((nfft & 7) == 0) b = (b & 7) * (nfft >> 3) + (b >> 3);
  const int c0 = b * G;
  const int gcnt = min(G, ncol - c0);
  T2* __restrict__ gphi = A.phi + c0;
  auto ld_hbm = [&](int k, int g) -> T2 {
    const T2 v = gphi[(size_t)k * ldphi + min(g, gcnt - 1)];
    return g < gcnt ? v : cmake<T2>(0, 0);
  };
  auto ld_lds = [&](int k, int g) -> T2 { return lds[lidx(k * G + g)]; };
  auto st_lds = [&](int k, int g, T2 v) { lds[lidx(k * G + g)] = v; };
  auto st_hbm = [&](int k, int g, T2 v) {
    if (g < gcnt) gphi[(size_t)k * ldphi + g] = v;
  };
  dft_run<T2, LMAX, 1>(twl + 0 * LMAX, G, nk, A.K3, 1, ld_hbm, st_lds);
  __syncthreads();
  dft_run<T2, LMAX, 1>(twl + 1 * LMAX, G, nk, A.K2, A.K3, ld_lds, st_lds);
  __syncthreads();
  dft_run<T2, LMAX, 1>(twl + 2 * LMAX, G, nk, A.K1, A.K2 * A.K3, ld_lds, st_hbm);
}

// Q[k][x] += coef * Im sum_a conj(PY(k,x,a)) PX(k,x,a); one thread per (k, x), each
// element updated by one thread per launch (deterministic over energies).
template <typename T2>
__global__ void k_wgrad(const T2* __restrict__ PX, const T2* __restrict__ PY, int nk, int nx, int na,
                        int ldphi, double coef, double* __restrict__ Q) {
  const size_t tot = (size_t)nk * nx;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(e / nx), x = (int)(e - (size_t)k * nx);
    const T2* px = PX + (size_t)k * ldphi + (size_t)x * na;
    const T2* py = PY + (size_t)k * ldphi + (size_t)x * na;
    double im = 0.0;
    for (int a = 0; a < na; ++a) {
      const T2 p = px[a], q = py[a];
      im += (double)q.x * (double)p.y - (double)q.y * (double)p.x;
    }
    Q[e] += coef * im;
  }
}

// ---------------------------------------------------------------------------
// Multiplet projection partials: apart(j = a*P+p, k) = sum_n c(n,k,p) (A psi)(n,k,a)
//   OCEAN_multiplet.f90:1730-1783 (no conjugation).  One workgroup per (k, s):
//   one wave per alpha, lanes over bands, 16 projectors per butterfly reduction.
//   Wave-level only (no barriers), so a workgroup of another kernel can run it.
// ---------------------------------------------------------------------------
constexpr int kProjChunk = 8;

__device__ __forceinline__ void mproj_k(const double2* __restrict__ c, const double2* __restrict__ psi,
                                        const double* __restrict__ allow, double2* __restrict__ apart,
                                        int nb, int nk, int na, int ns, int P, int k, int s, int c0,
                                        int cstep) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  const double2* cb = c + ((size_t)s * nk + k) * P * nb;
  for (int j = wid; j < acnt; j += nw) {
    const int a = s + j * astep;
    for (int p0 = c0 * kProjChunk; p0 < P; p0 += cstep * kProjChunk) {
      double v[2 * kProjChunk];
#pragma unroll
      for (int q = 0; q < 2 * kProjChunk; ++q) v[q] = 0.0;
      for (int n = lane; n < nb; n += 64) {
        const size_t off = ((size_t)a * nk + k) * nb + n;
        double2 pv = psi[off];
        const double m = allow[off];
        pv.x *= m; pv.y *= m;
#pragma unroll
        for (int q = 0; q < kProjChunk; ++q) {
          if (p0 + q < P) {
            const double2 cv = cb[(size_t)(p0 + q) * nb + n];
            v[2 * q] = fma(cv.x, pv.x, v[2 * q]); v[2 * q] = fma(-cv.y, pv.y, v[2 * q]);
            v[2 * q + 1] = fma(cv.x, pv.y, v[2 * q + 1]); v[2 * q + 1] = fma(cv.y, pv.x, v[2 * q + 1]);
          }
        }
      }
      int owner = 0;
      const double tot = wave_multi_sum<double, 2 * kProjChunk>(v, lane, &owner);
      if ((lane & (64 / (2 * kProjChunk) - 1)) == 0) {
        const int q = owner >> 1;
        if (p0 + q < P) {
          double* dst = reinterpret_cast<double*>(apart + (size_t)(a * P + p0 + q) * nk + k);
          dst[owner & 1] = tot;
        }
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_mproj(
    const double2* __restrict__ c, const double2* __restrict__ psi, const double* __restrict__ allow,
    double2* __restrict__ apart, int nb, int nk, int na, int ns, int P) {
  mproj_k(c, psi, allow, apart, nb, nk, na, ns, P, blockIdx.x, blockIdx.y, 0, 1);
}

// amp(j) = sum_k apart(j, k) for j in [j0, j1): one wave per amplitude, fixed order.
__device__ __forceinline__ void mred_sum(const double2* __restrict__ apart, double2* __restrict__ amp,
                                         int nk, int j0, int j1) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = j0 + wid; j < j1; j += nw) {
    const double2* row = apart + (size_t)j * nk;
    double sx = 0.0, sy = 0.0;
    for (int k0 = lane; k0 < nk; k0 += 8 * 64) {
      double2 v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = row[min(k0 + i * 64, nk - 1)];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (k0 + i * 64 < nk) { sx += v[i].x; sy += v[i].y; }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) amp[j] = make_double2(sx, sy);
  }
}

// h(i) = mul * sum_j K(i, j) amp(j): direct -mpm a plus exchange X a
// (OCEAN_multiplet.f90:1812-1936); one wave per row, coalesced rows of K.  With
// several right-hand sides K is block diagonal over blocks of nj0 amplitudes and a
// row sums its own block only (so a NaN in one right-hand side stays there).
__device__ __forceinline__ void mred_matvec(const double2* __restrict__ K, const double2* amp,
                                            double2* __restrict__ hmult, int nj, int nj0, double mul) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i = wid; i < nj; i += nw) {
    double2 acc = make_double2(0.0, 0.0);
    const int j0 = i - i % nj0;
    for (int j = j0 + lane; j < j0 + nj0; j += 64) cfma(acc, K[(size_t)i * nj + j], amp[j]);
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) hmult[i] = make_double2(acc.x * mul, acc.y * mul);
  }
}

// Stand-alone reduction (used when the long-range term is off): one workgroup.
__global__ __launch_bounds__(1024) void k_mred(
    const double2* __restrict__ apart, const double2* __restrict__ K, double2* __restrict__ hmult,
    int nk, int nj, int nj0, double mul) {
  __shared__ double2 amp[kMaxAlpha * 64];
  mred_sum(apart, amp, nk, 0, nj);
  __syncthreads();
  mred_matvec(K, amp, hmult, nj, nj0, mul);
}

// Reduction run by nred extra workgroups of another kernel's grid, one wave per
// amplitude: workgroup r sums amplitudes [r*kRedPerBlock, ...) over k into amp
// (global); the last workgroup to finish (device-scope counter) applies K, four
// rows per wave at a time with all their loads in flight.  The counter returns
// to zero for the next launch.  Loads are batched throughout: the whole tail is
// a few memory latencies, so it hides under the kernel it rides on.
constexpr int kRedPerBlock = 4;

__device__ void mred_tail(const double2* __restrict__ apart, double2* __restrict__ amp,
                          const double2* __restrict__ K, double2* __restrict__ hmult, int* counter,
                          int nk, int nj, int nj0, double mul, int r, int nred, double2* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  mred_sum(apart, amp, nk, r * kRedPerBlock, min(nj, (r + 1) * kRedPerBlock));
  __threadfence();
  __syncthreads();
  __shared__ int is_last;
  if (threadIdx.x == 0) is_last = (atomicAdd(counter, 1) == nred - 1);
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  for (int j = threadIdx.x; j < nj; j += blockDim.x) {
    // other workgroups (possibly on other XCDs) wrote amp: device-scope loads
    const double ax = __hip_atomic_load(&amp[j].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double ay = __hip_atomic_load(&amp[j].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    lds[j] = make_double2(ax, ay);
  }
  __syncthreads();
  // K is block diagonal over right-hand sides (blocks of nj0): each row sums its own
  // block, and a pass of R rows runs over the columns of the blocks they touch
  // four rows per pass: eight took the transform kernel this rides on to 198 VGPRs
  // (two waves per SIMD) from 79
  constexpr int R = 4;
  for (int i0 = wid * R; i0 < nj; i0 += nw * R) {
    double v[2 * R];
#pragma unroll
    for (int q = 0; q < 2 * R; ++q) v[q] = 0.0;
    const int iend = min(i0 + R, nj) - 1;
    const int c_lo = i0 - i0 % nj0, c_hi = iend - iend % nj0 + nj0;
    int rb[R];
#pragma unroll
    for (int q = 0; q < R; ++q) rb[q] = (min(i0 + q, nj - 1)) / nj0;
    for (int c0 = c_lo; c0 < c_hi; c0 += 64) {
      const int j = c0 + lane;
      const int jc = min(j, nj - 1);
      double2 kv[R];
#pragma unroll
      for (int q = 0; q < R; ++q) kv[q] = K[(size_t)min(i0 + q, nj - 1) * nj + jc];
      const double2 a = lds[jc];
      if (j < c_hi) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (j / nj0 != rb[q]) continue;
          v[2 * q] = fma(kv[q].x, a.x, v[2 * q]); v[2 * q] = fma(-kv[q].y, a.y, v[2 * q]);
          v[2 * q + 1] = fma(kv[q].x, a.y, v[2 * q + 1]); v[2 * q + 1] = fma(kv[q].y, a.x, v[2 * q + 1]);
        }
      }
    }
    int owner = 0;
    const double tot = wave_multi_sum<double, 2 * R>(v, lane, &owner);
    if ((lane & (64 / (2 * R) - 1)) == 0) {
      const int q = owner >> 1;
      if (i0 + q < nj) reinterpret_cast<double*>(hmult + i0 + q)[owner & 1] = tot * mul;
    }
  }
  if (threadIdx.x == 0) atomicExch(counter, 0);
}

// ---------------------------------------------------------------------------
// Adjoint projection fused with the whole epilogue of OCEAN_xact:
//   lr(n,k,a)  = - sum_x conj(u(n,k,x)) phi'(k,x,a)         long_range.f90:943-961
//   fg(n,k,a)  =   sum_p conj(c(n,k,p)) h(p,a)               OCEAN_multiplet.f90:1943-1998
//   out        =   E.psi + A.(fg + lr) + S psi                OCEAN_action.f90:337-372, :596-608
// Reads the transposed copy uT[s][k][x][n] (band index fastest): lanes run over
// bands, so every wave load is a contiguous band row, phi'(k,x,:) is the same
// for all lanes of a row (an LDS broadcast), and no cross-lane reduction is
// needed until the XS x-splits of the workgroup are summed once at the end.
// The projector rows c(:,k,p) have the same band-fastest shape and are streamed
// by the same threads (p split like x).  Every operand of the epilogue and the
// projector rows are loaded before the u stream starts, so their latency hides
// under it: with one k per workgroup the whole grid is resident at once and any
// load left to the end would be paid by every workgroup together.
// One workgroup per k; threads (n, xs) with n fastest.  JR alphas per reduction
// round through LDS (the reduction buffer reuses the phi' staging area).
// ---------------------------------------------------------------------------
template <typename T2, int NA>
__device__ __forceinline__ void adj_body(const ActArgs<T2>& A, int blk) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nb = A.nb, nk = A.nk, nx = A.nx, na = A.na, ns = A.ns, P = A.P;
  const int XC = A.XC, JR = A.JR, flags = A.flags, ldT = A.ldT, ldphi = A.ldphi;
  const T2* __restrict__ uT = A.uT;
  const T2* __restrict__ phi = A.phi;
  const double2* __restrict__ psi = A.psi;
  const double2* __restrict__ energy = A.energy;
  const double* __restrict__ allow = A.allow;
  const double2* __restrict__ cproj = A.cproj;
  const double2* __restrict__ hmult = A.hmult;
  const double2* __restrict__ S = A.S;
  double2* __restrict__ out = A.out;
  const int XS = blockDim.x / nb;  // x-splits (host guarantees nb <= blockDim.x)
  double2* res = reinterpret_cast<double2*>(smem);           // [na][nb]
  T2* ph = reinterpret_cast<T2*>(res + na * nb);              // [XC][NA] during the stream
  double2* red = res + na * nb;                               // [XS][JR][nb] afterwards
  // XCD-major workgroup order so that consecutive k share an XCD (phi' rows of
  // neighbouring k are adjacent in memory)
  int kk = blk;
  const int nblk = nk;
  if ((nblk & 7) == 0) kk = (kk & 7) * (nblk >> 3) + (kk >> 3);
  const int k = kk;
  const int t = threadIdx.x, bd = blockDim.x;
  const int n = t % nb, xs = t / nb;
  const bool active = xs < XS;
  const bool lr = (flags & F_LR) != 0;
  const bool fg = (flags & F_FG) != 0;
  const int nel = na * nb;
  const int na0 = A.na0;
  // epilogue operands of the first kEP elements owned by this thread
  constexpr int kEP = 2;
  double2 e_psi[kEP], e_en[kEP];
  double e_m[kEP];
#pragma unroll
  for (int i = 0; i < kEP; ++i) {
    const int e = min(t + i * bd, nel - 1);
    const int a = e / nb, nn = e - a * nb;
    const size_t off = ((size_t)a * nk + k) * nb + nn;
    e_psi[i] = psi[off];
    e_en[i] = energy[off];
    e_m[i] = allow[off];
  }

  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  for (int s = 0; s < ns; ++s) {
    // projector rows p = xs + i XS of this spin, issued first
    constexpr int kCP = 2;
    const double2* cb = cproj + ((size_t)s * nk + k) * P * nb + n;
    double2 cpre[kCP];
    if (fg && P > 0) {
#pragma unroll
      for (int i = 0; i < kCP; ++i) cpre[i] = cb[(size_t)min(xs + i * XS, P - 1) * nb];
    }
    T2 acc[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) acc[j] = cmake<T2>(0, 0);
    if (lr) {
      const T2* ub = uT + ((size_t)s * nk + k) * ldT;
      for (int x0 = 0; x0 < nx; x0 += XC) {
        const int xc = min(XC, nx - x0);
        const int tot = xc * acnt;
        __syncthreads();
        for (int e0 = t; e0 < tot; e0 += 4 * bd) {
          T2 v[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int e = min(e0 + i * bd, tot - 1);
            const int xl = e / acnt, j = e - xl * acnt;
            v[i] = phi[(size_t)k * ldphi + (size_t)(x0 + xl) * na + s + j * astep];
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int e = e0 + i * bd;
            if (e < tot) {
              const int xl = e / acnt, j = e - xl * acnt;
              ph[xl * NA + j] = v[i];
            }
          }
        }
        __syncthreads();
        if (active) {
          // kB independent row loads in flight per wave (addresses clamped, use
          // predicated): 128 bytes per lane in both precisions
          constexpr int kB = sizeof(T2) == 16 ? 8 : 16;
          for (int xb = xs; xb < xc; xb += kB * XS) {
            T2 uv[kB];
#pragma unroll
            for (int i = 0; i < kB; ++i) {
              const int xl = min(xb + i * XS, xc - 1);
              uv[i] = ld_stream(ub + (size_t)(x0 + xl) * nb + n);
            }
            __builtin_amdgcn_sched_barrier(0);
            // one row's NA phi' values at a time: with many alphas, letting the
            // compiler hoist the LDS reads of several rows costs occupancy
#pragma unroll
            for (int i = 0; i < kB; ++i) {
              const int xl = xb + i * XS;
              if (xl < xc) {
                if constexpr (NA >= 6) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < NA; ++j) cfmac(acc[j], uv[i], ph[xl * NA + j]);
              }
            }
          }
        }
      }
    }
    // this thread's share: fg - lr per alpha, in FP64
    double2 part[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) part[j] = make_double2(-(double)acc[j].x, -(double)acc[j].y);
    if (fg && active) {
#pragma unroll
      for (int i = 0; i < kCP; ++i) {
        const int p = xs + i * XS;
        if (p < P) {
#pragma unroll
          for (int j = 0; j < NA; ++j)
            if (j < acnt) cfmac(part[j], cpre[i], hmult[(s + j * astep) * P + p]);
        }
      }
      for (int p = xs + kCP * XS; p < P; p += XS) {
        const double2 cv = cb[(size_t)p * nb];
#pragma unroll
        for (int j = 0; j < NA; ++j)
          if (j < acnt) cfmac(part[j], cv, hmult[(s + j * astep) * P + p]);
      }
    }
    // sum the x-splits, JR alphas per round
    for (int j0 = 0; j0 < acnt; j0 += JR) {
      __syncthreads();
      if (active) {
#pragma unroll
        for (int j = 0; j < NA; ++j)
          if (j >= j0 && j < j0 + JR && j < acnt) red[(xs * JR + (j - j0)) * nb + n] = part[j];
      }
      __syncthreads();
      const int jn = min(JR, acnt - j0);
      for (int e = t; e < jn * nb; e += bd) {
        const int jj = e / nb, nn = e - jj * nb;
        double2 sum = make_double2(0.0, 0.0);
        for (int q = 0; q < XS; ++q) {
          const double2 r = red[(q * JR + jj) * nb + nn];
          sum.x += r.x;
          sum.y += r.y;
        }
        res[(s + (j0 + jj) * astep) * nb + nn] = sum;
      }
    }
  }
  __syncthreads();
  const bool back = (flags & F_BACK) != 0;
  auto finish = [&](int e, double2 pv, double2 ev, double am) {
    const int a = e / nb, nn = e - a * nb;
    const size_t off = ((size_t)a * nk + k) * nb + nn;
    double2 inter = res[e];
    inter.x *= am;
    inter.y *= am;
    if (flags & F_SO) {
      // S mixes the alphas of one right-hand side: rows a0 .. a0 + na0 - 1
      const int ar = a % na0, a0 = a - ar;
      for (int b = 0; b < na0; ++b) {
        double2 sab = S[ar * na0 + b];
        if (back) sab.y = -sab.y;
        cfma(inter, sab, psi[((size_t)(a0 + b) * nk + k) * nb + nn]);
      }
    }
    double2 r = inter;
    if (flags & F_E0) {
      if (back) ev.y = -ev.y;
      r.x = ev.x * pv.x - ev.y * pv.y + inter.x;
      r.y = ev.x * pv.y + ev.y * pv.x + inter.y;
    }
    out[off] = axmy_epilogue(A, pv, r);
  };
#pragma unroll
  for (int i = 0; i < kEP; ++i)
    if (t + i * bd < nel) finish(t + i * bd, e_psi[i], e_en[i], e_m[i]);
  for (int e = t + kEP * bd; e < nel; e += bd) {
    const int a = e / nb, nn = e - a * nb;
    const size_t off = ((size_t)a * nk + k) * nb + nn;
    finish(e, psi[off], energy[off], allow[off]);
  }
}

// One problem: one workgroup per k.
template <typename T2, int NA>
__global__ __launch_bounds__(512) void k_adj(const ActArgs<T2> A) {
  if (skipped(A)) return;
  adj_body<T2, NA>(A, blockIdx.x);
}

template <typename T2, int NA>
__global__ __launch_bounds__(512) void k_adj_batch(const ActArgs<T2>* __restrict__ As,
                                                   const int* __restrict__ off, int np) {
  const int p = find_problem(off, np, blockIdx.x);
  const ActArgs<T2> A = As[p];
  if (skipped(A)) return;
  adj_body<T2, NA>(A, blockIdx.x - off[p]);
}

}  // namespace og
