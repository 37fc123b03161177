// Shared definitions for the gfx950 BSE action library.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ocean_gpu.h"

#define OG_CHECK(expr)                                                         \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "ocean_gpu: %s failed: %s (%s:%d)\n", #expr,        \
                   hipGetErrorString(e_), __FILE__, __LINE__);                 \
      return OCEAN_GPU_ERR_HIP;                                                \
    }                                                                          \
  } while (0)

#define OG_TRY(expr)              \
  do {                            \
    int r_ = (expr);              \
    if (r_ != OCEAN_GPU_OK) return r_; \
  } while (0)

namespace og {

constexpr int kBlock = 256;      // threads per workgroup for streaming kernels
constexpr int kRedBlocks = 256;  // fixed partial count: deterministic reductions
constexpr int kMaxAlpha = 12;    // nalpha <= 4 (2 lc + 1) with lc <= 1
constexpr int kMaxDim = 24;      // largest k-mesh edge handled by the LDS DFT

// Tile shapes of the action kernels (defaults; the autotuner picks per batch group).
struct Tiles {
  int proj = 256;  // projection workgroup size (x-points per workgroup)
  int adj = 512;   // adjoint thread budget (x-splits = adj / nb)
  int G = 0;       // transform columns per strip, 0 = the largest that fits
  int adj_mf = -1; // many alphas on one spin: 1 the matrix-core adjoint with its fused
                   // epilogue, 0 the VALU adjoint, -1 the default (matrix cores)
};

// accumulator of one 16x16x4 FP64 matrix-core instruction (four rows per lane)
typedef double __attribute__((ext_vector_type(4))) og_d4;

// complex helpers (double2 / float2 as interleaved complex)
template <typename T2> struct Scalar;
template <> struct Scalar<double2> { using T = double; };
template <> struct Scalar<float2> { using T = float; };

template <typename T2>
__device__ __forceinline__ T2 cmake(typename Scalar<T2>::T r, typename Scalar<T2>::T i) {
  T2 z; z.x = r; z.y = i; return z;
}
// a * b
template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return cmake<T2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc += a * b
template <typename T2>
__device__ __forceinline__ void cfma(T2& acc, T2 a, T2 b) {
  acc.x = fma(a.x, b.x, acc.x); acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y); acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
template <typename T2>
__device__ __forceinline__ void cfmac(T2& acc, T2 a, T2 b) {
  acc.x = fma(a.x, b.x, acc.x); acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y); acc.y = fma(-a.y, b.x, acc.y);
}

// Streaming load of data read once per action (the Bloch-state slabs): the
// nontemporal hint keeps it from displacing reusable lines in L2 / MALL.
__device__ __forceinline__ double2 ld_stream(const double2* p) {
  typedef double v2d __attribute__((ext_vector_type(2)));
  const v2d t = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
  return make_double2(t.x, t.y);
}
// two adjacent complex floats (16-byte aligned)
__device__ __forceinline__ float4 ld_stream4(const float2* p) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f t = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
  return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  typedef float v2f __attribute__((ext_vector_type(2)));
  const v2f t = __builtin_nontemporal_load(reinterpret_cast<const v2f*>(p));
  return make_float2(t.x, t.y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of V values held by every lane over the 64 lanes of a wave, by halving:
// at each of the first log2(V) butterfly steps a lane keeps half of its values
// and receives the partner's copy of that half, so V values cost about V + 5
// shuffles instead of 6V.  On return lane l owns value rev3(l): the value index
// is read from the lane's high bits (bit 5 selects the upper half first).
// V must be a power of two <= 64; lanes hold the total in v[0].
template <typename T, int W, int BIT>
__device__ __forceinline__ void multi_sum_step(T* v, int lane, int& idx) {
  if constexpr (BIT >= 1) {
    if constexpr (W > 1) {
      constexpr int H = W / 2;
      const bool upper = (lane & BIT) != 0;
#pragma unroll
      for (int j = 0; j < H; ++j) {
        const T send = upper ? v[j] : v[j + H];
        const T keep = upper ? v[j + H] : v[j];
        v[j] = keep + __shfl_xor(send, BIT, 64);
      }
      idx = idx * 2 + (upper ? 1 : 0);
      multi_sum_step<T, H, BIT / 2>(v, lane, idx);
    } else {
      v[0] += __shfl_xor(v[0], BIT, 64);
      multi_sum_step<T, 1, BIT / 2>(v, lane, idx);
    }
  }
}

template <typename T, int V>
__device__ __forceinline__ T wave_multi_sum(T (&v)[V], int lane, int* owner) {
  static_assert((V & (V - 1)) == 0 && V <= 64, "V must be a power of two <= 64");
  int idx = 0;
  multi_sum_step<T, V, 32>(v, lane, idx);
  *owner = idx;
  return v[0];
}

// Deterministic sum of nblk partials part[b*nv + j] by one wave (fixed order).
__device__ __forceinline__ double wave_sum_parts(const double* part, int nblk, int nv, int j, int lane) {
  double s = 0.0;
  for (int b = lane; b < nblk; b += 64) s += part[(size_t)b * nv + j];
  return wave_sum(s);
}

// Block-wide sum of nv doubles per thread into out[0..nv) (valid in thread 0).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* >= 4*NV */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = wave_sum(v[j]);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) lds[wid * NV + j] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[w * NV + j];
      v[j] = s;
    }
  }
}

}  // namespace og

// ---------------------------------------------------------------------------
// The handle: module state of long_range / OCEAN_multiplet / OCEAN_energies,
// device-resident, plus the vector pool and scratch.
// ---------------------------------------------------------------------------
struct ocean_gpu_handle {
  ocean_bse_desc d{};
  int nb = 0, nk = 0, nx = 0, na = 0, ns = 0;
  int nrhs = 1;     // right-hand sides sharing this Hamiltonian (a site's polarisations)
  int na0 = 0;      // alphas of one right-hand side, 4 (2 lc + 1); na = nrhs na0
  int km[3] = {0, 0, 0};
  int P = 0;        // compact projector count sum_l (2l+1) nproj(l)
  // Padded leading dimensions: row strides that are large powers of two put the
  // concurrent row reads of all workgroups on the same HBM channels (measured
  // ~20% loss on MI355X), so u rows and phi k-rows carry 32 extra elements.
  int ldu = 0;      // u row stride (elements), >= nx
  int ldphi = 0;    // phi k-row stride (elements), >= nx * na
  int ldT = 0;      // uT k-slab stride (elements), >= nx * nb
  int precision = OCEAN_GPU_FP64;
  int device = 0;
  size_t lds_max = 65536;  // the device's LDS bytes per workgroup (160 KB on gfx950)
  int ncu = 256;           // compute units of the device
  hipStream_t stream = nullptr;

  // long-range state
  double2* u64 = nullptr;   // [s][k][n][x]  (projection)
  float2* u32 = nullptr;    // same layout, mixed form
  double2* uT64 = nullptr;  // [s][k][x][n]  (adjoint), k-slab stride ldT
  float2* uT32 = nullptr;
  double* W64 = nullptr;    // [x][k]
  float* W32 = nullptr;
  double2* phi64 = nullptr; // [k][x][a]
  float2* phi32 = nullptr;
  int32_t* w_i0 = nullptr;  // hat-weight bracket per (x,k) for dW/dptab
  double* w_w0 = nullptr;
  double* w_w1 = nullptr;
  int nptab = 0;

  // multiplet state
  double2* cproj = nullptr; // [s][k][p][n]
  double2* Kmat = nullptr;  // (P*na)^2, h_i = sum_j K[i*(P*na)+j] a_j, i=(a*P+p)
  double2* apart = nullptr; // [a*P+p][k] projection partials
  double2* amp = nullptr;   // [a*P+p]   summed over k
  double2* hmult = nullptr; // [a*P+p]   = mul K amp
  int* mcount = nullptr;    // arrival counter of the reduction workgroups (rests at 0)
  double2* Smat = nullptr;  // [a][b]
  int have_so = 0;

  // energies
  double2* energy = nullptr; // [a][k][n]
  double* allow = nullptr;   // [a][k][n]

  // OBF form of the long-range state (long_range.f90:262-455), FP64 only: the
  // band side runs the projection/adjoint kernels with x -> o over obf_u/obf_uT
  int no = 0;                   // num_obf
  int ldu_o = 0, ldT_o = 0, ldb = 0;
  double2* obf_u = nullptr;     // obf2u [k][n][o], row stride ldu_o
  double2* obf_uT = nullptr;    // obf2u [k][o][n], k stride ldT_o
  double2* obf_f = nullptr;     // obf [o][x]
  double2* obf_fT = nullptr;    // obf [x][o]
  double2* obf_phs = nullptr;   // make_phase's phase [x][k]
  double2* obf_beta = nullptr;  // band-side projection / back-projection [k][o][a], k stride ldb
  bool obf_on = false;

  bool u64_ok = false, u32_ok = false, w64_ok = false, w32_ok = false;
  bool has_mult = false, has_e = false;
  bool has_u() const { return precision == OCEAN_GPU_MIXED ? u32_ok : u64_ok; }
  bool has_w() const { return precision == OCEAN_GPU_MIXED ? w32_ok : w64_ok; }

  // scratch
  double* red = nullptr;     // reduction partials (kRedBlocks * 8 doubles)
  double* scal = nullptr;    // device scalars (Haydock a/b etc.), sized on demand
  size_t scal_n = 0;
  double* host_pin = nullptr;  // pinned staging for host planes
  size_t host_pin_n = 0;
  double2* tmp_in = nullptr;   // host-plane xact staging
  double2* tmp_out = nullptr;

  std::vector<double2*> vecs;  // device vectors (OCEAN_vector analogue)

  og::Tiles tiles;  // tile shapes of the action kernels

  // timing of the last repeat call
  double t_proj = 0, t_fft = 0, t_adj = 0, t_mult = 0, t_total = 0;

  size_t nvec() const { return (size_t)nb * nk * na; }
  size_t nvec0() const { return (size_t)nb * nk * na0; }  // one right-hand side
  size_t nu() const { return (size_t)ns * nk * nb * ldu; }
  size_t nphi() const { return (size_t)nk * ldphi; }
  size_t nuT() const { return (size_t)ns * nk * ldT; }
};
