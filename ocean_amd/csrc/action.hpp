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

// ---------------------------------------------------------------------------
// Projection: phi(k,x,a) = sum_n u(n,k,x) (A psi)(n,k,a)
//   long_range.f90:897-906 (no conjugation of u).  Grid (ceil(nx/256), nk, ns).
// ---------------------------------------------------------------------------
template <typename T2, int NA>
__global__ __launch_bounds__(kBlock) void k_proj(
    const T2* __restrict__ u, const double2* __restrict__ psi, const double* __restrict__ allow,
    T2* __restrict__ phi, int nb, int nk, int nx, int na, int ns) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  T2* ps = reinterpret_cast<T2*>(smem);  // [acnt][nb]
  const int k = blockIdx.y, s = blockIdx.z;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  for (int i = threadIdx.x; i < acnt * nb; i += blockDim.x) {
    const int j = i / nb, n = i - j * nb;
    const int a = s + j * astep;
    const size_t off = ((size_t)a * nk + k) * nb + n;
    const double2 p = psi[off];
    const double m = allow[off];
    ps[i] = cmake<T2>((T)(p.x * m), (T)(p.y * m));
  }
  __syncthreads();
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  T2 acc[NA];
#pragma unroll
  for (int j = 0; j < NA; ++j) acc[j] = cmake<T2>(0, 0);
  const T2* ub = u + ((size_t)(s * nk + k) * nb) * nx + x;
#pragma unroll 8
  for (int n = 0; n < nb; ++n) {
    const T2 uv = ub[(size_t)n * nx];
#pragma unroll
    for (int j = 0; j < NA; ++j)
      if (j < acnt) cfma(acc[j], uv, ps[j * nb + n]);
  }
  T2* out = phi + ((size_t)k * nx + x) * na;
#pragma unroll
  for (int j = 0; j < NA; ++j)
    if (j < acnt) out[s + j * astep] = acc[j];
}

// ---------------------------------------------------------------------------
// Batched 3-D DFT over the k-mesh, staged in LDS, fused with the W multiply:
//   phi' = (1/N_k) F_- [ W(:,x) . F_+ phi ]    long_range.f90:920-939
// F_+ = FFTW_BACKWARD (exp +2 pi i), F_- = FFTW_FORWARD, k3 fastest
// (Modules/FFT_wrapper.f90:67-72, :185-215).  One pencil per thread, L <= 16.
// ---------------------------------------------------------------------------
template <typename T2, int L>
__device__ __forceinline__ void dft_dim(T2* lds, int G, int nk, int S, int sign) {
  using T = typename Scalar<T2>::T;
  T2 tw[L];
#pragma unroll
  for (int m = 0; m < L; ++m) {
    T sn, cs;
    if constexpr (sizeof(T) == 8) sincospi((double)(2 * m) / (double)L, &sn, &cs);
    else sincospif((float)(2 * m) / (float)L, &sn, &cs);
    tw[m] = cmake<T2>(cs, sign * sn);
  }
  const int npen = (G * nk) / L;
  for (int q = threadIdx.x; q < npen; q += blockDim.x) {
    const int g = q % G, r = q / G;
    const int base = (r / S) * (S * L) + (r % S);
    T2 v[L];
#pragma unroll
    for (int j = 0; j < L; ++j) v[j] = lds[(base + j * S) * G + g];
#pragma unroll
    for (int m = 0; m < L; ++m) {
      T2 acc = cmake<T2>(0, 0);
#pragma unroll
      for (int j = 0; j < L; ++j) cfma(acc, v[j], tw[(j * m) % L]);
      lds[(base + m * S) * G + g] = acc;
    }
  }
}

template <typename T2>
__device__ void dft_run(T2* lds, int G, int nk, int L, int S, int sign) {
  switch (L) {
    case 1: break;
    case 2: dft_dim<T2, 2>(lds, G, nk, S, sign); break;
    case 3: dft_dim<T2, 3>(lds, G, nk, S, sign); break;
    case 4: dft_dim<T2, 4>(lds, G, nk, S, sign); break;
    case 5: dft_dim<T2, 5>(lds, G, nk, S, sign); break;
    case 6: dft_dim<T2, 6>(lds, G, nk, S, sign); break;
    case 7: dft_dim<T2, 7>(lds, G, nk, S, sign); break;
    case 8: dft_dim<T2, 8>(lds, G, nk, S, sign); break;
    case 9: dft_dim<T2, 9>(lds, G, nk, S, sign); break;
    case 10: dft_dim<T2, 10>(lds, G, nk, S, sign); break;
    case 12: dft_dim<T2, 12>(lds, G, nk, S, sign); break;
    case 16: dft_dim<T2, 16>(lds, G, nk, S, sign); break;
    default: break;  // rejected on the host
  }
}

template <typename T2>
__global__ __launch_bounds__(kBlock) void k_fft_w(
    T2* __restrict__ phi, const typename Scalar<T2>::T* __restrict__ W, int nk, int ncol, int na,
    int G, int K1, int K2, int K3) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  T2* lds = reinterpret_cast<T2*>(smem);  // [k][g]
  const int c0 = blockIdx.x * G;
  const int gcnt = min(G, ncol - c0);
  for (int e = threadIdx.x; e < nk * G; e += blockDim.x) {
    const int k = e / G, g = e - k * G;
    lds[e] = (g < gcnt) ? phi[(size_t)k * ncol + c0 + g] : cmake<T2>(0, 0);
  }
  __syncthreads();
  dft_run(lds, G, nk, K3, 1, +1);
  __syncthreads();
  dft_run(lds, G, nk, K2, K3, +1);
  __syncthreads();
  dft_run(lds, G, nk, K1, K2 * K3, +1);
  __syncthreads();
  for (int e = threadIdx.x; e < nk * G; e += blockDim.x) {
    const int k = e / G, g = e - k * G;
    if (g < gcnt) {
      const int x = (c0 + g) / na;
      const T w = W[(size_t)x * nk + k];
      T2 v = lds[e];
      v.x *= w; v.y *= w;
      lds[e] = v;
    }
  }
  __syncthreads();
  dft_run(lds, G, nk, K3, 1, -1);
  __syncthreads();
  dft_run(lds, G, nk, K2, K3, -1);
  __syncthreads();
  dft_run(lds, G, nk, K1, K2 * K3, -1);
  __syncthreads();
  const T inv = (T)1 / (T)nk;
  for (int e = threadIdx.x; e < nk * G; e += blockDim.x) {
    const int k = e / G, g = e - k * G;
    if (g < gcnt) {
      T2 v = lds[e];
      v.x *= inv; v.y *= inv;
      phi[(size_t)k * ncol + c0 + g] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Multiplet projection partials: apart(k,a,p) = sum_n c(n,k,p) (A psi)(n,k,a)
//   OCEAN_multiplet.f90:1730-1783 (no conjugation).  Grid (nk, ns).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mproj(
    const double2* __restrict__ c, const double2* __restrict__ psi, const double* __restrict__ allow,
    double2* __restrict__ apart, int nb, int nk, int na, int ns, int P) {
  const int k = blockIdx.x, s = blockIdx.y;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  for (int pair = wid; pair < acnt * P; pair += nw) {
    const int j = pair / P, p = pair - j * P;
    const int a = s + j * astep;
    double2 acc = make_double2(0.0, 0.0);
    const double2* crow = c + (((size_t)s * nk + k) * P + p) * nb;
    for (int n = lane; n < nb; n += 64) {
      const size_t off = ((size_t)a * nk + k) * nb + n;
      double2 pv = psi[off];
      const double m = allow[off];
      pv.x *= m; pv.y *= m;
      cfma(acc, crow[n], pv);
    }
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) apart[((size_t)k * na + a) * P + p] = acc;
  }
}

// Sum the partials over k and apply h = mul * K a (direct -mpm a plus exchange
// X a, OCEAN_multiplet.f90:1812-1936).  One workgroup.
__global__ __launch_bounds__(kBlock) void k_mred(
    const double2* __restrict__ apart, const double2* __restrict__ K, double2* __restrict__ hmult,
    int nk, int nj, double mul) {
  __shared__ double2 amp[kMaxAlpha * 64];
  for (int j = threadIdx.x; j < nj; j += blockDim.x) {
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < nk; ++k) {
      const double2 v = apart[(size_t)k * nj + j];
      s.x += v.x; s.y += v.y;
    }
    amp[j] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nj; i += blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < nj; ++j) cfma(acc, K[(size_t)i * nj + j], amp[j]);
    hmult[i] = make_double2(acc.x * mul, acc.y * mul);
  }
}

// ---------------------------------------------------------------------------
// Adjoint projection fused with the whole epilogue of OCEAN_xact:
//   lr(n,k,a)  = - sum_x conj(u(n,k,x)) phi'(k,x,a)         long_range.f90:943-961
//   fg(n,k,a)  =   sum_p conj(c(n,k,p)) h(p,a)               OCEAN_multiplet.f90:1943-1998
//   out        =   E.psi + A.(fg + lr) + S psi                OCEAN_action.f90:337-372, :596-608
// One workgroup per k; one wave per band row n, lanes over x, wave reduction.
// ---------------------------------------------------------------------------
template <typename T2, int NA>
__global__ __launch_bounds__(kBlock) void k_adj(
    const T2* __restrict__ u, const T2* __restrict__ phi, const double2* __restrict__ psi,
    const double2* __restrict__ energy, const double* __restrict__ allow,
    const double2* __restrict__ cproj, const double2* __restrict__ hmult,
    const double2* __restrict__ S, double2* __restrict__ out, int nb, int nk, int nx, int na,
    int ns, int P, int XC, int flags) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  double2* res = reinterpret_cast<double2*>(smem);  // [na][nb]
  T2* ph = reinterpret_cast<T2*>(res + na * nb);    // [XC][NA]
  const int k = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int e = threadIdx.x; e < na * nb; e += blockDim.x) res[e] = make_double2(0.0, 0.0);
  if (flags & F_LR) {
    for (int s = 0; s < ns; ++s) {
      const int astep = (ns == 2) ? 2 : 1;
      const int acnt = (ns == 2) ? na / 2 : na;
      for (int x0 = 0; x0 < nx; x0 += XC) {
        const int xc = min(XC, nx - x0);
        __syncthreads();
        for (int e = threadIdx.x; e < xc * acnt; e += blockDim.x) {
          const int xl = e / acnt, j = e - xl * acnt;
          ph[xl * NA + j] = phi[((size_t)k * nx + x0 + xl) * na + s + j * astep];
        }
        __syncthreads();
        for (int n = wid; n < nb; n += nw) {
          T2 acc[NA];
#pragma unroll
          for (int j = 0; j < NA; ++j) acc[j] = cmake<T2>(0, 0);
          const T2* urow = u + (((size_t)s * nk + k) * nb + n) * nx + x0;
#pragma unroll 4
          for (int xl = lane; xl < xc; xl += 64) {
            const T2 uv = urow[xl];
#pragma unroll
            for (int j = 0; j < NA; ++j)
              if (j < acnt) cfmac(acc[j], uv, ph[xl * NA + j]);
          }
#pragma unroll
          for (int j = 0; j < NA; ++j) {
            if (j < acnt) {
              const T sr = wave_sum(acc[j].x);
              const T si = wave_sum(acc[j].y);
              if (lane == 0) {
                double2& r = res[(s + j * astep) * nb + n];
                r.x += (double)sr;
                r.y += (double)si;
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  const bool back = (flags & F_BACK) != 0;
  for (int e = threadIdx.x; e < na * nb; e += blockDim.x) {
    const int a = e / nb, n = e - a * nb;
    const size_t off = ((size_t)a * nk + k) * nb + n;
    double2 inter = make_double2(-res[e].x, -res[e].y);
    if (flags & F_FG) {
      const int s = (ns == 2) ? (a & 1) : 0;
      const double2* cb = cproj + ((size_t)s * nk + k) * P * nb + n;
      for (int p = 0; p < P; ++p) cfmac(inter, cb[(size_t)p * nb], hmult[a * P + p]);
    }
    const double am = allow[off];
    inter.x *= am;
    inter.y *= am;
    if (flags & F_SO) {
      for (int b = 0; b < na; ++b) {
        double2 sab = S[a * na + b];
        if (back) sab.y = -sab.y;
        cfma(inter, sab, psi[((size_t)b * nk + k) * nb + n]);
      }
    }
    double2 r = inter;
    if (flags & F_E0) {
      double2 ev = energy[off];
      if (back) ev.y = -ev.y;
      const double2 pv = psi[off];
      r.x = ev.x * pv.x - ev.y * pv.y + inter.x;
      r.y = ev.x * pv.y + ev.y * pv.x + inter.y;
    }
    out[off] = r;
  }
}

}  // namespace og
