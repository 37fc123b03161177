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

// One pencil of length L: out[m] = sum_j v[j] w^(j m), w = exp(sign 2 pi i / L);
// tw[m] = w^m comes from LDS.  Power-of-two L: radix-2 decimation in time on data
// loaded in bit-reversed order; otherwise the direct sum.
template <typename T2, int L>
__device__ __forceinline__ void dft_dim(T2* lds, const T2* __restrict__ tw, int G, int nk, int S) {
  const int npen = (G * nk) / L;
  for (int q = threadIdx.x; q < npen; q += blockDim.x) {
    const int g = q % G, r = q / G;
    const int base = (r / S) * (S * L) + (r % S);
    T2 v[L];
    if constexpr ((L & (L - 1)) == 0) {
#pragma unroll
      for (int j = 0; j < L; ++j) v[bitrev<L>(j)] = lds[lidx((base + j * S) * G + g)];
#pragma unroll
      for (int len = 2; len <= L; len <<= 1) {
        constexpr int dummy = 0;
        (void)dummy;
        const int half = len / 2, step = L / len;
#pragma unroll
        for (int i = 0; i < L; i += len) {
#pragma unroll
          for (int j = 0; j < half; ++j) {
            const T2 t = cmul(tw[j * step], v[i + j + half]);
            const T2 u = v[i + j];
            v[i + j] = cmake<T2>(u.x + t.x, u.y + t.y);
            v[i + j + half] = cmake<T2>(u.x - t.x, u.y - t.y);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < L; ++m) lds[lidx((base + m * S) * G + g)] = v[m];
    } else {
#pragma unroll
      for (int j = 0; j < L; ++j) v[j] = lds[lidx((base + j * S) * G + g)];
#pragma unroll
      for (int m = 0; m < L; ++m) {
        T2 acc = cmake<T2>(0, 0);
#pragma unroll
        for (int j = 0; j < L; ++j) cfma(acc, v[j], tw[(j * m) % L]);
        lds[lidx((base + m * S) * G + g)] = acc;
      }
    }
  }
}

template <typename T2, int LMAX>
__device__ void dft_run(T2* lds, const T2* tw, int G, int nk, int L, int S) {
  switch (L) {
    case 1: break;
    case 2: dft_dim<T2, 2>(lds, tw, G, nk, S); break;
    case 3: dft_dim<T2, 3>(lds, tw, G, nk, S); break;
    case 4: dft_dim<T2, 4>(lds, tw, G, nk, S); break;
    case 5: dft_dim<T2, 5>(lds, tw, G, nk, S); break;
    case 6: dft_dim<T2, 6>(lds, tw, G, nk, S); break;
    case 7: dft_dim<T2, 7>(lds, tw, G, nk, S); break;
    case 8: dft_dim<T2, 8>(lds, tw, G, nk, S); break;
    default:
      if constexpr (LMAX > 8) {
        switch (L) {
          case 9: dft_dim<T2, 9>(lds, tw, G, nk, S); break;
          case 10: dft_dim<T2, 10>(lds, tw, G, nk, S); break;
          case 12: dft_dim<T2, 12>(lds, tw, G, nk, S); break;
          case 16: dft_dim<T2, 16>(lds, tw, G, nk, S); break;
          default: break;  // rejected on the host
        }
      }
      break;
  }
}

template <typename T2, int LMAX>
__global__ __launch_bounds__(256) void k_fft_w(
    T2* __restrict__ phi, const typename Scalar<T2>::T* __restrict__ W, int nk, int ncol, int na,
    int G, int K1, int K2, int K3) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  // twiddles: [dim][sign][m], dims ordered (K3, K2, K1), 2 * LMAX per dim
  T2* twl = reinterpret_cast<T2*>(smem);
  T2* lds = twl + 6 * LMAX;  // [k][g], padded
  const int Ks[3] = {K3, K2, K1};
  for (int t = threadIdx.x; t < 6 * LMAX; t += blockDim.x) {
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
  int b = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  const int c0 = b * G;
  const int gcnt = min(G, ncol - c0);
  for (int e = threadIdx.x; e < nk * G; e += blockDim.x) {
    const int k = e / G, g = e - k * G;
    lds[lidx(e)] = (g < gcnt) ? phi[(size_t)k * ncol + c0 + g] : cmake<T2>(0, 0);
  }
  __syncthreads();
  dft_run<T2, LMAX>(lds, twl + 0 * LMAX, G, nk, K3, 1);
  __syncthreads();
  dft_run<T2, LMAX>(lds, twl + 2 * LMAX, G, nk, K2, K3);
  __syncthreads();
  dft_run<T2, LMAX>(lds, twl + 4 * LMAX, G, nk, K1, K2 * K3);
  __syncthreads();
  for (int e = threadIdx.x; e < nk * G; e += blockDim.x) {
    const int k = e / G, g = e - k * G;
    if (g < gcnt) {
      const int x = (c0 + g) / na;
      const T w = W[(size_t)x * nk + k];
      T2 v = lds[lidx(e)];
      v.x *= w; v.y *= w;
      lds[lidx(e)] = v;
    }
  }
  __syncthreads();
  dft_run<T2, LMAX>(lds, twl + 1 * LMAX, G, nk, K3, 1);
  __syncthreads();
  dft_run<T2, LMAX>(lds, twl + 3 * LMAX, G, nk, K2, K3);
  __syncthreads();
  dft_run<T2, LMAX>(lds, twl + 5 * LMAX, G, nk, K1, K2 * K3);
  __syncthreads();
  const T inv = (T)1 / (T)nk;
  for (int e = threadIdx.x; e < nk * G; e += blockDim.x) {
    const int k = e / G, g = e - k * G;
    if (g < gcnt) {
      T2 v = lds[lidx(e)];
      v.x *= inv; v.y *= inv;
      phi[(size_t)k * ncol + c0 + g] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Multiplet projection partials: apart(j = a*P+p, k) = sum_n c(n,k,p) (A psi)(n,k,a)
//   OCEAN_multiplet.f90:1730-1783 (no conjugation).  Grid (nk, ns); one wave per
//   alpha, lanes over bands, 16 projectors per butterfly reduction.
// ---------------------------------------------------------------------------
constexpr int kProjChunk = 16;

__global__ __launch_bounds__(kBlock) void k_mproj(
    const double2* __restrict__ c, const double2* __restrict__ psi, const double* __restrict__ allow,
    double2* __restrict__ apart, int nb, int nk, int na, int ns, int P) {
  const int k = blockIdx.x, s = blockIdx.y;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int astep = (ns == 2) ? 2 : 1;
  const int acnt = (ns == 2) ? na / 2 : na;
  const double2* cb = c + ((size_t)s * nk + k) * P * nb;
  for (int j = wid; j < acnt; j += nw) {
    const int a = s + j * astep;
    for (int p0 = 0; p0 < P; p0 += kProjChunk) {
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

// Sum the partials over k (one wave per amplitude) and apply h = mul * K a:
// direct -mpm a plus exchange X a (OCEAN_multiplet.f90:1812-1936).  One workgroup;
// the matrix-vector product runs one wave per row with coalesced rows of K.
__global__ __launch_bounds__(1024) void k_mred(
    const double2* __restrict__ apart, const double2* __restrict__ K, double2* __restrict__ hmult,
    int nk, int nj, double mul) {
  __shared__ double2 amp[kMaxAlpha * 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = wid; j < nj; j += nw) {
    double sx = 0.0, sy = 0.0;
    for (int k = lane; k < nk; k += 64) {
      const double2 v = apart[(size_t)j * nk + k];
      sx += v.x;
      sy += v.y;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) amp[j] = make_double2(sx, sy);
  }
  __syncthreads();
  for (int i = wid; i < nj; i += nw) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = lane; j < nj; j += 64) cfma(acc, K[(size_t)i * nj + j], amp[j]);
    acc.x = wave_sum(acc.x);
    acc.y = wave_sum(acc.y);
    if (lane == 0) hmult[i] = make_double2(acc.x * mul, acc.y * mul);
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
    int ns, int P, int XC, int rpb, int flags) {
  using T = typename Scalar<T2>::T;
  constexpr int RPW = 1;
  constexpr int VV = 2 * NA * RPW;
  constexpr int V = (VV <= 4) ? 4 : (VV <= 8) ? 8 : (VV <= 16) ? 16 : (VV <= 32) ? 32 : 64;
  extern __shared__ __align__(16) unsigned char smem[];
  double* res = reinterpret_cast<double*>(smem);   // [na][rpb][2]
  T2* ph = reinterpret_cast<T2*>(res + 2 * na * rpb);  // [XC][NA]
  // 1-D grid of nk * nsplit workgroups walking k downwards: the projection
  // streamed u with k ascending, so the adjoint's first rows are the ones most
  // recently pulled through the Infinity Cache (and the next projection, going
  // up again, starts where this pass ends).  The splits of one k are adjacent.
  const int nsplit = (nb + rpb - 1) / rpb;
  const int k = nk - 1 - (int)(blockIdx.x / nsplit);
  const int n0 = (int)(blockIdx.x % nsplit) * rpb;
  const int nrow = min(rpb, nb - n0);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int e = threadIdx.x; e < 2 * na * rpb; e += blockDim.x) res[e] = 0.0;
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
        // RPW band rows per wave pass: RPW*8 loads in flight per lane and one
        // butterfly reduction over all RPW*2*NA values.
        for (int rl0 = wid * RPW; rl0 < nrow; rl0 += nw * RPW) {
          T2 acc[RPW][NA];
#pragma unroll
          for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int j = 0; j < NA; ++j) acc[r][j] = cmake<T2>(0, 0);
          for (int xb = lane; xb < xc; xb += 8 * 64) {
            T2 uv[RPW][8];
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
              const int rr = min(rl0 + r, nrow - 1);
              const T2* urow = u + (((size_t)s * nk + k) * nb + n0 + rr) * nx + x0;
#pragma unroll
              for (int i = 0; i < 8; ++i) {
                const int xl = xb + 64 * i;
                uv[r][i] = (xl < xc) ? urow[xl] : cmake<T2>(0, 0);
              }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
              const int xl = xb + 64 * i;
              if (xl < xc) {
#pragma unroll
                for (int j = 0; j < NA; ++j) {
                  if (j < acnt) {
                    const T2 pj = ph[xl * NA + j];
#pragma unroll
                    for (int r = 0; r < RPW; ++r) cfmac(acc[r][j], uv[r][i], pj);
                  }
                }
              }
            }
          }
          T v[V];
#pragma unroll
          for (int j = 0; j < V; ++j) v[j] = 0;
#pragma unroll
          for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int j = 0; j < NA; ++j) {
              v[(r * NA + j) * 2] = acc[r][j].x;
              v[(r * NA + j) * 2 + 1] = acc[r][j].y;
            }
          int owner = 0;
          const T tot = wave_multi_sum<T, V>(v, lane, &owner);
          if ((lane & (64 / V - 1)) == 0) {
            const int r = (owner >> 1) / NA, j = (owner >> 1) - r * NA, comp = owner & 1;
            if (r < RPW && j < acnt && rl0 + r < nrow)
              res[(((s + j * astep) * rpb) + rl0 + r) * 2 + comp] += (double)tot;
          }
        }
      }
    }
  }
  __syncthreads();
  const bool back = (flags & F_BACK) != 0;
  for (int e = threadIdx.x; e < na * nrow; e += blockDim.x) {
    const int a = e / nrow, rl = e - a * nrow;
    const int n = n0 + rl;
    const size_t off = ((size_t)a * nk + k) * nb + n;
    double2 inter = make_double2(-res[(a * rpb + rl) * 2], -res[(a * rpb + rl) * 2 + 1]);
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
