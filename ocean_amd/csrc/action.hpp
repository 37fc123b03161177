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
template <typename T2, int NA, int XV>
__global__ __launch_bounds__(kBlock) void k_proj(
    const T2* __restrict__ u, const double2* __restrict__ psi, const double* __restrict__ allow,
    T2* __restrict__ phi, int nb, int nk, int nx, int na, int ns, int ldu, int ldphi) {
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
  // XV consecutive x per thread: one 16-byte load per band row in both precisions
  const int x = (blockIdx.x * blockDim.x + threadIdx.x) * XV;
  if (x >= nx) return;
  T2 acc[XV][NA];
#pragma unroll
  for (int v = 0; v < XV; ++v)
#pragma unroll
    for (int j = 0; j < NA; ++j) acc[v][j] = cmake<T2>(0, 0);
  const T2* ub = u + ((size_t)(s * nk + k) * nb) * ldu + x;
  // Batches of kB band rows: all kB loads are issued before the first use, so
  // each wave keeps kB 1-KB row segments in flight.
  constexpr int kB = 8;
  for (int n0 = 0; n0 < nb; n0 += kB) {
    T2 uv[kB][XV];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int n = min(n0 + i, nb - 1);
      if constexpr (XV == 2) {
        const float4 q = *reinterpret_cast<const float4*>(ub + (size_t)n * ldu);
        uv[i][0] = cmake<T2>(q.x, q.y);
        uv[i][1] = cmake<T2>(q.z, q.w);
      } else {
        uv[i][0] = ub[(size_t)n * ldu];
      }
    }
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      if (n0 + i < nb) {
#pragma unroll
        for (int j = 0; j < NA; ++j) {
          const T2 pj = ps[j * nb + n0 + i];
#pragma unroll
          for (int v = 0; v < XV; ++v) cfma(acc[v][j], uv[i][v], pj);
        }
      }
    }
  }
#pragma unroll
  for (int v = 0; v < XV; ++v) {
    T2* out = phi + (size_t)k * ldphi + (size_t)(x + v) * na;
#pragma unroll
    for (int j = 0; j < NA; ++j) out[s + j * astep] = acc[v][j];
  }
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
    int G, int K1, int K2, int K3, int ldphi) {
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
  {
    // batched tile load: 8 independent loads per thread in flight
    const int tot = nk * G;
    for (int e0 = threadIdx.x; e0 < tot; e0 += 8 * blockDim.x) {
      T2 v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int e = min(e0 + i * (int)blockDim.x, tot - 1);
        const int k = e / G, g = e - k * G;
        v[i] = phi[(size_t)k * ldphi + c0 + min(g, gcnt - 1)];
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int e = e0 + i * (int)blockDim.x;
        if (e < tot) {
          const int g = e % G;
          lds[lidx(e)] = (g < gcnt) ? v[i] : cmake<T2>(0, 0);
        }
      }
    }
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
      phi[(size_t)k * ldphi + c0 + g] = v;
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
// Reads the transposed copy uT[s][k][x][n] (band index fastest): lanes run over
// bands, so every wave load is a contiguous band row, phi'(k,x,:) is the same
// for all lanes of a row (an LDS broadcast), and no cross-lane reduction is
// needed until the XS x-splits of the workgroup are summed once at the end.
// One workgroup per k; threads (n, xs) with n fastest.
// ---------------------------------------------------------------------------
template <typename T2, int NA>
__global__ __launch_bounds__(512) void k_adj(
    const T2* __restrict__ uT, const T2* __restrict__ phi, const double2* __restrict__ psi,
    const double2* __restrict__ energy, const double* __restrict__ allow,
    const double2* __restrict__ cproj, const double2* __restrict__ hmult,
    const double2* __restrict__ S, double2* __restrict__ out, int nb, int nk, int nx, int na,
    int ns, int P, int XC, int flags, int ldT, int ldphi) {
  using T = typename Scalar<T2>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const int XS = blockDim.x / nb;  // x-splits (host guarantees nb <= blockDim.x)
  double2* res = reinterpret_cast<double2*>(smem);           // [na][nb]
  double2* red = res + na * nb;                                // [XS][nb]
  T2* ph = reinterpret_cast<T2*>(red + XS * nb);               // [XC][NA]
  // XCD-major workgroup order so that consecutive k share an XCD (phi' rows of
  // neighbouring k are adjacent in memory)
  int kk = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) kk = (kk & 7) * (nblk >> 3) + (kk >> 3);
  const int k = kk;
  const int t = threadIdx.x;
  const int n = t % nb, xs = t / nb;
  const bool active = xs < XS;
  for (int e = t; e < na * nb; e += blockDim.x) res[e] = make_double2(0.0, 0.0);
  if (flags & F_LR) {
    for (int s = 0; s < ns; ++s) {
      const int astep = (ns == 2) ? 2 : 1;
      const int acnt = (ns == 2) ? na / 2 : na;
      T2 acc[NA];
#pragma unroll
      for (int j = 0; j < NA; ++j) acc[j] = cmake<T2>(0, 0);
      const T2* ub = uT + ((size_t)s * nk + k) * ldT;
      for (int x0 = 0; x0 < nx; x0 += XC) {
        const int xc = min(XC, nx - x0);
        __syncthreads();
        for (int e = t; e < xc * acnt; e += blockDim.x) {
          const int xl = e / acnt, j = e - xl * acnt;
          ph[xl * NA + j] = phi[(size_t)k * ldphi + (size_t)(x0 + xl) * na + s + j * astep];
        }
        __syncthreads();
        if (active) {
          // kB independent row loads in flight per wave (addresses clamped, use predicated)
          constexpr int kB = 8;
          for (int xb = xs; xb < xc; xb += kB * XS) {
            T2 uv[kB];
#pragma unroll
            for (int i = 0; i < kB; ++i) {
              const int xl = min(xb + i * XS, xc - 1);
              uv[i] = ub[(size_t)(x0 + xl) * nb + n];
            }
#pragma unroll
            for (int i = 0; i < kB; ++i) {
              const int xl = xb + i * XS;
              if (xl < xc) {
#pragma unroll
                for (int j = 0; j < NA; ++j) cfmac(acc[j], uv[i], ph[xl * NA + j]);
              }
            }
          }
        }
      }
      // sum the x-splits, one alpha at a time through LDS
#pragma unroll
      for (int j = 0; j < NA; ++j) {
        __syncthreads();
        if (active) red[xs * nb + n] = make_double2((double)acc[j].x, (double)acc[j].y);
        __syncthreads();
        if (t < nb) {
          double2 sum = make_double2(0.0, 0.0);
          for (int q = 0; q < XS; ++q) { sum.x += red[q * nb + t].x; sum.y += red[q * nb + t].y; }
          res[(s + j * astep) * nb + t] = sum;
        }
      }
    }
  }
  __syncthreads();
  const bool back = (flags & F_BACK) != 0;
  for (int e = t; e < na * nb; e += blockDim.x) {
    const int a = e / nb, nn = e - a * nb;
    const size_t off = ((size_t)a * nk + k) * nb + nn;
    double2 inter = make_double2(-res[e].x, -res[e].y);
    if (flags & F_FG) {
      const int s = (ns == 2) ? (a & 1) : 0;
      const double2* cb = cproj + ((size_t)s * nk + k) * P * nb + nn;
      for (int p = 0; p < P; ++p) cfmac(inter, cb[(size_t)p * nb], hmult[a * P + p]);
    }
    const double am = allow[off];
    inter.x *= am;
    inter.y *= am;
    if (flags & F_SO) {
      for (int b = 0; b < na; ++b) {
        double2 sab = S[a * na + b];
        if (back) sab.y = -sab.y;
        cfma(inter, sab, psi[((size_t)b * nk + k) * nb + nn]);
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
