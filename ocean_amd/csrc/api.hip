// C-ABI of the gfx950 core-level BSE library (declared in include/ocean_gpu.h).
#include <algorithm>
#include <complex>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <type_traits>

#include "action.hpp"
#include "common.hpp"
#include "gemm64.hpp"
#include "obf.hpp"
#include "setup.hpp"
#include "vec.hpp"

using og::kBlock;
using og::kRedBlocks;
using H = ocean_gpu_handle;
using cplx = std::complex<double>;

namespace {

constexpr double kEV2Hartree = 0.036749322175666;   // OCEAN2/OCEAN_constants.f90:15
constexpr double kHartree2eV = 27.211386245980;     // OCEAN2/OCEAN_constants.f90:16
constexpr double kAlpha = 0.0072973525643;           // OCEAN2/OCEAN_constants.f90:18
constexpr double kAlphaInv = 137.035999177;           // OCEAN2/OCEAN_constants.f90:19
constexpr double kBohr = 0.529177210544;             // OCEAN2/OCEAN_constants.f90:20 (Angstrom)
constexpr double kAu2sec = 2.4188843265864e-17;      // OCEAN2/OCEAN_constants.f90:22
constexpr double kPi = 3.1415926535897932384626433832795028841971693993751;  // :29

int grid_for(size_t n) {
  size_t g = (n + kBlock - 1) / kBlock;
  return (int)std::min<size_t>(std::max<size_t>(g, 1), 4096);
}

template <typename T>
int dmalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return OCEAN_GPU_ERR_NOMEM;
  }
  return OCEAN_GPU_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

int ensure_scal(H* h, size_t n) {
  if (h->scal_n >= n) return OCEAN_GPU_OK;
  dfree(h->scal);
  OG_TRY(dmalloc(&h->scal, n));
  h->scal_n = n;
  return OCEAN_GPU_OK;
}

int ensure_pin(H* h, size_t ndouble) {
  if (h->host_pin_n >= ndouble) return OCEAN_GPU_OK;
  if (h->host_pin) (void)hipHostFree(h->host_pin);
  h->host_pin = nullptr;
  if (hipHostMalloc((void**)&h->host_pin, ndouble * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    h->host_pin_n = 0;
    return OCEAN_GPU_ERR_NOMEM;
  }
  h->host_pin_n = ndouble;
  return OCEAN_GPU_OK;
}

// host planes (n) -> device interleaved complex
int put_planes(H* h, const double* re, const double* im, double2* dst, size_t n) {
  OG_TRY(ensure_pin(h, 2 * n));
  double* p = h->host_pin;
  for (size_t i = 0; i < n; ++i) {
    p[2 * i] = re[i];
    p[2 * i + 1] = im ? im[i] : 0.0;
  }
  OG_CHECK(hipMemcpyAsync(dst, p, n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  OG_CHECK(hipStreamSynchronize(h->stream));
  return OCEAN_GPU_OK;
}

int get_planes(H* h, const double2* src, double* re, double* im, size_t n) {
  OG_TRY(ensure_pin(h, 2 * n));
  double* p = h->host_pin;
  OG_CHECK(hipMemcpyAsync(p, src, n * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
  OG_CHECK(hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < n; ++i) {
    re[i] = p[2 * i];
    im[i] = p[2 * i + 1];
  }
  return OCEAN_GPU_OK;
}

int na_tmpl(int acnt) {
  if (acnt <= 2) return 2;
  if (acnt <= 4) return 4;
  if (acnt <= 6) return 6;
  if (acnt <= 12) return 12;
  return 24;
}

// k-mesh edges of the LDS transform: 1..10, 12 and 16 as before, and the 2^a 3^b 5^c
// edges 15, 18, 20 and 24 that the reference's bundled FFT also takes
// (Modules/jlmfft.f:64-74) by direct sums in the 24-edge class.
bool dim_ok(int L) {
  return (L >= 1 && L <= 10) || L == 12 || L == 15 || L == 16 || L == 18 || L == 20 || L == 24;
}

// Transform instantiation class of a k mesh: largest edge <= 8, <= 16 or <= 24.
int edge_class(const int km[3]) {
  const int m = std::max(km[0], std::max(km[1], km[2]));
  return m <= 8 ? 8 : (m <= 16 ? 16 : 24);
}

// Launches with more than 64 KB of dynamic LDS (large k meshes, many bands) state
// their size to the runtime first; the device limit is checked when the shape is set.
template <typename K>
int lds_opt_in(K* kernel, size_t bytes) {
  if (bytes <= 65536) return OCEAN_GPU_OK;
  OG_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
  return OCEAN_GPU_OK;
}

int flags_of(const H* h, int backwards, const int32_t* hflag) {
  int hf[6] = {1, 0, 0, 0, 1, 1};
  if (hflag)
    for (int i = 0; i < 6; ++i) hf[i] = hflag[i];
  int f = 0;
  if (h->d.long_range && hf[0] == 1) f |= og::F_LR;
  if (h->d.mult && hf[4] == 1) f |= og::F_FG;
  if (h->d.mult && hf[5] == 1 && h->have_so) f |= og::F_SO;
  if (h->d.e0) f |= og::F_E0;
  if (backwards) f |= og::F_BACK;
  return f;
}

int check_ready(const H* h, int flags) {
  if (!h->has_e) return OCEAN_GPU_ERR_STATE;
  if ((flags & og::F_LR) && !((h->obf_on || h->has_u()) && h->has_w())) return OCEAN_GPU_ERR_STATE;
  if ((flags & (og::F_FG | og::F_SO)) && !h->has_mult) return OCEAN_GPU_ERR_STATE;
  return OCEAN_GPU_OK;
}

// --------------------------------------------------------------------------
// The action, device pointer to device pointer, on h->stream.
// --------------------------------------------------------------------------
using og::Tiles;

constexpr size_t kLdsBudget = 64 * 1024;  // per workgroup: two adjoint workgroups per CU

template <typename T2>
struct Geom {
  size_t sm_proj = 0, sm_fft = 0, sm_adj = 0;
  int adj_threads = 0, nblk_proj = 0, nblk_fft = 0, nblk_adj = 0, lmax = 8;
};

// The x extent and strides the projection/adjoint kernels run over: the x mesh
// for the Bloch form, the OBF index for the band side of the OBF form.
struct BandSide {
  int nx, ldu, ldT, ldphi;
};

// Fill the kernel arguments and launch geometry of one problem.  The
// multiplet term rides on the long-range kernels when both are on: its partials
// in the last x-tile workgroup of k_proj, its k-sum and K apply in workgroups
// prepended to k_fft_w; both only read psi, so one stream carries the action.
template <typename T2, int NA>
int make_args(const H* h, const double2* in, double2* out, int flags, const Tiles& t, const T2* u,
              const T2* uT, T2* phi, const typename og::Scalar<T2>::T* W, og::ActArgs<T2>& a,
              Geom<T2>& g, const BandSide* bs = nullptr) {
  using T = typename og::Scalar<T2>::T;
  const int nb = h->nb, nk = h->nk, na = h->na, ns = h->ns;
  const int nx = bs ? bs->nx : h->nx;
  const int acnt = (ns == 2) ? na / 2 : na;
  a = og::ActArgs<T2>{};
  a.u = u; a.uT = uT; a.phi = phi; a.W = W;
  a.psi = in; a.out = out;
  a.energy = h->energy; a.allow = h->allow;
  a.cproj = h->cproj; a.apart = h->apart; a.amp = h->amp; a.Kmat = h->Kmat; a.hmult = h->hmult;
  a.mcount = h->mcount; a.S = h->Smat;
  a.mul = h->d.inter_scale / ((double)nk * h->d.celvol);
  a.nb = nb; a.nk = nk; a.nx = nx; a.na = na; a.ns = ns; a.P = h->P; a.nj = na * h->P; a.na0 = h->na0;
  a.ldu = bs ? bs->ldu : h->ldu;
  a.ldT = bs ? bs->ldT : h->ldT;
  a.ldphi = bs ? bs->ldphi : h->ldphi;
  a.flags = flags;
  const bool lr = (flags & og::F_LR) != 0, fg = (flags & og::F_FG) != 0;
  // projection
  if (t.proj < 64 || t.proj > 512 || (t.proj & 63)) return OCEAN_GPU_ERR_ARG;
  // FP32 streaming workgroups cover two x points per thread; the FP32 matrix-core
  // form (many alphas) covers 32 per wave and stages B as nb4 rows of proj_brow floats
  const bool mf = sizeof(T) == 4 && NA >= 6;
  const bool mf64 = sizeof(T) == 8 && NA >= 6;
  const int xper = mf ? 32 * (t.proj / 64) : mf64 ? 16 * og::kProjMT64 * (t.proj / 64)
                                                  : t.proj * (sizeof(T) == 4 ? 2 : 1);
  a.px = (nx + xper - 1) / xper;
  // the FP64 matrix-core form (many alphas) forms A psi per lane and stages nothing
  g.sm_proj = mf ? (size_t)((nb + 3) & ~3) * og::proj_brow<NA>() * sizeof(float)
              : mf64 ? 0 : (size_t)acnt * nb * sizeof(T2);
  // the FP64 matrix-core form stages A psi (nb acnt complex) where it fits 64 KB, up to
  // twelve alphas (measured: L2,3 232 -> 222 us; at 24 alphas 121 -> 133 us)
  if (mf64 && NA <= 12 && (size_t)nb * acnt * sizeof(double2) <= 65536) {
    a.bstage = 1;
    g.sm_proj = (size_t)nb * acnt * sizeof(double2);
  }
  g.nblk_proj = a.px * nk * ns;
  // transform
  a.ncol = nx * na;
  const int cap = (int)(57344 / sizeof(T2)) / nk;
  // default strip: 2048 elements; stacked right-hand sides in FP32 on a 512-point
  // k mesh take twice that (measured: six polarisations 173.3 -> 172.0 us, three
  // 127.2 -> 125.8 us per action; one right-hand side and L2,3 are not helped)
  int Gdef = std::min((256 * 8) / nk, cap);
  if (sizeof(T) == 4 && h->nrhs > 1 && nk >= 512) Gdef = std::min(2 * Gdef, cap);
  int G = t.G > 0 ? t.G : Gdef;
  G = std::max(1, std::min(std::min(G, std::max(cap, 1)), a.ncol));
  a.G = G;
  a.nfft = (a.ncol + G - 1) / G;
  a.nred = (lr && fg) ? (a.nj + og::kRedPerBlock - 1) / og::kRedPerBlock : 0;
  a.K1 = h->km[0]; a.K2 = h->km[1]; a.K3 = h->km[2];
  g.lmax = edge_class(h->km);
  const size_t tile = (size_t)nk * G;
  g.sm_fft = (6 * (size_t)g.lmax + tile + tile / 8 + 1) * sizeof(T2) +
             (size_t)((G - 1) / na + 2) * nk * sizeof(T);
  g.sm_fft = std::max(g.sm_fft, (size_t)a.nj * sizeof(double2));
  if (g.sm_fft > h->lds_max || g.sm_proj > h->lds_max) return OCEAN_GPU_ERR_ARG;  // k mesh or bands too large
  g.nblk_fft = a.nred + a.nfft;
  // adjoint: x-splits, phi' chunk and reduction rounds sharing one LDS area
  if (nb > t.adj) return OCEAN_GPU_ERR_ARG;
  const int XS = std::max(1, t.adj / nb);
  g.adj_threads = nb * XS;
  const size_t resb = (size_t)na * nb * sizeof(double2);
  // two adjoint workgroups per CU while the result rows fit half the LDS; past that
  // the device's whole per-workgroup LDS (many bands at the L2,3 edge)
  size_t lds_total = kLdsBudget;
  if (resb + (size_t)XS * nb * sizeof(double2) > lds_total) lds_total = h->lds_max;
  if (resb + (size_t)XS * nb * sizeof(double2) > lds_total) return OCEAN_GPU_ERR_ARG;
  const size_t budget = lds_total - resb;
  const size_t per_j = (size_t)XS * nb * sizeof(double2);
  a.JR = (int)std::max<size_t>(1, std::min<size_t>(acnt, budget / per_j));
  int XC = nx;
  const int xcap = (int)(budget / (NA * sizeof(T2)));
  if (XC > xcap) XC = std::max(XS, (xcap / XS) * XS);
  a.XC = XC;
  g.sm_adj = resb + std::max((size_t)XC * NA * sizeof(T2), (size_t)a.JR * per_j);
  g.nblk_adj = nk;
  return OCEAN_GPU_OK;
}

// Many alphas on one spin: the adjoint as a GEMM per k on the matrix cores with the
// whole epilogue in the same kernel (k_adj_lr), where the VALU adjoint is bound by
// its nalpha LDS broadcasts of phi' per band row (measured faster in both precisions
// at 12 alphas).  The tile can force either form.
template <typename T2, int NA>
bool adj_mf(const H* h, int flags, const Tiles& t) {
  if (NA < 6 || h->ns != 1 || h->na != NA) return false;
  if (NA > 12) return true;  // 24 alphas: the only adjoint (with lr off it runs the epilogue alone)
  if (!(flags & og::F_LR)) return false;
  return t.adj_mf != 0;
}

// The device state of one precision form.
template <typename T2> struct StateOf;
template <> struct StateOf<double2> {
  static const double2* u(const H* h) { return h->u64; }
  static const double2* uT(const H* h) { return h->uT64; }
  static double2* phi(const H* h) { return h->phi64; }
  static const double* W(const H* h) { return h->W64; }
};
template <> struct StateOf<float2> {
  static const float2* u(const H* h) { return h->u32; }
  static const float2* uT(const H* h) { return h->uT32; }
  static float2* phi(const H* h) { return h->phi32; }
  static const float* W(const H* h) { return h->W32; }
};

// A handle's tiles for one precision: FP32 streaming projection threads take two x
// points each, so half as many threads keep the x points per workgroup of FP64; the
// FP32 matrix-core form for many alphas (32 x points per wave) takes twice as many.
template <typename T2, int NA>
Tiles tiles_for(const H* h) {
  Tiles t = h->tiles;
  if (sizeof(typename og::Scalar<T2>::T) == 4) {
    if (NA < 6) t.proj = std::max(64, t.proj / 2);
    else t.proj = std::min(512, 2 * t.proj);  // fewer, wider workgroups per k share the B staging
  }
  return t;
}

template <typename T2, int NA>
int xact_t(H* h, const double2* in, double2* out, int flags, T2* u, T2* uT, T2* phi,
           const typename og::Scalar<T2>::T* W, hipEvent_t* ev, const int* skip, const double2* axmy,
           const og::GmFin* fin) {
  og::ActArgs<T2> a;
  Geom<T2> g;
  const Tiles tl = tiles_for<T2, NA>(h);
  OG_TRY((make_args<T2, NA>(h, in, out, flags, tl, u, uT, phi, W, a, g)));
  a.skip = skip;
  if (axmy) {
    a.axmy = 1;
    a.axmy_c = *axmy;
  }
  // a pending GMRES finish rides on the projection (the caller passes one only when
  // the long-range term, and so k_proj, runs)
  if (fin) a.fin = *fin;
  if (ev) OG_CHECK(hipEventRecord(ev[0], h->stream));
  if (flags & og::F_LR) {
    OG_TRY(lds_opt_in(og::k_proj<T2, NA>, g.sm_proj));
    og::k_proj<T2, NA><<<g.nblk_proj, tl.proj, g.sm_proj, h->stream>>>(a);
    OG_CHECK(hipGetLastError());
    if (ev) OG_CHECK(hipEventRecord(ev[1], h->stream));
    if (g.lmax == 8) {
      og::k_fft_w<T2, 8><<<g.nblk_fft, kBlock, g.sm_fft, h->stream>>>(a);
    } else if (g.lmax == 16) {
      OG_TRY(lds_opt_in(og::k_fft_w<T2, 16>, g.sm_fft));
      og::k_fft_w<T2, 16><<<g.nblk_fft, kBlock, g.sm_fft, h->stream>>>(a);
    } else {
      OG_TRY(lds_opt_in(og::k_fft_w<T2, 24>, g.sm_fft));
      og::k_fft_w<T2, 24><<<g.nblk_fft, kBlock, g.sm_fft, h->stream>>>(a);
    }
    OG_CHECK(hipGetLastError());
    if (ev) OG_CHECK(hipEventRecord(ev[2], h->stream));
  } else if (ev) {
    OG_CHECK(hipEventRecord(ev[1], h->stream));
    OG_CHECK(hipEventRecord(ev[2], h->stream));
  }
  if ((flags & og::F_FG) && !(flags & og::F_LR)) {
    og::k_mproj<<<dim3(h->nk, h->ns), kBlock, 0, h->stream>>>(h->cproj, in, h->allow, h->apart, h->nb,
                                                             h->nk, h->na, h->ns, h->P);
    OG_CHECK(hipGetLastError());
    og::k_mred<<<1, 1024, 0, h->stream>>>(h->apart, h->Kmat, h->hmult, h->nk, a.nj, h->na0 * h->P, a.mul);
    OG_CHECK(hipGetLastError());
  }
  if (ev) OG_CHECK(hipEventRecord(ev[3], h->stream));
  bool mf = false;
  if constexpr (NA >= 6) {
    mf = adj_mf<T2, NA>(h, flags, tl);
    // 24 alphas (several right-hand sides on one spin) have the matrix-core adjoint only
    if (NA > 12 && !mf) return OCEAN_GPU_ERR_UNSUPPORTED;
    if (mf) {
      // the deep load ring where two workgroups per CU hold the whole grid
      const int nwg = ((h->nb + 63) / 64) * h->nk;
      if (nwg <= 2 * h->ncu)
        og::k_adj_lr<T2, NA, true><<<nwg, 256, 0, h->stream>>>(a);
      else
        og::k_adj_lr<T2, NA, false><<<nwg, 256, 0, h->stream>>>(a);
    }
  }
  if constexpr (NA <= 12) {
    if (!mf) {
      OG_TRY(lds_opt_in(og::k_adj<T2, NA>, g.sm_adj));
      og::k_adj<T2, NA><<<g.nblk_adj, g.adj_threads, g.sm_adj, h->stream>>>(a);
    }
  }
  OG_CHECK(hipGetLastError());
  if (ev) OG_CHECK(hipEventRecord(ev[4], h->stream));
  return OCEAN_GPU_OK;
}

// The action with the OBF long-range term (long_range.f90:262-455):
//   k_proj over obf2u        beta(k,o,a) = sum_n obf2u (A psi)        (:357-360)
//   k_obf_fwd                phi(k,x,a)  = sum_o beta obf             (:373-382)
//   k_fft_w with the phase   Y = conj(phs) F_- W F_+ (phs phi) / N_k  (:388-406)
//   k_obf_bwd                b(k,o,a) = -N_k sum_x conj(obf) Y        (:412-421)
//   k_adj over obf2u         hpsi = E psi + A (-sum_o conj(obf2u) b) + S psi
// The -N_k undoes the transform's 1/N_k (cfft is unnormalised, Modules/jlmfft.f:86-89)
// and turns the adjoint's subtraction into lr_act_obf2's addition (:448-451).  The
// multiplet term is not added: lr_act_obf2 overwrites the vector it was
// accumulated into (:445-446).  Timing events: proj = band-side projection, fft =
// the x side (both obf products and the transform), adj = the adjoint.
template <int NA>
int xact_obf_t(H* h, const double2* in, double2* out, int flags, hipEvent_t* ev, const int* skip,
               const double2* axmy) {
  flags &= ~og::F_FG;
  og::ActArgs<double2> a, ab;
  Geom<double2> g, gb;
  const Tiles tl = tiles_for<double2, NA>(h);
  OG_TRY((make_args<double2, NA>(h, in, out, flags, tl, nullptr, nullptr, h->phi64, h->W64, a, g)));
  a.phs = h->obf_phs;
  const BandSide bs{h->no, h->ldu_o, h->ldT_o, h->ldb};
  OG_TRY((make_args<double2, NA>(h, in, out, flags, tl, h->obf_u, h->obf_uT, h->obf_beta, nullptr, ab, gb,
                                 &bs)));
  a.skip = skip;
  ab.skip = skip;
  if (axmy) {
    ab.axmy = 1;
    ab.axmy_c = *axmy;
  }
  const int nk = h->nk, nx = h->nx, no = h->no;
  if (ev) OG_CHECK(hipEventRecord(ev[0], h->stream));
  og::k_proj<double2, NA><<<gb.nblk_proj, tl.proj, gb.sm_proj, h->stream>>>(ab);
  OG_CHECK(hipGetLastError());
  if (ev) OG_CHECK(hipEventRecord(ev[1], h->stream));
  // 16-column tiles of the flattened (k, alpha, re/im) output columns
  const int ntiles = (2 * nk * NA + 15) / 16;
  og::k_obf_fwd<NA, 2><<<dim3((nx + 63) / 64, (ntiles + 7) / 8), 256, 0, h->stream>>>(
      h->obf_beta, h->ldb, h->obf_f, h->phi64, h->ldphi, nk, nx, no);
  OG_CHECK(hipGetLastError());
  if (g.lmax == 8) {
    og::k_fft_w<double2, 8, true><<<g.nblk_fft, kBlock, g.sm_fft, h->stream>>>(a);
  } else if (g.lmax == 16) {
    OG_TRY(lds_opt_in(og::k_fft_w<double2, 16, true>, g.sm_fft));
    og::k_fft_w<double2, 16, true><<<g.nblk_fft, kBlock, g.sm_fft, h->stream>>>(a);
  } else {
    OG_TRY(lds_opt_in(og::k_fft_w<double2, 24, true>, g.sm_fft));
    og::k_fft_w<double2, 24, true><<<g.nblk_fft, kBlock, g.sm_fft, h->stream>>>(a);
  }
  OG_CHECK(hipGetLastError());
  og::k_obf_bwd<NA, 1><<<dim3((no + 63) / 64, ntiles), 256, 0, h->stream>>>(
      h->phi64, h->ldphi, h->obf_fT, h->obf_beta, h->ldb, nk, nx, no, -(double)nk);
  OG_CHECK(hipGetLastError());
  if (ev) {
    OG_CHECK(hipEventRecord(ev[2], h->stream));
    OG_CHECK(hipEventRecord(ev[3], h->stream));
  }
  OG_TRY(lds_opt_in(og::k_adj<double2, NA>, gb.sm_adj));
  og::k_adj<double2, NA><<<gb.nblk_adj, gb.adj_threads, gb.sm_adj, h->stream>>>(ab);
  OG_CHECK(hipGetLastError());
  if (ev) OG_CHECK(hipEventRecord(ev[4], h->stream));
  return OCEAN_GPU_OK;
}

// skip (nullable): every kernel returns at once while *skip != 0; axmy (nullable):
// out = c in - H in with c = *axmy, formed in the adjoint's epilogue.
int xact_dev(H* h, const double2* in, double2* out, int flags, hipEvent_t* ev = nullptr,
             const int* skip = nullptr, const double2* axmy = nullptr, const og::GmFin* fin = nullptr) {
  const int acnt = (h->ns == 2) ? h->na / 2 : h->na;
  const int NA = na_tmpl(acnt);
  const bool mixed = (h->precision == OCEAN_GPU_MIXED);
  if (h->obf_on && (flags & og::F_LR)) {
    if (mixed || h->ns != 1) return OCEAN_GPU_ERR_STATE;
    if (NA == 4 && h->na == 4) return xact_obf_t<4>(h, in, out, flags, ev, skip, axmy);
    if (NA == 12 && h->na == 12) return xact_obf_t<12>(h, in, out, flags, ev, skip, axmy);
    return OCEAN_GPU_ERR_ARG;
  }
#define OG_DISPATCH(NAV)                                                              \
  if (NA == NAV)                                                                      \
    return mixed ? xact_t<float2, NAV>(h, in, out, flags, h->u32, h->uT32, h->phi32, h->W32, ev, skip, axmy, fin) \
                 : xact_t<double2, NAV>(h, in, out, flags, h->u64, h->uT64, h->phi64, h->W64, ev, skip, axmy, fin);
  OG_DISPATCH(2)
  OG_DISPATCH(4)
  OG_DISPATCH(6)
  OG_DISPATCH(12)
  OG_DISPATCH(24)
#undef OG_DISPATCH
  return OCEAN_GPU_ERR_ARG;
}

int ensure_precision_buffers(H* h) {
  const size_t nu = h->nu();
  const size_t nphi = h->nphi();
  const size_t nw = (size_t)h->nk * h->nx;
  const size_t nuT = h->nuT();
  if (h->precision == OCEAN_GPU_MIXED) {
    if (!h->uT32) OG_TRY(dmalloc(&h->uT32, nuT));
    if (!h->u32) OG_TRY(dmalloc(&h->u32, nu));
    if (!h->phi32) OG_TRY(dmalloc(&h->phi32, nphi));
    if (!h->W32) OG_TRY(dmalloc(&h->W32, nw));
  } else {
    if (!h->uT64) OG_TRY(dmalloc(&h->uT64, nuT));
    if (!h->u64) OG_TRY(dmalloc(&h->u64, nu));
    if (!h->phi64) OG_TRY(dmalloc(&h->phi64, nphi));
    if (!h->W64) OG_TRY(dmalloc(&h->W64, nw));
  }
  return OCEAN_GPU_OK;
}

double2* vec_ptr(H* h, int32_t id) {
  if (id < 0 || (size_t)id >= h->vecs.size()) return nullptr;
  return h->vecs[id];
}

int reduce_to_host(H* h, int nv, double* out) {
  og::k_finish<<<1, 64, 0, h->stream>>>(h->red, kRedBlocks, nv, h->red + kRedBlocks * 8);
  OG_CHECK(hipGetLastError());
  OG_CHECK(hipMemcpyAsync(out, h->red + kRedBlocks * 8, nv * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
  OG_CHECK(hipStreamSynchronize(h->stream));
  return OCEAN_GPU_OK;
}

int dev_dot(H* h, const double2* p, const double2* q, double* r, double* i) {
  og::k_dot_part<<<kRedBlocks, kBlock, 0, h->stream>>>(p, q, h->nvec(), h->red);
  OG_CHECK(hipGetLastError());
  double o[2];
  OG_TRY(reduce_to_host(h, 2, o));
  *r = o[0];
  if (i) *i = o[1];
  return OCEAN_GPU_OK;
}

int dev_nrm(H* h, const double2* p, double* v) {
  og::k_nrm_part<<<kRedBlocks, kBlock, 0, h->stream>>>(p, h->nvec(), h->red);
  OG_CHECK(hipGetLastError());
  return reduce_to_host(h, 1, v);
}

// ---------------------------------------------------------------------------
// GMRES helpers
// ---------------------------------------------------------------------------
// <W_i|v1>, <W_i|v2> for i < m: partials [blk][i][4].  Eight W rows per pass with
// all their loads in flight; the 32 sums of a pass reduce by one butterfly per
// wave and one LDS round across the waves.
__global__ __launch_bounds__(kBlock) void k_multidot2(const int* __restrict__ skip,
                                                      const double2* const* __restrict__ Wl, int m,
                                                      const double2* __restrict__ v1,
                                                      const double2* __restrict__ v2, size_t n,
                                                      double* __restrict__ part) {
  if (*skip) return;
  constexpr int R = 8;
  __shared__ double lds[(kBlock / 64) * 4 * R];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i0 = 0; i0 < m; i0 += R) {
    double v[4 * R];
#pragma unroll
    for (int q = 0; q < 4 * R; ++q) v[q] = 0.0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
      const double2 b = v1[e], c = v2[e];
      double2 a[R];
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] = Wl[min(i0 + q, m - 1)][e];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        v[4 * q + 0] = fma(a[q].x, b.x, v[4 * q + 0]); v[4 * q + 0] = fma(a[q].y, b.y, v[4 * q + 0]);
        v[4 * q + 1] = fma(a[q].x, b.y, v[4 * q + 1]); v[4 * q + 1] = fma(-a[q].y, b.x, v[4 * q + 1]);
        v[4 * q + 2] = fma(a[q].x, c.x, v[4 * q + 2]); v[4 * q + 2] = fma(a[q].y, c.y, v[4 * q + 2]);
        v[4 * q + 3] = fma(a[q].x, c.y, v[4 * q + 3]); v[4 * q + 3] = fma(-a[q].y, c.x, v[4 * q + 3]);
      }
    }
    int owner = 0;
    const double tot = og::wave_multi_sum<double, 4 * R>(v, lane, &owner);
    if ((lane & (64 / (4 * R) - 1)) == 0) lds[wid * 4 * R + owner] = tot;
    __syncthreads();
    if (threadIdx.x < 4 * R) {
      double sum = 0.0;
      for (int w = 0; w < nw; ++w) sum += lds[w * 4 * R + threadIdx.x];
      const int q = threadIdx.x >> 2;
      if (i0 + q < m) part[((size_t)blockIdx.x * m + i0 + q) * 4 + (threadIdx.x & 3)] = sum;
    }
    __syncthreads();
  }
}

// out[j] = sum_b part[b*nv + j]: one wave per value, lane-strided partial sums
// then a butterfly (fixed order, so every run gives the same bits).
__global__ void k_multifinish(const int* __restrict__ skip, const double* __restrict__ part, int nblk, int nv,
                              double* __restrict__ out) {
  // the flag is read with the partials (one round trip) and tested before the store
  const int sk = *skip;
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= nv) return;
  const double s = og::wave_sum_parts(part, nblk, nv, j, lane);
  if (lane == 0 && !sk) out[j] = s;
}

// x += sum_i C_i U_i over the cycle's cumulative least-squares coefficients (the sum of
// the per-iteration updates of update_gmres, OCEAN_gmres.f90:336-341, formed once
// when the basis is about to be replaced or the solve has converged); terms with
// C_i = 0 are skipped (their U_i may not have been formed).  skip may be null.
__global__ __launch_bounds__(kBlock) void k_materialize(const int* __restrict__ skip,
                                                        const double2* const* __restrict__ Ul,
                                                        const double2* __restrict__ C, int m,
                                                        double2* __restrict__ x, size_t n) {
  if (skip && *skip) return;
  constexpr int R = 8;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    double2 xv = x[e];
    for (int i0 = 0; i0 < m; i0 += R) {
      double2 w[R], ci[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        ci[q] = C[min(i0 + q, m - 1)];
        const bool use = i0 + q < m && (ci[q].x != 0.0 || ci[q].y != 0.0);
        w[q] = use ? Ul[i0 + q][e] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (i0 + q < m && (ci[q].x != 0.0 || ci[q].y != 0.0)) {
          xv.x += ci[q].x * w[q].x; xv.y += ci[q].x * w[q].y;
          xv.x += -ci[q].y * w[q].y; xv.y += ci[q].y * w[q].x;
        }
      }
    }
    x[e] = xv;
  }
}

// Least-squares step of update_gmres on the device (OCEAN_gmres.f90:241-344).
// The Gram matrix grows by one bordered row per call (cur = 1, 2, ... after every
// restart or recycled energy), so its LDL^H factor (no pivoting) grows by one row:
//   L(j,k) = (C(j,k) - sum_{m<k} L(j,m) conj(L(k,m)) D(m)) / D(k),  k < j = cur-1
//   D(j)   =  C(j,j) - sum_{k<j} L(j,k) conj(L(j,k)) D(k)
// with D and the strictly lower triangle of L (packed row by row: L(i, k) at
// i (i - 1) / 2 + k) kept in global memory between calls; then coef = -C^-1 <au|g> by
// the two triangular solves.  One wave: lane l owns indices l, l + 64, .. (kLsSlots of
// them), and every triangular sweep is right-looking, so each of its cur steps is
// one broadcast of the finished value plus one FMA per owned index; the strictly
// lower triangle of L is staged in LDS (packed rows) when it fits, otherwise read
// from global memory.  A zero pivot leaves its 1-based index in *info (zhptrf's
// info) and the coefficients at 0.  The coefficients are also added to the cycle's
// cumulative ones C (x = x_b + sum_i C_i U_i, formed when the basis is replaced).
constexpr int kLsSlots = 4;
constexpr int kMaxGmresDepth = 64 * kLsSlots;
constexpr int kLsPacked = 3916;  // LDS-staged rows 1 .. cur-2: (cur-2)(cur-1)/2 entries, cur <= 90

__device__ __forceinline__ double2 ls_bcast(const double2 (&v)[kLsSlots], int k) {
  const int q = k >> 6, src = k & 63;
  double2 s = v[0];
#pragma unroll
  for (int i = 1; i < kLsSlots; ++i)
    if (q == i) s = v[i];
  return make_double2(__shfl(s.x, src, 64), __shfl(s.y, src, 64));
}

__global__ __launch_bounds__(64) void k_ls_update(int* __restrict__ skip, const double* __restrict__ fin,
                                                  double2* __restrict__ L, double* __restrict__ D,
                                                  double2* __restrict__ coef, double2* __restrict__ C, int cur,
                                                  int depth, int* __restrict__ info) {
  // every input is loaded before the flag is tested, so the flag, the Gram column
  // and the stored factor arrive in one round trip
  const int sk = *skip;
  __shared__ double2 Ls[kLsPacked];
  __shared__ double Dl[kMaxGmresDepth];
  const int lane = threadIdx.x;
  const int j = cur - 1;
  const int npk = j * (j - 1) / 2;  // strictly lower rows 1 .. j - 1, packed row by row
  const bool in_lds = npk <= kLsPacked;
  // C(j, k) = conj(<au_k|au_j>) for k < j (the lower triangle of the Hermitian
  // matrix whose upper column j holds <au_k|au_j>); C(j, j) = |au_j|^2; the
  // right-hand side -<au_i|g>; D(0 .. j-1)
  double2 t[kLsSlots], lj[kLsSlots], r[kLsSlots];
  double dv[kLsSlots];
#pragma unroll
  for (int q = 0; q < kLsSlots; ++q) {
    const int i = lane + 64 * q;
    t[q] = i < j ? make_double2(fin[4 * i], -fin[4 * i + 1]) : make_double2(0.0, 0.0);
    r[q] = i < cur ? make_double2(-fin[4 * i + 2], -fin[4 * i + 3]) : make_double2(0.0, 0.0);
    dv[q] = i < j ? D[i] : 0.0;
    lj[q] = make_double2(0.0, 0.0);
  }
  // rows 1 .. j - 1 of L (contiguous in the packed layout); row j is formed here and
  // kept in registers.  The first four loads per lane are issued with the others.
  double2 v0[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v0[q] = L[min(lane + 64 * q, max(npk - 1, 0))];
  if (sk) return;
  if (in_lds) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (lane + 64 * q < npk) Ls[lane + 64 * q] = v0[q];
    for (int e0 = lane + 4 * 64; e0 < npk; e0 += 4 * 64) {
      double2 v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = L[min(e0 + 64 * q, npk - 1)];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (e0 + 64 * q < npk) Ls[e0 + 64 * q] = v[q];
    }
  }
#pragma unroll
  for (int q = 0; q < kLsSlots; ++q)
    if (lane + 64 * q < j) Dl[lane + 64 * q] = dv[q];
  __syncthreads();
  // The sweeps, compiled once for L staged in LDS and once for L in global memory: a
  // run-time choice between the two address spaces would make every L read a flat
  // load that waits on the vector memory counter inside the serial sweeps.
  auto sweeps = [&](auto in_lds_tag) {
    constexpr bool kInLds = decltype(in_lds_tag)::value;
    auto Lat = [&](int i, int k) -> double2 {  // L(i, k), k < i < j
      if constexpr (kInLds) return Ls[i * (i - 1) / 2 + k];
      else return L[(size_t)i * (i - 1) / 2 + k];
    };
    // row j: after step k, t(k) is final and L(j, k) = t(k) / D(k) updates every t(kk), kk > k
    double dsum = 0.0;
    for (int k = 0; k < j; ++k) {
      const double2 tk = ls_bcast(t, k);
      const double dk = Dl[k];
      const double2 a = make_double2(tk.x / dk, tk.y / dk);
      if (lane == 0) dsum -= (a.x * a.x + a.y * a.y) * dk;
  #pragma unroll
      for (int q = 0; q < kLsSlots; ++q) {
        const int kk = lane + 64 * q;
        if (kk == k) lj[q] = a;
        if (kk > k && kk < j) {
          const double2 b = Lat(kk, k);
          // a * conj(b) * d
          t[q].x -= (a.x * b.x + a.y * b.y) * dk;
          t[q].y -= (a.y * b.x - a.x * b.y) * dk;
        }
      }
    }
  #pragma unroll
    for (int q = 0; q < kLsSlots; ++q) {
      const int k = lane + 64 * q;
      if (k < j) L[(size_t)npk + k] = lj[q];  // row j at j (j - 1) / 2
    }
    if (lane == 0) {
      const double dj = fin[4 * j] + dsum;
      D[j] = dj;
      Dl[j] = dj;
    }
    __syncthreads();
    int bad = 0;
    for (int k = 0; k <= j; ++k)
      if (Dl[k] == 0.0) { bad = k + 1; break; }
    if (bad) {
      if (lane == 0) {
        *info = bad;
        *skip = 2;  // the solve stops here: later enqueued work is dropped
      }
      for (int i = lane; i < cur; i += 64) coef[i] = make_double2(0.0, 0.0);
      return;
    }
    // coef = -C^-1 rvec: L y = -rvec, then D, then L^H coef = y (row j of L from the
    // registers of its owners)
    for (int k = 0; k < cur; ++k) {
      const double2 v = ls_bcast(r, k);
      const double2 ljk = ls_bcast(lj, k);
  #pragma unroll
      for (int q = 0; q < kLsSlots; ++q) {
        const int i = lane + 64 * q;
        if (i > k && i < cur) {
          const double2 l = i == j ? ljk : Lat(i, k);
          r[q].x -= l.x * v.x - l.y * v.y;
          r[q].y -= l.x * v.y + l.y * v.x;
        }
      }
    }
  #pragma unroll
    for (int q = 0; q < kLsSlots; ++q) {
      const int i = lane + 64 * q;
      if (i < cur) { r[q].x /= Dl[i]; r[q].y /= Dl[i]; }
    }
    for (int k = cur - 1; k > 0; --k) {
      const double2 v = ls_bcast(r, k);
  #pragma unroll
      for (int q = 0; q < kLsSlots; ++q) {
        const int i = lane + 64 * q;
        if (i < k) {
          const double2 l = k == j ? lj[q] : Lat(k, i);  // conj(l) * v
          r[q].x -= l.x * v.x + l.y * v.y;
          r[q].y -= l.x * v.y - l.y * v.x;
        }
      }
    }
  #pragma unroll
    for (int q = 0; q < kLsSlots; ++q) {
      const int i = lane + 64 * q;
      if (i < cur) {
        coef[i] = r[q];
        double2 ci = C[i];
        ci.x += r[q].x; ci.y += r[q].y;
        C[i] = ci;
      }
    }
  };
  if (in_lds) sweeps(std::true_type{});
  else sweeps(std::false_type{});
}

// pcdiv = (E - (h1_r + i(h1_i + g))) / ((E - h1)^2 + g^2)  (OCEAN_psi_min_set_prec, :2907-2994)
__global__ void k_set_prec(const int* __restrict__ skip, const double2* __restrict__ h1, double2* __restrict__ pc,
                           double E, double gprc, size_t n) {
  if (*skip) return;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const double2 hv = h1[e];
    const double dr = E - hv.x, di = -hv.y;
    const double2 ct = make_double2(dr * dr - di * di + gprc * gprc, 2.0 * dr * di);
    const double2 num = make_double2(E - hv.x, -(hv.y + gprc));
    pc[e] = og::c_div(num, ct);
  }
}

// a .* psi for the sfact copy (OCEAN_energies_sfact_copy, OCEAN_energies.f90:40-53)
__global__ void k_mask(double2* __restrict__ y, const double2* __restrict__ x, const double* __restrict__ m, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    double2 v = x[e];
    v.x *= m[e]; v.y *= m[e];
    y[e] = v;
  }
}

using og::GmCtl;
using og::GmFin;
using og::kResX;
using og::kResU;
using og::kDotU;

__global__ void k_gm_reset(GmCtl* __restrict__ c) {
  if (threadIdx.x == 0) {
    c->done = 0;
    c->conv_it = 0;
  }
}

// y += (ar + i ai) x unless the solve has stopped (recycling step of make_recycle)
__global__ void k_axpy_gated(const int* __restrict__ skip, double ar, double ai, const double2* __restrict__ x,
                             double2* __restrict__ y, size_t n) {
  if (*skip) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 xv = x[i];
    double2 yv = y[i];
    yv.x += ar * xv.x; yv.y += ar * xv.y;
    yv.x += -ai * xv.y; yv.y += ai * xv.x;
    y[i] = yv;
  }
}

// The residual kernels write the sums of each workgroup (v[3], block-summed in thread
// 0) as partials; og::gm_finish sums them in a fixed order, in k_resid_fin or in the
// next action's projection.  (A last-workgroup-to-arrive finish inside the kernels
// serialised a device-scope atomic per workgroup.)
__device__ __forceinline__ void resid_part(double (&v)[3], double* __restrict__ part) {
  __shared__ double lds[4 * 3];
  og::block_sum<3>(v, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = v[0]; part[3 * blockIdx.x + 1] = v[1]; part[3 * blockIdx.x + 2] = v[2];
  }
}

__global__ __launch_bounds__(256) void k_resid_fin(const GmFin F) { og::gm_finish(F); }

// g = ax - rhs for one element and its term of |g|^2
__device__ __forceinline__ double2 resid_g(size_t e, double2 a, const double2* __restrict__ rhs,
                                           double2* __restrict__ g, double& gg) {
  const double2 b = rhs[e];
  const double2 gv = make_double2(a.x - b.x, a.y - b.y);
  g[e] = gv;
  gg = fma(gv.x, gv.x, gg); gg = fma(gv.y, gv.y, gg);
  return b;
}

// <b|v> term of one element into (vr, vi)
__device__ __forceinline__ void dot_term(double2 b, double2 xv, double& vr, double& vi) {
  vr = fma(b.x, xv.x, vr); vr = fma(b.y, xv.y, vr);
  vi = fma(b.x, xv.y, vi); vi = fma(-b.y, xv.x, vi);
}

// The residual step with x materialised (after an action has recomputed ax from x):
// g = ax - rhs and the partials of |g|^2 and <rhs|x> (k_resid_fin, kResX).
__global__ __launch_bounds__(kBlock) void k_resid_step(GmCtl* __restrict__ c, const double2* __restrict__ ax,
                                                       const double2* __restrict__ rhs, const double2* __restrict__ x,
                                                       double2* __restrict__ g, size_t n, double* __restrict__ part) {
  if (c->done) return;
  double v[3] = {0, 0, 0};
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const double2 b = resid_g(e, ax[e], rhs, g, v[0]);
    dot_term(b, x[e], v[1], v[2]);
  }
  resid_part(v, part);
}

// The first basis vector of a cycle, U = g . pc (OCEAN_gmres.f90:760-766), and the
// partials of <rhs|U> (k_resid_fin, kDotU).
__global__ __launch_bounds__(kBlock) void k_pc_apply(GmCtl* __restrict__ c, double2* __restrict__ u,
                                                     const double2* __restrict__ g, const double2* __restrict__ pc,
                                                     const double2* __restrict__ rhs, size_t n,
                                                     double* __restrict__ part) {
  if (c->done) return;
  double v[3] = {0, 0, 0};
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const double2 uv = og::cmul(g[e], pc[e]);
    u[e] = uv;
    dot_term(rhs[e], uv, v[1], v[2]);
  }
  resid_part(v, part);
}

// ax += sum_i c_i AU_i (update_gmres, OCEAN_gmres.f90:336-341; eight rows per step
// with their loads in flight, the sum in i order), then in the same pass the residual
// step: g = ax - rhs, |g|^2, the next basis vector unext = g . pc (nullable) and the
// partials of <rhs|unext>; k_resid_fin (kResU) forms <rhs|x> from the cumulative
// coefficients C (already holding this update) without forming x.
__global__ __launch_bounds__(kBlock) void k_axpy_resid(GmCtl* __restrict__ c, const double2* const* __restrict__ AUl,
                                                       const double2* __restrict__ coef, int m,
                                                       double2* __restrict__ ax, const double2* __restrict__ rhs,
                                                       double2* __restrict__ g, const double2* __restrict__ pc,
                                                       double2* __restrict__ unext, size_t n,
                                                       double* __restrict__ part) {
  // the flag is read with the first loads and tested before the first store
  const int done = c->done;
  constexpr int R = 8;
  double v[3] = {0, 0, 0};
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const double2 z = make_double2(0.0, 0.0);
  // two elements (e0 and e0 + stride, the grid-stride order) per step, so that twice
  // the rows are in flight
  for (size_t e0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < n; e0 += 2 * stride) {
    const size_t e1 = e0 + stride;
    const bool has1 = e1 < n;
    double2 av0 = ax[e0], av1 = has1 ? ax[e1] : z;
    for (int i0 = 0; i0 < m; i0 += R) {
      double2 a0[R], a1[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const double2* row = AUl[min(i0 + q, m - 1)];
        a0[q] = row[e0];
        a1[q] = has1 ? row[e1] : z;
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (i0 + q < m) {
          const double2 ci = coef[i0 + q];
          av0.x += ci.x * a0[q].x; av0.y += ci.x * a0[q].y;
          av0.x += -ci.y * a0[q].y; av0.y += ci.y * a0[q].x;
          av1.x += ci.x * a1[q].x; av1.y += ci.x * a1[q].y;
          av1.x += -ci.y * a1[q].y; av1.y += ci.y * a1[q].x;
        }
      }
    }
    if (done) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const size_t e = h ? e1 : e0;
      if (h && !has1) break;
      const double2 av = h ? av1 : av0;
      ax[e] = av;
      const double2 b = resid_g(e, av, rhs, g, v[0]);
      if (unext) {
        const double2 gv = make_double2(av.x - b.x, av.y - b.y);
        const double2 uv = og::cmul(gv, pc[e]);
        unext[e] = uv;
        dot_term(b, uv, v[1], v[2]);
      }
    }
  }
  if (done) return;
  resid_part(v, part);
}

// After the basis coefficients have been folded into x (k_materialize): <rhs|x_b> +=
// sum_i C_i <rhs|U_i> and C = 0.  One wave; skip may be null.
__global__ __launch_bounds__(64) void k_fold_c(const int* __restrict__ skip, GmCtl* __restrict__ c,
                                               double2* __restrict__ C, const double2* __restrict__ rU, int m) {
  if (skip && *skip) return;
  const int lane = threadIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int i = lane; i < m; i += 64) {
    const double2 ci = C[i];
    if (ci.x != 0.0 || ci.y != 0.0) {
      const double2 u = rU[i];
      ar += ci.x * u.x - ci.y * u.y;
      ai += ci.x * u.y + ci.y * u.x;
    }
    C[i] = make_double2(0.0, 0.0);
  }
  ar = og::wave_sum(ar);
  ai = og::wave_sum(ai);
  if (lane == 0) {
    c->rxr += ar;
    c->rxi += ai;
  }
}

}  // namespace

// ===========================================================================
extern "C" {

const char* ocean_gpu_strerror(int code) {
  switch (code) {
    case OCEAN_GPU_OK: return "ok";
    case OCEAN_GPU_ERR_NAN: return "NaN detected in the Haydock recursion";
    case OCEAN_GPU_ERR_ARG: return "invalid argument or inconsistent sizes";
    case OCEAN_GPU_ERR_STATE: return "required state (bloch, W, multiplet or energies) not loaded";
    case OCEAN_GPU_ERR_HIP: return "HIP runtime error";
    case OCEAN_GPU_ERR_NOMEM: return "device allocation failed";
    case OCEAN_GPU_ERR_UNSUPPORTED: return "unsupported option";
    case OCEAN_GPU_ERR_NOCONV: return "GMRES did not converge within its iteration cap";
    default: return code > 0 ? "GMRES least-squares pivot is zero (zhptrf info)" : "unknown error";
  }
}

}  // extern "C"

namespace {

int create_impl(const ocean_bse_desc* d, int nrhs, ocean_gpu_handle** out) {
  if (!d || !out) return OCEAN_GPU_ERR_ARG;
  *out = nullptr;
  if (nrhs < 1) return OCEAN_GPU_ERR_ARG;
  if (d->num_bands <= 0 || d->nspn < 1 || d->nspn > 2 || d->lc < 0 || d->lc > 1) return OCEAN_GPU_ERR_ARG;
  for (int i = 0; i < 3; ++i)
    if (d->kmesh[i] <= 0 || d->xmesh[i] <= 0 || !dim_ok(d->kmesh[i])) return OCEAN_GPU_ERR_ARG;
  if (d->lmin < 0 || d->lmax < d->lmin || d->lmax - d->lmin >= 8) return OCEAN_GPU_ERR_ARG;
  if (d->lvl > d->lmin || d->lvh < d->lmax) return OCEAN_GPU_ERR_ARG;
  if (d->celvol <= 0.0) return OCEAN_GPU_ERR_ARG;
  H* h = new H();
  h->d = *d;
  h->nb = d->num_bands;
  h->km[0] = d->kmesh[0]; h->km[1] = d->kmesh[1]; h->km[2] = d->kmesh[2];
  h->nk = d->kmesh[0] * d->kmesh[1] * d->kmesh[2];
  h->nx = d->xmesh[0] * d->xmesh[1] * d->xmesh[2];
  h->na0 = 4 * (2 * d->lc + 1);
  h->nrhs = nrhs;
  h->na = nrhs * h->na0;
  h->ns = d->nspn;
  // the right-hand sides run as na alphas of the action kernels, whose alpha counts
  // per spin are 2, 4, 6 and 12, and 24 on one spin (matrix-core adjoint only)
  {
    const int acnt = h->ns == 2 ? h->na / 2 : h->na;
    if (nrhs > 1 && (acnt > 24 || na_tmpl(acnt) != acnt || (acnt == 24 && h->ns != 1))) {
      delete h;
      return OCEAN_GPU_ERR_UNSUPPORTED;
    }
  }
  h->ldu = (h->nx + 33) & ~1;  // even, so FP32 rows pair up into 16-byte loads
  h->ldphi = h->nx * h->na + 32;
  h->ldT = h->nx * h->nb + 32;
  h->precision = d->precision;
  {
    int dev = 0, lds = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
      h->lds_max = (size_t)lds;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      h->ncu = ncu;
  }
  // the adjoint keeps one k's result (nalpha x nb complex) and one x-split row in LDS
  if (h->nb > 512 || (size_t)(h->na + 1) * h->nb * sizeof(double2) > h->lds_max) {
    delete h;
    return OCEAN_GPU_ERR_ARG;
  }
  int P = 0;
  for (int l = d->lmin; l <= d->lmax; ++l) {
    if (d->nproj[l - d->lmin] < 0) { delete h; return OCEAN_GPU_ERR_ARG; }
    P += (2 * l + 1) * d->nproj[l - d->lmin];
  }
  h->P = P;
  if (h->na * std::max(P, 1) > og::kMaxAlpha * 64) { delete h; return OCEAN_GPU_ERR_ARG; }
  if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
    ocean_gpu_destroy(h);
    return OCEAN_GPU_ERR_HIP;
  }
  int rc = OCEAN_GPU_OK;
  const size_t nv = h->nvec();
  if (rc == 0) rc = dmalloc(&h->energy, nv);
  if (rc == 0) rc = dmalloc(&h->allow, nv);
  if (rc == 0) rc = dmalloc(&h->red, (size_t)kRedBlocks * 8 + 64);
  if (rc == 0) rc = dmalloc(&h->tmp_in, nv);
  if (rc == 0) rc = dmalloc(&h->tmp_out, nv);
  if (rc == 0) rc = ensure_precision_buffers(h);
  if (rc != 0) {
    ocean_gpu_destroy(h);
    return rc;
  }
  *out = h;
  return OCEAN_GPU_OK;
}

}  // namespace

extern "C" {

int ocean_gpu_create(const ocean_bse_desc* d, ocean_gpu_handle** out) { return create_impl(d, 1, out); }

int ocean_gpu_create_multi(const ocean_bse_desc* d, int32_t nrhs, ocean_gpu_handle** out) {
  return create_impl(d, nrhs, out);
}

int ocean_gpu_nrhs(ocean_gpu_handle* h, int32_t* nrhs) {
  if (!h || !nrhs) return OCEAN_GPU_ERR_ARG;
  *nrhs = h->nrhs;
  return OCEAN_GPU_OK;
}

int ocean_gpu_destroy(ocean_gpu_handle* h) {
  if (!h) return OCEAN_GPU_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  dfree(h->u64); dfree(h->u32); dfree(h->uT64); dfree(h->uT32); dfree(h->W64); dfree(h->W32); dfree(h->phi64); dfree(h->phi32);
  dfree(h->w_i0); dfree(h->w_w0); dfree(h->w_w1);
  dfree(h->obf_u); dfree(h->obf_uT); dfree(h->obf_f); dfree(h->obf_fT); dfree(h->obf_phs); dfree(h->obf_beta);
  dfree(h->cproj); dfree(h->Kmat); dfree(h->apart); dfree(h->hmult); dfree(h->Smat);
  dfree(h->amp); dfree(h->mcount);
  dfree(h->energy); dfree(h->allow); dfree(h->red); dfree(h->scal); dfree(h->tmp_in); dfree(h->tmp_out);
  for (auto& v : h->vecs) dfree(v);
  if (h->host_pin) (void)hipHostFree(h->host_pin);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return OCEAN_GPU_OK;
}

int ocean_gpu_set_precision(ocean_gpu_handle* h, int32_t precision) {
  if (!h || (precision != OCEAN_GPU_FP64 && precision != OCEAN_GPU_MIXED)) return OCEAN_GPU_ERR_ARG;
  if (precision == h->precision) return OCEAN_GPU_OK;
  if (h->obf_on) return OCEAN_GPU_ERR_STATE;  // the OBF form is FP64 only
  h->precision = precision;
  OG_TRY(ensure_precision_buffers(h));
  const size_t nu = h->nu();
  const size_t nw = (size_t)h->nk * h->nx;
  // FP64 -> FP32 conversion of loaded state; the FP32 state cannot be promoted
  // back (an FP64 reload is required), so the FP64 copy is kept when it exists.
  if (precision == OCEAN_GPU_MIXED) {
    if (h->u64_ok && !h->u32_ok) {
      og::k_cast<float><<<4096, kBlock, 0, h->stream>>>((const double*)h->u64, (float*)h->u32, 2 * nu);
      og::k_cast<float><<<4096, kBlock, 0, h->stream>>>((const double*)h->uT64, (float*)h->uT32,
                                                       2 * h->nuT());
      h->u32_ok = true;
    }
    if (h->w64_ok && !h->w32_ok) {
      og::k_cast<float><<<4096, kBlock, 0, h->stream>>>(h->W64, h->W32, nw);
      h->w32_ok = true;
    }
  }
  OG_CHECK(hipGetLastError());
  OG_CHECK(hipStreamSynchronize(h->stream));
  return OCEAN_GPU_OK;
}

static int load_u_common(H* h, const double* re, const double* im, const int32_t* xoff, bool dev = false) {
  if (!h || !re || !im) return OCEAN_GPU_ERR_ARG;
  const size_t nu = (size_t)h->ns * h->nk * h->nb * h->nx;
  double *dre = nullptr, *dim_ = nullptr;
  int* dxo = nullptr;
  int rc = OCEAN_GPU_OK;
  if (dev) {
    // caller's device planes are read in place
    dre = const_cast<double*>(re);
    dim_ = const_cast<double*>(im);
  } else {
    OG_TRY(dmalloc(&dre, nu));
    if (dmalloc(&dim_, nu) != 0) { dfree(dre); return OCEAN_GPU_ERR_NOMEM; }
    if (hipMemcpy(dre, re, nu * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dim_, im, nu * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0 && xoff) {
    rc = dmalloc(&dxo, (size_t)3 * h->nx);
    if (rc == 0 && hipMemcpy(dxo, xoff, 3 * h->nx * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0) {
    // the FP64 form keeps an FP64 copy; the mixed form keeps FP32 only
    dim3 g((h->nb + 31) / 32, (h->nx + 31) / 32, h->ns * h->nk);
    if (h->precision == OCEAN_GPU_MIXED) {
      if (xoff)
        og::k_load_u_phase<float2><<<g, 256, 0, h->stream>>>(dre, dim_, dxo, h->u32, h->nb, h->nk, h->nx,
            h->km[0], h->km[1], h->km[2], h->d.xmesh[0], h->d.xmesh[1], h->d.xmesh[2], h->ldu);
      else
        og::k_load_u<float2><<<g, 256, 0, h->stream>>>(dre, dim_, h->u32, h->nb, h->nk, h->nx, h->ldu);
    } else {
      if (xoff)
        og::k_load_u_phase<double2><<<g, 256, 0, h->stream>>>(dre, dim_, dxo, h->u64, h->nb, h->nk, h->nx,
            h->km[0], h->km[1], h->km[2], h->d.xmesh[0], h->d.xmesh[1], h->d.xmesh[2], h->ldu);
      else
        og::k_load_u<double2><<<g, 256, 0, h->stream>>>(dre, dim_, h->u64, h->nb, h->nk, h->nx, h->ldu);
    }
    const size_t tot = (size_t)h->ns * h->nk * h->nx * h->nb;
    const int gT = (int)std::min<size_t>((tot + 255) / 256, 65536);
    if (h->precision == OCEAN_GPU_MIXED)
      og::k_load_uT<float2><<<gT, 256, 0, h->stream>>>(dre, dim_, xoff ? dxo : nullptr, h->uT32, h->nb, h->nk,
          h->nx, h->ns, h->km[0], h->km[1], h->km[2], h->d.xmesh[0], h->d.xmesh[1], h->d.xmesh[2], h->ldT);
    else
      og::k_load_uT<double2><<<gT, 256, 0, h->stream>>>(dre, dim_, xoff ? dxo : nullptr, h->uT64, h->nb, h->nk,
          h->nx, h->ns, h->km[0], h->km[1], h->km[2], h->d.xmesh[0], h->d.xmesh[1], h->d.xmesh[2], h->ldT);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  if (!dev) { dfree(dre); dfree(dim_); }
  dfree(dxo);
  if (rc == 0) {
    if (h->precision == OCEAN_GPU_MIXED) { h->u32_ok = true; h->u64_ok = false; }
    else { h->u64_ok = true; h->u32_ok = false; }
    h->obf_on = false;
  }
  return rc;
}

int ocean_gpu_load_bloch(ocean_gpu_handle* h, const double* u_re, const double* u_im) {
  return load_u_common(h, u_re, u_im, nullptr);
}

int ocean_gpu_load_bloch_dev(ocean_gpu_handle* h, const double* u_re, const double* u_im) {
  return load_u_common(h, u_re, u_im, nullptr, true);
}

int ocean_gpu_load_bloch_raw(ocean_gpu_handle* h, const double* u_re, const double* u_im,
                             const int32_t* xoff) {
  if (!xoff) return OCEAN_GPU_ERR_ARG;
  return load_u_common(h, u_re, u_im, xoff);
}

int ocean_gpu_load_obf(ocean_gpu_handle* h, int32_t num_obf, const double* obf_re, const double* obf_im,
                       const double* obf2u_re, const double* obf2u_im, const int32_t* xoff) {
  if (!h || num_obf <= 0 || !obf_re || !obf_im || !obf2u_re || !obf2u_im || !xoff) return OCEAN_GPU_ERR_ARG;
  // OCEAN_obf_load handles nspn = 1 only (OCEAN_obf.f90:669-673); the form is FP64
  if (h->ns != 1 || h->precision != OCEAN_GPU_FP64) return OCEAN_GPU_ERR_STATE;
  if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  const int nb = h->nb, nk = h->nk, nx = h->nx, no = num_obf;
  h->obf_on = false;
  dfree(h->obf_u); dfree(h->obf_uT); dfree(h->obf_f); dfree(h->obf_fT); dfree(h->obf_phs); dfree(h->obf_beta);
  h->no = no;
  h->ldu_o = (no + 33) & ~1;
  h->ldT_o = no * nb + 32;
  h->ldb = no * h->na + 32;
  const size_t n2u = (size_t)nb * no * nk, nf = (size_t)no * nx;
  OG_TRY(dmalloc(&h->obf_u, (size_t)nk * nb * h->ldu_o));
  OG_TRY(dmalloc(&h->obf_uT, (size_t)nk * h->ldT_o));
  OG_TRY(dmalloc(&h->obf_f, nf));
  OG_TRY(dmalloc(&h->obf_fT, nf));
  OG_TRY(dmalloc(&h->obf_phs, (size_t)nk * nx));
  OG_TRY(dmalloc(&h->obf_beta, (size_t)nk * h->ldb));
  OG_CHECK(hipMemsetAsync(h->obf_u, 0, (size_t)nk * nb * h->ldu_o * sizeof(double2), h->stream));
  OG_CHECK(hipMemsetAsync(h->obf_uT, 0, (size_t)nk * h->ldT_o * sizeof(double2), h->stream));
  OG_CHECK(hipMemsetAsync(h->obf_beta, 0, (size_t)nk * h->ldb * sizeof(double2), h->stream));
  double* dpl = nullptr;
  int* dxo = nullptr;
  const size_t npl = std::max(n2u, nf);
  OG_TRY(dmalloc(&dpl, 2 * npl));
  int rc = dmalloc(&dxo, (size_t)3 * nx);
  if (rc == 0 && (hipMemcpy(dpl, obf2u_re, n2u * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(dpl + npl, obf2u_im, n2u * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(dxo, xoff, 3 * (size_t)nx * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    og::k_obf_load2u<<<grid_for(n2u), kBlock, 0, h->stream>>>(dpl, dpl + npl, h->obf_u, h->obf_uT, nb, no, nk,
                                                             h->ldu_o, h->ldT_o);
    og::k_obf_phase<<<grid_for((size_t)nk * nx), kBlock, 0, h->stream>>>(dxo, h->obf_phs, nk, nx, h->km[0],
        h->km[1], h->km[2], h->d.xmesh[0], h->d.xmesh[1], h->d.xmesh[2]);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0 && (hipMemcpy(dpl, obf_re, nf * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(dpl + npl, obf_im, nf * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    og::k_obf_loadf<<<grid_for(nf), kBlock, 0, h->stream>>>(dpl, dpl + npl, h->obf_f, h->obf_fT, no, nx);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(dpl);
  dfree(dxo);
  if (rc == 0) h->obf_on = true;
  return rc;
}

static int finish_w(H* h, const double* dW64) {
  const size_t nw = (size_t)h->nk * h->nx;
  if (h->precision == OCEAN_GPU_MIXED) {
    og::k_cast<float><<<grid_for(nw), kBlock, 0, h->stream>>>(dW64, h->W32, nw);
    OG_CHECK(hipGetLastError());
    h->w32_ok = true;
    h->w64_ok = false;
  } else {
    OG_CHECK(hipMemcpyAsync(h->W64, dW64, nw * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    h->w64_ok = true;
    h->w32_ok = false;
  }
  OG_CHECK(hipStreamSynchronize(h->stream));
  return OCEAN_GPU_OK;
}

int ocean_gpu_load_w(ocean_gpu_handle* h, const double* W) {
  if (!h || !W) return OCEAN_GPU_ERR_ARG;
  const size_t nw = (size_t)h->nk * h->nx;
  double* dW = nullptr;
  OG_TRY(dmalloc(&dW, nw));
  int rc = OCEAN_GPU_OK;
  if (hipMemcpy(dW, W, nw * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) rc = finish_w(h, dW);
  dfree(dW);
  return rc;
}

int ocean_gpu_build_w(ocean_gpu_handle* h, const double* rdist, const double* rtab, const double* ptab,
                      int32_t nptab, double epsilon0, double scale) {
  if (!h || !rdist || !rtab || !ptab || nptab < 2 || epsilon0 == 0.0) return OCEAN_GPU_ERR_ARG;
  for (int i = 1; i < nptab; ++i)
    if (!(rtab[i] > rtab[i - 1])) return OCEAN_GPU_ERR_ARG;
  const size_t nw = (size_t)h->nk * h->nx;
  if (!h->w_i0) { OG_TRY(dmalloc(&h->w_i0, nw)); OG_TRY(dmalloc(&h->w_w0, nw)); OG_TRY(dmalloc(&h->w_w1, nw)); }
  double *dr = nullptr, *dt = nullptr, *dW = nullptr;
  OG_TRY(dmalloc(&dr, nw));
  int rc = dmalloc(&dt, 2 * (size_t)nptab);
  if (rc == 0) rc = dmalloc(&dW, nw);
  if (rc == 0 && (hipMemcpy(dr, rdist, nw * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(dt, rtab, nptab * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(dt + nptab, ptab, nptab * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    og::k_build_w<<<grid_for(nw), kBlock, 0, h->stream>>>(dr, dt, dt + nptab, nptab, 1.0 / epsilon0, scale, nw,
                                                         dW, h->w_i0, h->w_w0, h->w_w1);
    if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0) rc = finish_w(h, dW);
  if (rc == 0) h->nptab = nptab;
  dfree(dr); dfree(dt); dfree(dW);
  return rc;
}

int ocean_gpu_load_mult(ocean_gpu_handle* h, const double* mpcr, const double* mpci, const double* mpm,
                        const double* mhr, const double* mhi, const double* somelr, const double* someli,
                        int32_t have_so) {
  if (!h || !mpcr || !mpci || !mpm || !mhr || !mhi || !somelr || !someli) return OCEAN_GPU_ERR_ARG;
  const ocean_bse_desc& d = h->d;
  // the blocks of one right-hand side (na0 alphas); with several right-hand sides K
  // is block diagonal over them and S is applied per right-hand side
  const int nb = h->nb, nk = h->nk, ns = h->ns, na = h->na0, P = h->P, lc = d.lc;
  const int lmin = d.lmin, lmax = d.lmax, nl = lmax - lmin + 1;
  int npmax = 0;
  for (int l = lmin; l <= lmax; ++l) npmax = std::max(npmax, d.nproj[l - lmin]);
  const int nm = 2 * lmax + 1;
  // compact projector index p = (l, m, nu) in the reference's ip order (OCEAN_multiplet.f90:451-462)
  std::vector<int> p_l, p_m, p_nu, p_off(nl + 1, 0);
  for (int l = lmin; l <= lmax; ++l) {
    p_off[l - lmin] = (int)p_l.size();
    for (int m = -l; m <= l; ++m)
      for (int nu = 0; nu < d.nproj[l - lmin]; ++nu) { p_l.push_back(l); p_m.push_back(m); p_nu.push_back(nu); }
  }
  // c[s][k][p][n]
  std::vector<double2> c((size_t)ns * nk * P * nb);
  for (int s = 0; s < ns; ++s)
    for (int k = 0; k < nk; ++k)
      for (int p = 0; p < P; ++p) {
        const int li = p_l[p] - lmin, mi = p_m[p] + lmax, nu = p_nu[p];
        const size_t base = ((((size_t)s * nl + li) * nm + mi) * npmax + nu) * nk + k;
        for (int n = 0; n < nb; ++n) {
          const size_t ref = base * nb + n;
          c[(((size_t)s * nk + k) * P + p) * nb + n] = make_double2(mpcr[ref], mpci[ref]);
        }
      }
  // K = -mpm (direct, diagonal in (l,m,alpha)) + X (exchange), row i = a*P + p
  const int nj = na * P;
  std::vector<cplx> K((size_t)nj * nj, cplx(0, 0));
  for (int a = 0; a < na; ++a)
    for (int p = 0; p < P; ++p)
      for (int q = 0; q < P; ++q) {
        if (p_l[p] != p_l[q] || p_m[p] != p_m[q]) continue;
        const int li = p_l[p] - lmin;
        const double v = mpm[((size_t)li * npmax + p_nu[q]) * npmax + p_nu[p]];  // mpm(mu=nu_p, nu=nu_q, l)
        K[(size_t)(a * P + p) * nj + (a * P + q)] -= v;
      }
  size_t jb = 0;
  for (int l = d.lvl; l <= d.lvh; ++l) {
    const int npl = (l >= lmin && l <= lmax) ? d.nproj[l - lmin] : 0;
    const int mham = 4 * (2 * lc + 1) * (2 * l + 1) * npl;
    if (l >= lmin && l <= lmax) {
      for (int ii = 0; ii < mham; ++ii) {
        const int nu_i = ii % npl, m_i = (ii / npl) % (2 * l + 1) - l, a_i = ii / (npl * (2 * l + 1));
        const int p_i = p_off[l - lmin] + (m_i + l) * npl + nu_i;
        for (int jj = 0; jj < mham; ++jj) {
          const int nu_j = jj % npl, m_j = (jj / npl) % (2 * l + 1) - l, a_j = jj / (npl * (2 * l + 1));
          const int p_j = p_off[l - lmin] + (m_j + l) * npl + nu_j;
          const size_t j1 = jb + (size_t)jj + (size_t)ii * mham;  // mh(jj, ii)
          K[(size_t)(a_i * P + p_i) * nj + (a_j * P + p_j)] += cplx(mhr[j1], mhi[j1]);
        }
      }
    }
    jb += (size_t)mham * mham;
  }
  const int njt = h->na * P;  // = nrhs nj
  std::vector<double2> Kd((size_t)njt * njt, make_double2(0.0, 0.0));
  for (int r = 0; r < h->nrhs; ++r)
    for (int i = 0; i < nj; ++i)
      for (int j = 0; j < nj; ++j)
        Kd[(size_t)(r * nj + i) * njt + r * nj + j] = make_double2(K[(size_t)i * nj + j].real(),
                                                                   K[(size_t)i * nj + j].imag());
  std::vector<double2> S((size_t)na * na);
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < na; ++b) S[a * na + b] = make_double2(somelr[(size_t)b * na + a], someli[(size_t)b * na + a]);
  dfree(h->cproj); dfree(h->Kmat); dfree(h->apart); dfree(h->hmult); dfree(h->Smat);
  dfree(h->amp); dfree(h->mcount);
  OG_TRY(dmalloc(&h->cproj, c.size()));
  OG_TRY(dmalloc(&h->amp, (size_t)njt));
  OG_TRY(dmalloc(&h->mcount, (size_t)1));
  OG_CHECK(hipMemset(h->mcount, 0, sizeof(int)));
  OG_TRY(dmalloc(&h->Kmat, Kd.size()));
  OG_TRY(dmalloc(&h->apart, (size_t)nk * njt));
  OG_TRY(dmalloc(&h->hmult, (size_t)njt));
  OG_TRY(dmalloc(&h->Smat, S.size()));
  OG_CHECK(hipMemcpy(h->cproj, c.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice));
  OG_CHECK(hipMemcpy(h->Kmat, Kd.data(), Kd.size() * sizeof(double2), hipMemcpyHostToDevice));
  OG_CHECK(hipMemcpy(h->Smat, S.data(), S.size() * sizeof(double2), hipMemcpyHostToDevice));
  h->have_so = have_so ? 1 : 0;
  h->has_mult = true;
  return OCEAN_GPU_OK;
}

int ocean_gpu_mult_setup(ocean_gpu_handle* h, const double* slater, double scfac, int32_t tdlda, double xi_ev,
                         const double* life_time_ev, double* mhr, double* mhi, double* somelr, double* someli,
                         int32_t* have_so) {
  if (!h || !slater || !mhr || !mhi || !somelr || !someli || !have_so) return OCEAN_GPU_ERR_ARG;
  const ocean_bse_desc& d = h->d;
  const int lc = d.lc, na = h->na0;
  // per valence l: the block size and where its Slater matrices and block start
  std::vector<int> npl_l, nsl_l;
  std::vector<size_t> jb_l, sb_l;
  size_t jtot = 0, stot = 0;
  for (int l = d.lvl; l <= d.lvh; ++l) {
    const int npl = (l >= d.lmin && l <= d.lmax) ? d.nproj[l - d.lmin] : 0;
    const int kfh = std::min(2 * lc, 2 * l);
    const int nF = tdlda ? 1 : (kfh >= 2 ? kfh / 2 : 0), nG = (lc + l - std::abs(lc - l)) / 2 + 1;
    const size_t mham = (size_t)4 * (2 * lc + 1) * (2 * l + 1) * npl;
    npl_l.push_back(npl);
    nsl_l.push_back((nF + nG) * npl * npl);
    jb_l.push_back(jtot);
    sb_l.push_back(stot);
    jtot += mham * mham;
    stot += (size_t)(nF + nG) * npl * npl;
  }
  double *dsl = nullptr, *dmh = nullptr, *dso = nullptr;
  OG_TRY(dmalloc(&dsl, std::max<size_t>(stot, 1)));
  int rc = dmalloc(&dmh, 2 * std::max<size_t>(jtot, 1));
  if (rc == 0) rc = dmalloc(&dso, 2 * (size_t)na * na);
  if (rc == 0 && stot && hipMemcpy(dsl, slater, stot * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = OCEAN_GPU_ERR_HIP;
  // nbsemhsetup2 works in eV; the module converts the blocks to Hartree (OCEAN_multiplet.f90:388-389)
  for (size_t i = 0; rc == 0 && i < npl_l.size(); ++i) {
    const int npl = npl_l[i];
    if (npl == 0) continue;
    const int l = d.lvl + (int)i;
    const size_t mham = (size_t)4 * (2 * lc + 1) * (2 * l + 1) * npl;
    og::k_mham<<<grid_for(mham * mham), kBlock, 0, h->stream>>>(lc, l, npl, dsl + sb_l[i], scfac, kEV2Hartree,
                                                               tdlda ? 1 : 0,
                                                               dmh + jb_l[i], dmh + jtot + jb_l[i]);
    if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  const double xi = xi_ev * kEV2Hartree;
  // core_broaden.ipt: lifetimes of the j = lc + 1/2 and lc - 1/2 levels (OCEAN_multiplet.f90:613-627)
  double l_alpha = 0.0, l_beta = 0.0;
  if (life_time_ev) {
    const double t1 = life_time_ev[0] * kEV2Hartree, t2 = life_time_ev[1] * kEV2Hartree;
    l_alpha = 2.0 / (double)(2 * lc + 1) * (t1 - t2);
    l_beta = (double)(lc + 1) / (double)(2 * lc + 1) * t1 + (double)lc / (double)(2 * lc + 1) * t2;
  }
  if (rc == 0) {
    og::k_soprep<<<1, kBlock, 0, h->stream>>>(lc, xi, l_alpha, l_beta, dso, dso + (size_t)na * na);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0 && jtot &&
      (hipMemcpy(mhr, dmh, jtot * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(mhi, dmh + jtot, jtot * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0 && (hipMemcpy(somelr, dso, (size_t)na * na * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(someli, dso + (size_t)na * na, (size_t)na * na * sizeof(double),
                            hipMemcpyDeviceToHost) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  *have_so = (xi >= 1.0e-12 || life_time_ev) ? 1 : 0;  // soprep's switch, OCEAN_multiplet.f90:612-615
  dfree(dsl); dfree(dmh); dfree(dso);
  return rc;
}

int ocean_gpu_load_energies(ocean_gpu_handle* h, const double* e_re, const double* e_im, const double* allow) {
  if (!h || !e_re || !allow) return OCEAN_GPU_ERR_ARG;
  // one right-hand side's arrays; every right-hand side of the handle sees the same
  const size_t nv0 = h->nvec0();
  OG_TRY(put_planes(h, e_re, e_im, h->energy, nv0));
  OG_CHECK(hipMemcpy(h->allow, allow, nv0 * sizeof(double), hipMemcpyHostToDevice));
  for (int r = 1; r < h->nrhs; ++r) {
    OG_CHECK(hipMemcpy(h->energy + r * nv0, h->energy, nv0 * sizeof(double2), hipMemcpyDeviceToDevice));
    OG_CHECK(hipMemcpy(h->allow + r * nv0, h->allow, nv0 * sizeof(double), hipMemcpyDeviceToDevice));
  }
  h->has_e = true;
  return OCEAN_GPU_OK;
}

int ocean_gpu_xact(ocean_gpu_handle* h, const double* psi_re, const double* psi_im, double* hpsi_re,
                   double* hpsi_im, int32_t backwards, const int32_t* hflag) {
  if (!h || !psi_re || !psi_im || !hpsi_re || !hpsi_im) return OCEAN_GPU_ERR_ARG;
  const int f = flags_of(h, backwards, hflag);
  OG_TRY(check_ready(h, f));
  const size_t nv = h->nvec();
  OG_TRY(put_planes(h, psi_re, psi_im, h->tmp_in, nv));
  OG_TRY(xact_dev(h, h->tmp_in, h->tmp_out, f));
  return get_planes(h, h->tmp_out, hpsi_re, hpsi_im, nv);
}

// ---- vectors ----
int ocean_gpu_vec_new(ocean_gpu_handle* h, int32_t* id) {
  if (!h || !id) return OCEAN_GPU_ERR_ARG;
  double2* p = nullptr;
  OG_TRY(dmalloc(&p, h->nvec()));
  og::k_fill<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(p, make_double2(0.0, 0.0), h->nvec());
  OG_CHECK(hipGetLastError());
  for (size_t i = 0; i < h->vecs.size(); ++i)
    if (!h->vecs[i]) { h->vecs[i] = p; *id = (int32_t)i; return OCEAN_GPU_OK; }
  h->vecs.push_back(p);
  *id = (int32_t)(h->vecs.size() - 1);
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_free(ocean_gpu_handle* h, int32_t id) {
  if (!h || !vec_ptr(h, id)) return OCEAN_GPU_ERR_ARG;
  (void)hipStreamSynchronize(h->stream);
  dfree(h->vecs[id]);
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_put(ocean_gpu_handle* h, int32_t id, const double* re, const double* im) {
  double2* p = h ? vec_ptr(h, id) : nullptr;
  if (!p || !re || !im) return OCEAN_GPU_ERR_ARG;
  return put_planes(h, re, im, p, h->nvec());
}

int ocean_gpu_vec_get(ocean_gpu_handle* h, int32_t id, double* re, double* im) {
  double2* p = h ? vec_ptr(h, id) : nullptr;
  if (!p || !re || !im) return OCEAN_GPU_ERR_ARG;
  return get_planes(h, p, re, im, h->nvec());
}

int ocean_gpu_vec_dot(ocean_gpu_handle* h, int32_t p, int32_t q, double* rval, double* ival) {
  if (!h || !vec_ptr(h, p) || !vec_ptr(h, q) || !rval) return OCEAN_GPU_ERR_ARG;
  return dev_dot(h, vec_ptr(h, p), vec_ptr(h, q), rval, ival);
}

int ocean_gpu_vec_nrm(ocean_gpu_handle* h, int32_t p, double* val) {
  if (!h || !vec_ptr(h, p) || !val) return OCEAN_GPU_ERR_ARG;
  return dev_nrm(h, vec_ptr(h, p), val);
}

int ocean_gpu_vec_axpy(ocean_gpu_handle* h, double rval, double ival, int32_t x, int32_t y) {
  if (!h || !vec_ptr(h, x) || !vec_ptr(h, y)) return OCEAN_GPU_ERR_ARG;
  og::k_axpy<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(rval, ival, vec_ptr(h, x), vec_ptr(h, y), h->nvec());
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_scal(ocean_gpu_handle* h, double rval, int32_t x) {
  if (!h || !vec_ptr(h, x)) return OCEAN_GPU_ERR_ARG;
  og::k_scal<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(rval, vec_ptr(h, x), h->nvec());
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_axmy(ocean_gpu_handle* h, double rval, double ival, int32_t x, int32_t y) {
  if (!h || !vec_ptr(h, x) || !vec_ptr(h, y)) return OCEAN_GPU_ERR_ARG;
  og::k_axmz<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(rval, ival, vec_ptr(h, x), vec_ptr(h, y), vec_ptr(h, y),
                                                          h->nvec());
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_axmz(ocean_gpu_handle* h, double rval, double ival, int32_t x, int32_t y, int32_t z) {
  if (!h || !vec_ptr(h, x) || !vec_ptr(h, y) || !vec_ptr(h, z)) return OCEAN_GPU_ERR_ARG;
  og::k_axmz<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(rval, ival, vec_ptr(h, x), vec_ptr(h, z), vec_ptr(h, y),
                                                          h->nvec());
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_divide(ocean_gpu_handle* h, int32_t x, double rval, double ival, int32_t which) {
  if (!h || !vec_ptr(h, x) || which < 1 || which > 3) return OCEAN_GPU_ERR_ARG;
  og::k_divide<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(rval, ival, which, vec_ptr(h, x), h->nvec());
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int ocean_gpu_vec_copy(ocean_gpu_handle* h, int32_t dst, int32_t src) {
  if (!h || !vec_ptr(h, dst) || !vec_ptr(h, src)) return OCEAN_GPU_ERR_ARG;
  og::k_copy<<<grid_for(h->nvec()), kBlock, 0, h->stream>>>(vec_ptr(h, dst), vec_ptr(h, src), h->nvec());
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int ocean_gpu_xact_vec(ocean_gpu_handle* h, int32_t in, int32_t out, int32_t backwards, const int32_t* hflag) {
  if (!h || !vec_ptr(h, in) || !vec_ptr(h, out) || in == out) return OCEAN_GPU_ERR_ARG;
  const int f = flags_of(h, backwards, hflag);
  OG_TRY(check_ready(h, f));
  OG_TRY(xact_dev(h, vec_ptr(h, in), vec_ptr(h, out), f));
  OG_CHECK(hipStreamSynchronize(h->stream));
  return OCEAN_GPU_OK;
}

int ocean_gpu_xact_vec_repeat(ocean_gpu_handle* h, int32_t in, int32_t out, int32_t nrep) {
  if (!h || !vec_ptr(h, in) || !vec_ptr(h, out) || in == out || nrep <= 0) return OCEAN_GPU_ERR_ARG;
  const int f = flags_of(h, 0, nullptr);
  OG_TRY(check_ready(h, f));
  hipEvent_t e0, e1;
  OG_CHECK(hipEventCreate(&e0));
  OG_CHECK(hipEventCreate(&e1));
  OG_CHECK(hipEventRecord(e0, h->stream));
  int rc = OCEAN_GPU_OK;
  for (int r = 0; r < nrep && rc == 0; ++r) rc = xact_dev(h, vec_ptr(h, in), vec_ptr(h, out), f);
  OG_CHECK(hipEventRecord(e1, h->stream));
  OG_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  OG_CHECK(hipEventElapsedTime(&ms, e0, e1));
  h->t_total = ms / nrep;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int ocean_gpu_profile_xact(ocean_gpu_handle* h, int32_t in, int32_t out, int32_t nrep) {
  if (!h || !vec_ptr(h, in) || !vec_ptr(h, out) || in == out || nrep <= 0) return OCEAN_GPU_ERR_ARG;
  const int f = flags_of(h, 0, nullptr);
  OG_TRY(check_ready(h, f));
  // All repetitions are enqueued back to back with events between the kernel
  // families and one synchronisation at the end, so the stream never idles and
  // no family absorbs host launch latency; repetition 0 is discarded.
  const int R = nrep + 1;
  std::vector<hipEvent_t> ev((size_t)R * 5);
  for (auto& e : ev) OG_CHECK(hipEventCreate(&e));
  int rc = OCEAN_GPU_OK;
  for (int r = 0; r < R && rc == 0; ++r) rc = xact_dev(h, vec_ptr(h, in), vec_ptr(h, out), f, &ev[(size_t)r * 5]);
  if (rc == 0) OG_CHECK(hipEventSynchronize(ev.back()));
  double t[4] = {0, 0, 0, 0};
  for (int r = 1; r < R && rc == 0; ++r)
    for (int i = 0; i < 4; ++i) {
      float ms = 0.f;
      OG_CHECK(hipEventElapsedTime(&ms, ev[(size_t)r * 5 + i], ev[(size_t)r * 5 + i + 1]));
      t[i] += ms;
    }
  h->t_proj = t[0] / nrep; h->t_fft = t[1] / nrep; h->t_mult = t[2] / nrep; h->t_adj = t[3] / nrep;
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int ocean_gpu_last_timing(ocean_gpu_handle* h, const char* name, double* ms) {
  if (!h || !name || !ms) return OCEAN_GPU_ERR_ARG;
  const std::string n(name);
  if (n == "proj") *ms = h->t_proj;
  else if (n == "fft") *ms = h->t_fft;
  else if (n == "adj") *ms = h->t_adj;
  else if (n == "mult") *ms = h->t_mult;
  else if (n == "total") *ms = h->t_total;
  else return OCEAN_GPU_ERR_ARG;
  return OCEAN_GPU_OK;
}

// ---- Haydock ----
}  // extern "C"

namespace {

// Early-exit settings of the Haydock loop (sys%earlyExit, haydockConvergeSpacing,
// haydockConvergeThreshold; spectrum grid of calc_spect_core in Hartree).
struct EarlyExit {
  int spacing = 0;
  double threshold = 0.0, kpref = 0.0, gam = 0.0, el = 0.0, eh = 0.0;
  int ne = 0;
};

int haydock_run(H* h, const double* psi0_re, const double* psi0_im, int niter, const EarlyExit* ex,
                double* real_a, double* real_b, double* imag_a, int32_t* niter_done) {
  if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;  // ocean_gpu_haydock_multi
  const int f = flags_of(h, 0, nullptr);
  OG_TRY(check_ready(h, f));
  const size_t nv = h->nvec();
  const size_t n1 = (size_t)niter + 1;
  constexpr int kConvBlocks = 64;
  OG_TRY(ensure_scal(h, 3 * n1 + 3 * (size_t)kRedBlocks + 2 * n1 + 4 * (size_t)kConvBlocks));
  double* da = h->scal;
  double* dai = da + n1;
  double* db = dai + n1;
  double* part1 = db + n1;
  double* part2 = part1 + 2 * kRedBlocks;
  double* conv = part2 + kRedBlocks;      // [iter][relArea, maxDiff]
  double* cpart = conv + 2 * n1;
  double2 *v0 = nullptr, *v1 = nullptr, *v2 = nullptr;
  OG_TRY(dmalloc(&v0, nv));
  if (dmalloc(&v1, nv) || dmalloc(&v2, nv)) { dfree(v0); dfree(v1); return OCEAN_GPU_ERR_NOMEM; }
  int rc = put_planes(h, psi0_re, psi0_im, v0, nv);
  double2 *psi = v0, *old = v1, *hv = v2;
  int done = niter;
  // early exit: the convergence figures are formed on the device every iteration
  // and read back every kCheck iterations; the reference's decision is replayed in
  // order, so the iteration it stops at is the reference's (later ones are dropped)
  constexpr int kCheck = 8;
  std::vector<double> hconv(2 * n1, 0.0);
  bool prev_conv = false, stop = false;
  int checked = 0;
  auto replay = [&](int upto) -> int {
    OG_CHECK(hipMemcpyAsync(hconv.data(), conv, 2 * n1 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    OG_CHECK(hipStreamSynchronize(h->stream));
    for (int it = checked + 1; it <= upto && !stop; ++it) {
      if (it <= ex->spacing) continue;
      if (hconv[2 * (size_t)it] < ex->threshold) {
        if (prev_conv) { stop = true; done = it; }
        else prev_conv = true;
      } else {
        prev_conv = false;
      }
    }
    checked = upto;
    return OCEAN_GPU_OK;
  };
  if (rc == 0) {
    og::k_fill<<<grid_for(nv), kBlock, 0, h->stream>>>(old, make_double2(0.0, 0.0), nv);
    og::k_fill_d<<<grid_for(5 * n1), kBlock, 0, h->stream>>>(da, 0.0, 3 * n1);
    og::k_fill_d<<<grid_for(2 * n1), kBlock, 0, h->stream>>>(conv, 0.0, 2 * n1);
    for (int it = 1; it <= niter && rc == 0 && !stop; ++it) {
      rc = xact_dev(h, psi, hv, f);
      if (rc) break;
      og::k_hay1<<<kRedBlocks, kBlock, 0, h->stream>>>(hv, old, psi, db + (it - 1), nv, part1);
      og::k_hay2<<<kRedBlocks, kBlock, 0, h->stream>>>(hv, psi, part1, kRedBlocks, nv, da + (it - 1),
                                                      dai + (it - 1), part2);
      og::k_hay3<<<kRedBlocks, kBlock, 0, h->stream>>>(hv, part2, kRedBlocks, nv, db + it);
      if (ex && it > ex->spacing) {
        const int nb = std::min(kConvBlocks, std::max(1, (ex->ne + kBlock - 1) / kBlock));
        og::k_conv_part<<<nb, kBlock, 0, h->stream>>>(da, db, dai, it, it - ex->spacing, ex->kpref, ex->gam,
                                                      ex->el, ex->eh, ex->ne, cpart);
        og::k_conv_fin<<<1, 64, 0, h->stream>>>(cpart, nb, conv + 2 * (size_t)it);
      }
      if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
      double2* t = old;
      old = psi;
      psi = hv;
      hv = t;
      if (rc == 0 && ex && (it % kCheck == 0 || it == niter)) rc = replay(it);
    }
  }
  std::vector<double> hb(3 * n1);
  if (rc == 0) {
    if (hipMemcpyAsync(hb.data(), da, hb.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(v0); dfree(v1); dfree(v2);
  if (rc) return rc;
  for (int i = 0; i < done; ++i)
    if (hb[i] != hb[i]) { done = i + 1; rc = OCEAN_GPU_ERR_NAN; break; }
  // a(0:done-1) and b(0:done) belong to the recursion that stopped at `done`
  for (size_t i = 0; i < n1; ++i) {
    real_a[i] = (int)i < done ? hb[i] : 0.0;
    real_b[i] = (int)i <= done ? hb[2 * n1 + i] : 0.0;
    if (imag_a) imag_a[i] = (int)i < done ? hb[n1 + i] : 0.0;
  }
  if (niter_done) *niter_done = done;
  return rc;
}

}  // namespace

extern "C" {

int ocean_gpu_haydock(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im, int32_t niter,
                      double* real_a, double* real_b, double* imag_a, int32_t* niter_done) {
  if (!h || !psi0_re || !psi0_im || niter <= 0 || !real_a || !real_b) return OCEAN_GPU_ERR_ARG;
  return haydock_run(h, psi0_re, psi0_im, niter, nullptr, real_a, real_b, imag_a, niter_done);
}

int ocean_gpu_haydock_early(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im, int32_t niter,
                            int32_t spacing, double threshold, double kpref, double gam, double el, double eh,
                            int32_t ne, double* real_a, double* real_b, double* imag_a, int32_t* niter_done) {
  if (!h || !psi0_re || !psi0_im || niter <= 0 || !real_a || !real_b || spacing < 1 || ne < 1)
    return OCEAN_GPU_ERR_ARG;
  EarlyExit ex;
  ex.spacing = spacing;
  ex.threshold = threshold;
  ex.kpref = kpref;
  ex.gam = gam;
  ex.el = el;
  ex.eh = eh;
  ex.ne = ne;
  return haydock_run(h, psi0_re, psi0_im, niter, &ex, real_a, real_b, imag_a, niter_done);
}

// The recursions of every right-hand side of the handle at once: one action over the
// stacked vector per iteration (each u stream serves all nrhs right-hand sides), then
// the three-term steps of OCEAN_hay_ab_twoterm (OCEAN2/OCEAN_haydock.f90:824-938) per
// right-hand side in one launch each, with the fixed-order reductions of the
// one-right-hand-side recursion.
int ocean_gpu_haydock_multi(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im, int32_t niter,
                            double* real_a, double* real_b, double* imag_a, int32_t* niter_done) {
  if (!h || !psi0_re || !psi0_im || niter <= 0 || !real_a || !real_b) return OCEAN_GPU_ERR_ARG;
  const int f = flags_of(h, 0, nullptr);
  OG_TRY(check_ready(h, f));
  const int R = h->nrhs;
  const size_t nv = h->nvec(), nv0 = h->nvec0();
  const size_t n1 = (size_t)niter + 1;
  const size_t nscal = 3 * n1 + 3 * (size_t)kRedBlocks;  // a, ai, b, part1 (2 kRedBlocks), part2
  double2* V[3] = {nullptr, nullptr, nullptr};
  double* S = nullptr;
  int* dcount = nullptr;  // [0] iteration number, [1] its copy for the third step
  og::HayArgs* dhay[3] = {nullptr, nullptr, nullptr};
  int rc = OCEAN_GPU_OK;
  for (int i = 0; i < 3 && rc == 0; ++i) rc = dmalloc(&V[i], nv);
  if (rc == 0) rc = dmalloc(&S, (size_t)R * nscal);
  if (rc == 0) rc = dmalloc(&dcount, (size_t)2);
  for (int i = 0; i < 3 && rc == 0; ++i) rc = dmalloc(&dhay[i], (size_t)R);
  if (rc == 0) rc = put_planes(h, psi0_re, psi0_im, V[0], nv);
  // rotation r of an iteration: (psi, old, hv) = (V0, V1, V2), (V2, V0, V1), (V1, V2, V0)
  const int rot[3][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}};
  for (int r = 0; r < 3 && rc == 0; ++r) {
    std::vector<og::HayArgs> ha((size_t)R);
    for (int q = 0; q < R; ++q) {
      double* sq = S + (size_t)q * nscal;
      ha[q] = og::HayArgs{V[rot[r][2]] + q * nv0, V[rot[r][1]] + q * nv0, V[rot[r][0]] + q * nv0,
                          sq + 3 * n1, sq + 3 * n1 + 2 * kRedBlocks, sq, sq + n1, sq + 2 * n1, nv0};
    }
    if (hipMemcpy(dhay[r], ha.data(), sizeof(og::HayArgs) * R, hipMemcpyHostToDevice) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  // psi_old = 0, coefficients 0 (b_0 = 0), iteration 1
  if (rc == 0) {
    og::k_fill<<<grid_for(nv), kBlock, 0, h->stream>>>(V[1], make_double2(0.0, 0.0), nv);
    og::k_fill_d<<<grid_for((size_t)R * nscal), kBlock, 0, h->stream>>>(S, 0.0, (size_t)R * nscal);
    const int init[2] = {1, 0};
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(dcount, init, sizeof(init), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  for (int it = 1; it <= niter && rc == 0; ++it) {
    const int r = (it - 1) % 3;
    rc = xact_dev(h, V[rot[r][0]], V[rot[r][2]], f);
    if (rc) break;
    og::k_hay1_batch<<<R * kRedBlocks, kBlock, 0, h->stream>>>(dhay[r], dcount);
    og::k_hay2_batch<<<R * kRedBlocks, kBlock, 0, h->stream>>>(dhay[r], dcount);
    og::k_hay3_batch<<<R * kRedBlocks, kBlock, 0, h->stream>>>(dhay[r], dcount);
    if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0 && hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  // coefficients back to the host, NaN check per right-hand side as the single recursion
  int rc_nan = OCEAN_GPU_OK;
  std::vector<double> hb(3 * n1);
  for (int q = 0; q < R && rc == 0; ++q) {
    if (hipMemcpy(hb.data(), S + (size_t)q * nscal, hb.size() * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      rc = OCEAN_GPU_ERR_HIP;
      break;
    }
    int done = niter;
    for (int i = 0; i < niter; ++i)
      if (hb[i] != hb[i]) { done = i + 1; rc_nan = OCEAN_GPU_ERR_NAN; break; }
    for (size_t i = 0; i < n1; ++i) {
      real_a[q * n1 + i] = (int)i < done ? hb[i] : 0.0;
      real_b[q * n1 + i] = (int)i <= done ? hb[2 * n1 + i] : 0.0;
      if (imag_a) imag_a[q * n1 + i] = (int)i < done ? hb[n1 + i] : 0.0;
    }
    if (niter_done) niter_done[q] = done;
  }
  for (auto& v : V) dfree(v);
  for (auto& d : dhay) dfree(d);
  dfree(S);
  dfree(dcount);
  return rc ? rc : rc_nan;
}

// The non-Hermitian (bi-Lanczos) recursion of OCEAN2's complex_haydock branch
// (OCEAN2/OCEAN_haydock.f90:60-160): both vectors start from psi0 and the allow mask,
// each iteration applies the action to psi and its Hermitian conjugate (the
// backwards action) to bpsi, then haydock_abc_1 (:1250-1433).  Vectors stay on the
// device; the iteration number is a kernel argument.
int ocean_gpu_haydock_abc(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im, int32_t niter,
                          double* real_a, double* imag_a, double* real_b, double* imag_b, double* real_c,
                          double* imag_c, int32_t* niter_done) {
  if (!h || !psi0_re || !psi0_im || niter <= 0 || !real_a || !imag_a || !real_b || !imag_b || !real_c ||
      !imag_c)
    return OCEAN_GPU_ERR_ARG;
  if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  const int f = flags_of(h, 0, nullptr);
  OG_TRY(check_ready(h, f));
  const int fb = flags_of(h, 1, nullptr);
  const size_t nv = h->nvec();
  const int n1 = niter + 1;
  double2* V[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // psi, old, h; bpsi, bold, bh
  double* S = nullptr;  // coef [6][n1], part1, part2
  int rc = OCEAN_GPU_OK;
  for (int i = 0; i < 6 && rc == 0; ++i) rc = dmalloc(&V[i], nv);
  if (rc == 0) rc = dmalloc(&S, 6 * (size_t)n1 + 4 * (size_t)kRedBlocks);
  if (rc == 0) rc = put_planes(h, psi0_re, psi0_im, V[0], nv);
  if (rc == 0) {
    // psi = bpsi = allow . psi0; old = bold = 0; coefficients 0 (b(0) = c(0) = 0)
    k_mask<<<grid_for(nv), kBlock, 0, h->stream>>>(V[0], V[0], h->allow, nv);
    og::k_copy<<<grid_for(nv), kBlock, 0, h->stream>>>(V[3], V[0], nv);
    og::k_fill<<<grid_for(nv), kBlock, 0, h->stream>>>(V[1], make_double2(0.0, 0.0), nv);
    og::k_fill<<<grid_for(nv), kBlock, 0, h->stream>>>(V[4], make_double2(0.0, 0.0), nv);
    og::k_fill_d<<<grid_for(6 * (size_t)n1), kBlock, 0, h->stream>>>(S, 0.0, 6 * (size_t)n1);
    if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  int ip = 0, io = 1, ih = 2;  // rotation of (psi, old, h); the backward set is +3
  for (int it = 1; it <= niter && rc == 0; ++it) {
    rc = xact_dev(h, V[ip], V[ih], f);
    if (rc == 0) rc = xact_dev(h, V[3 + ip], V[3 + ih], fb);
    if (rc) break;
    og::AbcArgs A{V[ih], V[3 + ih], V[io], V[3 + io], V[ip], V[3 + ip], h->allow, S,
                  S + 6 * (size_t)n1, S + 6 * (size_t)n1 + 2 * kRedBlocks, nv, n1};
    og::k_abc1<<<kRedBlocks, kBlock, 0, h->stream>>>(A);
    og::k_abc2<<<kRedBlocks, kBlock, 0, h->stream>>>(A, it);
    og::k_abc3<<<kRedBlocks, kBlock, 0, h->stream>>>(A, it);
    if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
    // old <- psi, psi <- h (the old old buffer takes the next action)
    const int t = io;
    io = ip;
    ip = ih;
    ih = t;
  }
  std::vector<double> hc(6 * (size_t)n1);
  if (rc == 0 && (hipMemcpyAsync(hc.data(), S, hc.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream) !=
                      hipSuccess ||
                  hipStreamSynchronize(h->stream) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  for (auto& v : V) dfree(v);
  dfree(S);
  if (rc) return rc;
  int done = niter;
  for (int i = 0; i < niter; ++i)
    if (hc[i] != hc[i]) { done = i + 1; rc = OCEAN_GPU_ERR_NAN; break; }
  double* outs[6] = {real_a, imag_a, real_b, imag_b, real_c, imag_c};
  for (int q = 0; q < 6; ++q)
    for (int i = 0; i < n1; ++i) {
      const bool keep = q < 2 ? i < done : i <= done;  // a(0:done-1), b and c (0:done)
      outs[q][i] = keep ? hc[(size_t)q * n1 + i] : 0.0;
    }
  if (niter_done) *niter_done = done;
  return rc;
}

// The continued fraction of ocean_gpu_spectrum with complex b and c (the
// bi-Lanczos coefficients), as write_core evaluates them (OCEAN2/OCEAN_haydock.f90:
// 1846-1847, :1870).
int ocean_gpu_spectrum_abc(const double* real_a, const double* imag_a, const double* real_b, const double* imag_b,
                           const double* real_c, const double* imag_c, int32_t niter, double kpref, double gam,
                           double el, double eh, int32_t ne, int32_t literal_cmplx, double* out) {
  if (!real_a || !imag_a || !real_b || !imag_b || !real_c || !imag_c || !out || niter <= 0 || ne <= 0)
    return OCEAN_GPU_ERR_ARG;
  const size_t n1 = (size_t)niter + 1;
  double* d = nullptr;
  OG_TRY(dmalloc(&d, 6 * n1 + 5 * ((size_t)ne + 1)));
  const double* src[6] = {real_a, real_b, imag_a, imag_b, real_c, imag_c};  // device order ra, rb, ia, ib, rc, ic
  int rc = OCEAN_GPU_OK;
  for (int q = 0; q < 6 && rc == 0; ++q)
    if (hipMemcpy(d + q * n1, src[q], n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    og::k_cfrac<<<(ne + 1 + 127) / 128, 128>>>(d, d + n1, d + 2 * n1, niter, kpref, gam, el, eh, ne, literal_cmplx,
                                               kHartree2eV, d + 6 * n1, d + 3 * n1, d + 4 * n1, d + 5 * n1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d + 6 * n1, 5 * ((size_t)ne + 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(d);
  return rc;
}

// ---- spectrum, OCEAN 3.2.1 write_core (nine columns / Voigt) ----
int ocean_gpu_write_core(const double* real_a, const double* real_b, const double* imag_a, int32_t niter,
                         double kpref, const ocean_spect_desc* sd, double* out, int32_t* ncols) {
  if (!real_a || !real_b || !out || !sd || niter <= 0 || sd->ne <= 0 || !(sd->eh > sd->el)) return OCEAN_GPU_ERR_ARG;
  const int ne = sd->ne;
  const size_t n1 = (size_t)niter + 1;
  const bool voigt = sd->gam_gauss > sd->gam0 / 10.0;
  const int nc = voigt ? 5 : 9;
  if (ncols) *ncols = nc;
  // write_core_Voigt's fine grid (OCEAN_haydock.f90:1910-1917), in its order
  double estep = 0.0, prefactor = 0.0, inv2s2 = 0.0;
  int nratio = 1, nstep = 0;
  if (voigt) {
    const double de = (sd->eh - sd->el) / (double)ne;
    estep = std::min(sd->gam0, sd->gam_gauss) / 2.0;
    nratio = (int)((de - estep) / estep) + 1;
    estep = de / (double)nratio;
    nstep = (int)((2.99999 * sd->gam_gauss) / estep) + 1;
    prefactor = estep / (sd->gam_gauss * 2.506628274631);
    inv2s2 = 1.0 / (2.0 * sd->gam_gauss * sd->gam_gauss);
    // a grid coarser than six widths would shift past the reference's fine array
    if (nratio > 2 * nstep || nratio < 1) return OCEAN_GPU_ERR_ARG;
  }
  double* d = nullptr;
  OG_TRY(dmalloc(&d, 3 * n1 + (size_t)nc * (ne + 1)));
  std::vector<double> ia(n1, 0.0);
  if (imag_a) std::copy(imag_a, imag_a + n1, ia.begin());
  int rc = OCEAN_GPU_OK;
  if (hipMemcpy(d, real_a, n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n1, real_b, n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + 2 * n1, ia.data(), n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    if (voigt)
      og::k_write_core_voigt<<<(ne + 1 + 63) / 64, 64>>>(d, d + n1, d + 2 * n1, niter, kpref, sd->gam0, sd->el, sd->eh,
                                                         ne, sd->ebase_ev, sd->nval, sd->eps, sd->literal_cmplx,
                                                         estep, nratio, nstep, prefactor, inv2s2, kHartree2eV,
                                                         d + 3 * n1);
    else
      og::k_write_core<<<(ne + 1 + 127) / 128, 128>>>(d, d + n1, d + 2 * n1, niter, kpref, sd->gam0, sd->el, sd->eh,
                                                      ne, sd->ebase_ev, sd->old_broaden, sd->nval, sd->eps,
                                                      sd->ucvol, sd->literal_cmplx,
                                                      og::SpectConst{kHartree2eV, kEV2Hartree, kAlpha, kBohr, kPi,
                                                                     kAu2sec},
                                                      d + 3 * n1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d + 3 * n1, (size_t)nc * (ne + 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(d);
  return rc;
}

// ---- o_spect (OCEAN2/o_spect.f90) ----
int ocean_gpu_ospect(const double* real_a, const double* real_b, int32_t iter, double kpref,
                     const ocean_ospect_desc* od, double* out, int32_t* nrows, int32_t* ncols) {
  if (!real_a || !real_b || !od || !out || iter <= 0 || od->ne <= 0) return OCEAN_GPU_ERR_ARG;
  const int ne = od->ne;
  const int nr = od->is_val ? ne : ne + 1, nc = od->is_val ? 8 : 5;
  if (nrows) *nrows = nr;
  if (ncols) *ncols = nc;
  // spect.in is in eV (o_spect.f90:30-33)
  const double el = od->el_ev * kEV2Hartree, eh = od->eh_ev * kEV2Hartree;
  const double gam0 = od->gam0_ev * kEV2Hartree, ebase = od->ebase_ev * kEV2Hartree;
  const size_t n1 = (size_t)iter + 1;
  double* d = nullptr;
  OG_TRY(dmalloc(&d, 2 * n1 + (size_t)nr * nc));
  int rc = OCEAN_GPU_OK;
  if (hipMemcpy(d, real_a, n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n1, real_b, n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    if (od->is_val)
      og::k_ospect_val<<<(nr + 63) / 64, 64>>>(d, d + n1, iter, kpref, od->ucvol, el, eh, ne, gam0, kHartree2eV,
                                               kBohr * kAlphaInv * 1000.0, d + 2 * n1);
    else
      og::k_ospect_core<<<(nr + 63) / 64, 64>>>(d, d + n1, iter, kpref, el, eh, ne, gam0, ebase, od->nval,
                                                od->eps0, kHartree2eV, d + 2 * n1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d + 2 * n1, (size_t)nr * nc * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(d);
  return rc;
}

// ---- Lanczos eigenvalues (redtrid) ----
int ocean_gpu_tridiag_eigs(const double* real_a, const double* real_b, int32_t n, double* w) {
  if (!real_a || !real_b || !w || n < 0) return OCEAN_GPU_ERR_ARG;
  const int m = n + 1;
  // b(1:n) couple rows i-1 and i; b(0) is unused
  std::vector<double> bb((size_t)m, 0.0);
  for (int i = 1; i < m; ++i) bb[i] = real_b[i];
  double lo = real_a[0], hi = real_a[0], amax = 0.0;
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(bb[i]) : 0.0) + (i + 1 < m ? std::fabs(bb[i + 1]) : 0.0);
    lo = std::min(lo, real_a[i] - r);
    hi = std::max(hi, real_a[i] + r);
    amax = std::max(amax, std::max(std::fabs(real_a[i]), std::fabs(bb[i])));
  }
  const double span = std::max(hi - lo, 1e-300);
  lo -= 1e-14 * span;
  hi += 1e-14 * span;
  const double pivmin = std::max(amax * amax, 1e-300) * 2.2250738585072014e-308 / 1.1102230246251565e-16;
  double* d = nullptr;
  OG_TRY(dmalloc(&d, 3 * (size_t)m));
  int rc = OCEAN_GPU_OK;
  if (hipMemcpy(d, real_a, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + m, bb.data(), m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    og::k_tridiag_eigs<<<(m + 63) / 64, 64>>>(d, d + m, m, lo, hi, pivmin, d + 2 * m);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(w, d + 2 * m, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(d);
  return rc;
}

// ---- spectrum ----
int ocean_gpu_spectrum(const double* real_a, const double* real_b, const double* imag_a, int32_t niter,
                       double kpref, double gam, double el, double eh, int32_t ne, int32_t literal_cmplx,
                       double* out) {
  if (!real_a || !real_b || !out || niter <= 0 || ne <= 0) return OCEAN_GPU_ERR_ARG;
  double *d = nullptr;
  const size_t n1 = (size_t)niter + 1;
  OG_TRY(dmalloc(&d, 3 * n1 + 5 * ((size_t)ne + 1)));
  std::vector<double> ia(n1, 0.0);
  if (imag_a) std::copy(imag_a, imag_a + n1, ia.begin());
  int rc = OCEAN_GPU_OK;
  if (hipMemcpy(d, real_a, n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n1, real_b, n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + 2 * n1, ia.data(), n1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    og::k_cfrac<<<(ne + 1 + 127) / 128, 128>>>(d, d + n1, d + 2 * n1, niter, kpref, gam, el, eh, ne,
                                               literal_cmplx, kHartree2eV, d + 3 * n1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d + 3 * n1, 5 * ((size_t)ne + 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(d);
  return rc;
}

}  // extern "C"

// ---- the continued fraction without allocation ----
// ocean_gpu_spectrum stages a, b, imag_a and the output in a buffer it allocates per
// call.  The allocation-free form stages them in this file-scope array instead (each
// GPU holds its own copy until the process ends).  Host code never passes the
// array's name to a kernel or copy: it is a host-side stand-in, so every use goes
// through the device address that hipGetSymbolAddress returns for the current GPU.
namespace {

constexpr size_t kNoAllocDoubles = (size_t)1 << 19;  // 4 MiB per GPU
__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of the current GPU's scratch array, looked up once per GPU.
int scratch_for_device(double** out) {
  int dev = 0;
  OG_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return OCEAN_GPU_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    OG_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return OCEAN_GPU_OK;
}

}  // namespace

extern "C" {

int ocean_gpu_spectrum_noalloc(const double* real_a, const double* real_b, const double* imag_a, int32_t niter,
                               double kpref, double gam, double el, double eh, int32_t ne, int32_t literal_cmplx,
                               double* out) {
  if (!real_a || !real_b || !out || niter <= 0 || ne <= 0) return OCEAN_GPU_ERR_ARG;
  const size_t n1 = (size_t)niter + 1;
  if (3 * n1 + 5 * ((size_t)ne + 1) > kNoAllocDoubles) return OCEAN_GPU_ERR_ARG;
  double* d = nullptr;
  OG_TRY(scratch_for_device(&d));
  OG_CHECK(hipMemcpy(d, real_a, n1 * sizeof(double), hipMemcpyHostToDevice));
  OG_CHECK(hipMemcpy(d + n1, real_b, n1 * sizeof(double), hipMemcpyHostToDevice));
  if (imag_a) {
    OG_CHECK(hipMemcpy(d + 2 * n1, imag_a, n1 * sizeof(double), hipMemcpyHostToDevice));
  } else {
    // the array keeps the previous call's contents: a missing imag_a is zero
    og::k_fill_d<<<(int)((n1 + kBlock - 1) / kBlock), kBlock>>>(d + 2 * n1, 0.0, n1);
    OG_CHECK(hipGetLastError());
  }
  og::k_cfrac<<<(ne + 1 + 127) / 128, 128>>>(d, d + n1, d + 2 * n1, niter, kpref, gam, el, eh, ne,
                                             literal_cmplx, kHartree2eV, d + 3 * n1);
  OG_CHECK(hipGetLastError());
  OG_CHECK(hipMemcpy(out, d + 3 * n1, 5 * ((size_t)ne + 1) * sizeof(double), hipMemcpyDeviceToHost));
  return OCEAN_GPU_OK;
}

// ---- GMRES ----
}  // extern "C"

namespace {

// Recycled GMRES over the energy list for the device right-hand side rhs (already
// masked).  Solutions go to xs_dev (nE vectors, nullable) and/or the host planes.
int gmres_run(H* h, const double2* rhs_dev, int nE, const double* energies_eV, double gres_eV,
              double gprc_eV, double ffff, int depth, double kpref, double* out, double2* xs_dev,
              double* x_re, double* x_im) {
  if (depth > kMaxGmresDepth) return OCEAN_GPU_ERR_ARG;
  if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  const int f = flags_of(h, 0, nullptr);
  OG_TRY(check_ready(h, f));
  const size_t nv = h->nvec();
  const int grid = grid_for(nv);
  const double gres = gres_eV * kEV2Hartree, gprc = gprc_eV * kEV2Hartree;
  hipStream_t st = h->stream;
  // vectors: rhs, h1, x, ax, g, pcdiv, u[1..depth], au[1..depth]
  const int nfix = 6;
  std::vector<double2*> V(nfix + 2 * depth, nullptr);
  int rc = OCEAN_GPU_OK;
  for (auto& p : V)
    if (rc == 0) rc = dmalloc(&p, nv);
  double2** dlist = nullptr;
  double2* dcoef = nullptr;
  double* part = nullptr;
  if (rc == 0) rc = dmalloc(&dlist, 2 * (size_t)depth + 2);
  if (rc == 0) rc = dmalloc(&dcoef, (size_t)depth);
  if (rc == 0) rc = dmalloc(&part, (size_t)kRedBlocks * 4 * depth + 4 * (size_t)depth + 3 * (size_t)kRedBlocks + 16);
  double2* dL = nullptr;
  double* dD = nullptr;
  double2* dC = nullptr;   // cumulative least-squares coefficients of the cycle
  double2* dRU = nullptr;  // <rhs|U_i>
  GmCtl* ctl = nullptr;
  GmCtl* hctl = nullptr;
  if (rc == 0) rc = dmalloc(&dL, (size_t)depth * (depth - 1) / 2 + 1);
  if (rc == 0) rc = dmalloc(&dD, (size_t)depth);
  if (rc == 0) rc = dmalloc(&dC, (size_t)depth);
  if (rc == 0) rc = dmalloc(&dRU, (size_t)depth);
  if (rc == 0) rc = dmalloc(&ctl, (size_t)1);
  if (rc == 0 && hipHostMalloc((void**)&hctl, sizeof(GmCtl), hipHostMallocDefault) != hipSuccess)
    rc = OCEAN_GPU_ERR_NOMEM;
  auto cleanup = [&]() {
    for (auto& p : V) dfree(p);
    dfree(dlist); dfree(dcoef); dfree(part); dfree(dL); dfree(dD); dfree(dC); dfree(dRU); dfree(ctl);
    if (hctl) (void)hipHostFree(hctl);
  };
  if (rc) { cleanup(); return rc; }
  int* skip = &ctl->done;
  int* info = &ctl->info;
  double2 *rhs = V[0], *h1 = V[1], *x = V[2], *ax = V[3], *g = V[4], *pc = V[5];
  double2** U = &V[nfix];            // U[i], i = 0..depth-1  (reference u_matrix(i+1))
  double2** AU = &V[nfix + depth];   // AU[i]
  std::vector<double2*> hlist(2 * depth);
  for (int i = 0; i < depth; ++i) { hlist[i] = AU[i]; hlist[depth + i] = U[i]; }

  auto launched = [&]() -> int {
    if (hipGetLastError() != hipSuccess) return OCEAN_GPU_ERR_HIP;
    return OCEAN_GPU_OK;
  };
  OG_CHECK(hipMemcpyAsync(rhs, rhs_dev, nv * sizeof(double2), hipMemcpyDeviceToDevice, st));
  k_gm_reset<<<1, 64, 0, st>>>(ctl);
  og::k_fill<<<1, 64, 0, st>>>(reinterpret_cast<double2*>(ctl), make_double2(0.0, 0.0), sizeof(GmCtl) / 16);
  og::k_fill<<<grid, kBlock, 0, st>>>(x, make_double2(1.0, 0.0), nv);
  og::k_fill<<<1, 64, 0, st>>>(dC, make_double2(0.0, 0.0), (size_t)depth);
  rc = launched();
  if (rc == 0) rc = xact_dev(h, x, h1, f);  // h1 = H . 1
  if (rc == 0) rc = hipMemcpy(dlist, hlist.data(), 2 * depth * sizeof(double2*), hipMemcpyHostToDevice) == hipSuccess
                        ? OCEAN_GPU_OK : OCEAN_GPU_ERR_HIP;

  // A finish of a residual step that the next action's projection runs (within a
  // restart cycle the next kernel after the residual step is always that projection)
  const bool ride = (f & og::F_LR) && !h->obf_on;
  GmFin pend{};
  bool have_pend = false;
  // outv = (E + i gres) in - H in (formed in the action's epilogue), dropped once
  // the solve has stopped
  auto axmy_xact = [&](double2* in, double2* outv, double E) -> int {
    const double2 c = make_double2(E, gres);
    const GmFin* fp = have_pend ? &pend : nullptr;
    have_pend = false;
    OG_TRY(xact_dev(h, in, outv, f, nullptr, skip, &c, fp));
    return launched();
  };
  // x is kept as x_b + sum_i C_i U_i: each iteration's update goes into the
  // cumulative coefficients C and into ax, and <rhs|x> for the stopping test comes
  // from <rhs|x_b> and <rhs|U_i>; x itself is formed (materialise) only where the
  // basis is about to be replaced, at a restart, and when the solve has converged.
  // gated: dropped once the solve has stopped (the converged x is formed by the host's
  // ungated call after reading the state)
  auto materialise = [&](bool gated) -> int {
    k_materialize<<<grid, kBlock, 0, st>>>(gated ? skip : nullptr, (const double2* const*)(dlist + depth), dC, depth,
                                           x, nv);
    k_fold_c<<<1, 64, 0, st>>>(gated ? skip : nullptr, ctl, dC, dRU, depth);
    return launched();
  };
  // g = ax - rhs with x materialised, its sums and (with check) the stopping test of
  // iteration it
  double* rpart = part + (size_t)kRedBlocks * 4 * depth + 4 * (size_t)depth;
  auto fin_of = [&](int it, int check, int mode, int m, int slot) -> GmFin {
    GmFin F{};
    F.c = ctl; F.part = rpart; F.C = dC; F.rU = dRU; F.ffff = ffff;
    F.nblk = kRedBlocks; F.it = it; F.check = check; F.mode = mode; F.m = m; F.slot = slot;
    return F;
  };
  auto resid = [&](int it, int check) -> int {
    k_resid_step<<<kRedBlocks, kBlock, 0, st>>>(ctl, ax, rhs, x, g, nv, rpart);
    k_resid_fin<<<1, 256, 0, st>>>(fin_of(it, check, kResX, 0, -1));
    return launched();
  };
  // update_gmres(cur): Gram column cur, least-squares solve and the x / ax updates
  // (OCEAN_gmres.f90:241-344); with rit >= 0 the ax update carries the residual step
  // of iteration rit, and unext (nullable) receives the next basis vector U[cur];
  // without, the update ends a cycle: x is materialised (ax is recomputed from it)
  auto update = [&](int cur, int rit = -1, int check = 0, double2* unext = nullptr) -> int {
    k_multidot2<<<kRedBlocks, kBlock, 0, st>>>(skip, (const double2* const*)dlist, cur, AU[cur - 1], g, nv, part);
    double* fin = part + (size_t)kRedBlocks * 4 * depth;
    k_multifinish<<<(4 * cur + 3) / 4, 256, 0, st>>>(skip, part, kRedBlocks, 4 * cur, fin);
    k_ls_update<<<1, 64, 0, st>>>(skip, fin, dL, dD, dcoef, dC, cur, depth, info);
    if (rit >= 0) {
      k_axpy_resid<<<kRedBlocks, kBlock, 0, st>>>(ctl, (const double2* const*)dlist, dcoef, cur, ax, rhs, g, pc,
                                                  unext, nv, rpart);
      const GmFin F = fin_of(rit, check, kResU, cur, unext ? cur : -1);
      if (ride && check && unext) {
        // the next kernel is the next iteration's projection: the finish rides on it
        pend = F;
        have_pend = true;
      } else {
        k_resid_fin<<<1, 256, 0, st>>>(F);
      }
      return launched();
    }
    OG_TRY(launched());
    return materialise(true);
  };
  auto read_ctl = [&]() -> int {
    OG_CHECK(hipMemcpyAsync(hctl, ctl, sizeof(GmCtl), hipMemcpyDeviceToHost, st));
    OG_CHECK(hipStreamSynchronize(st));
    return hctl->done == 2 ? (hctl->info > 0 ? hctl->info : OCEAN_GPU_ERR_HIP) : OCEAN_GPU_OK;
  };
  auto set_prec = [&](double E) -> int {
    k_set_prec<<<grid, kBlock, 0, st>>>(skip, h1, pc, E, gprc, nv);
    return launched();
  };
  // iterations enqueued ahead of each read-back of the device state
  constexpr int kAhead = 8;

  int global_iter = 0, prev_iter = 0;
  double prev_energy = -1e300;
  for (int step = 0; step < nE && rc == 0; ++step) {
    int complete_iter = 0;
    const double E = energies_eV[step] * kEV2Hartree;
    // a finish still pending belongs to an iteration enqueued after the previous
    // energy converged (its kernels were dropped): it must not run in this energy
    have_pend = false;
    k_gm_reset<<<1, 64, 0, st>>>(ctl);
    if ((rc = launched())) break;
    if (prev_iter == 0) {
      prev_energy = E;
      if ((rc = set_prec(E))) break;
    }
    bool converged = false;
    if (prev_iter == 0) {
      if (step > 0) {
        if ((rc = axmy_xact(x, ax, E))) break;
      } else {
        og::k_fill<<<grid, kBlock, 0, st>>>(x, make_double2(0.0, 0.0), nv);
        og::k_fill<<<grid, kBlock, 0, st>>>(ax, make_double2(0.0, 0.0), nv);
      }
      if ((rc = resid(0, 0))) break;
    } else {
      // make_recycle: au += (E - E_prev) u, then the updates over the kept basis
      const double delta = E - energies_eV[step - 1] * kEV2Hartree;
      for (int i = 0; i < prev_iter; ++i) k_axpy_gated<<<grid, kBlock, 0, st>>>(skip, delta, 0.0, U[i], AU[i], nv);
      if ((rc = axmy_xact(x, ax, E))) break;
      if ((rc = resid(0, 0))) break;
      for (int it = 1; it <= prev_iter && rc == 0; ++it) rc = update(it, 0, 0, nullptr);
      if (rc || (rc = read_ctl())) break;
      const double gval = hctl->gval, ival = hctl->ival;
      if (gval < ffff) {
        const double rel = -gval / ival;
        if (rel < 0.001 || gval < ffff / 10.0) converged = true;
      }
      if (!converged && std::fabs(prev_energy - E) > std::fabs(gprc) / 4.0) {
        // a new cycle replaces the kept basis: form x from it first
        if ((rc = materialise(false))) break;
        prev_iter = 0;
        prev_energy = E;
        if ((rc = set_prec(E))) break;
      }
    }
    while (!converged && rc == 0) {
      if (complete_iter >= 50 * depth) {  // 50 restarts at one energy
        rc = OCEAN_GPU_ERR_NOCONV;
        break;
      }
      const int cycle_start = prev_iter;
      const bool new_prec_at_end = (E != prev_energy);
      int it = prev_iter + 1;
      while (it <= depth && rc == 0) {
        const int iend = std::min(depth, it + kAhead - 1);
        for (; it <= iend && rc == 0; ++it) {
          // the first basis vector of a cycle from g and the current preconditioner;
          // later ones were formed by the previous iteration's residual step
          if (it == cycle_start + 1) {
            k_pc_apply<<<kRedBlocks, kBlock, 0, st>>>(ctl, U[it - 1], g, pc, rhs, nv, rpart);
            k_resid_fin<<<1, 256, 0, st>>>(fin_of(0, 0, kDotU, 0, it - 1));
          }
          if ((rc = axmy_xact(U[it - 1], AU[it - 1], E))) break;
          if (it < depth) {
            rc = update(it, it, 1, U[it]);
          } else {
            // the end of a cycle recomputes ax from x before the residual
            if ((rc = update(it))) break;
            if ((rc = axmy_xact(x, ax, E))) break;
            if (new_prec_at_end && (rc = set_prec(E))) break;
            rc = resid(it, 1);
          }
        }
        if (rc || (rc = read_ctl())) break;
        if (hctl->done == 1) {
          const int c = hctl->conv_it;
          complete_iter += c - cycle_start;
          global_iter += c - cycle_start;
          if (c == depth) prev_energy = E;
          prev_iter = (c == depth) ? 0 : c;
          converged = true;
          break;
        }
      }
      if (rc || converged) break;
      complete_iter += depth - cycle_start;
      global_iter += depth - cycle_start;
      prev_energy = E;
      prev_iter = 0;
    }
    if (rc) break;
    // the converged x (C is zero if it was formed at a restart)
    if ((rc = materialise(false))) break;
    const double gval = hctl->gval, rval = hctl->rval, ival = hctl->ival;
    double* o = out + 6 * (size_t)step;
    o[0] = energies_eV[step];
    o[1] = 1.0 - rval * kpref;
    o[2] = -ival * kpref;
    o[3] = -gval / ival;
    o[4] = complete_iter;
    o[5] = global_iter;
    if (x_re && x_im) rc = get_planes(h, x, x_re + step * nv, x_im + step * nv, nv);
    if (xs_dev && rc == 0)
      rc = hipMemcpyAsync(xs_dev + step * nv, x, nv * sizeof(double2), hipMemcpyDeviceToDevice, st) == hipSuccess
               ? OCEAN_GPU_OK : OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0 && hipStreamSynchronize(st) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  cleanup();
  return rc;
}

// rhs = allow .* hay_vec (the reference's sfact copy, allow%i = allow%r for core)
int gmres_rhs(H* h, const double* hay_re, const double* hay_im, double2** rhs) {
  const size_t nv = h->nvec();
  double2* tmp = nullptr;
  OG_TRY(dmalloc(rhs, nv));
  int rc = dmalloc(&tmp, nv);
  if (rc == 0) rc = put_planes(h, hay_re, hay_im, tmp, nv);
  if (rc == 0) {
    k_mask<<<grid_for(nv), kBlock, 0, h->stream>>>(*rhs, tmp, h->allow, nv);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  dfree(tmp);
  if (rc) dfree(*rhs);
  return rc;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// Per-energy part of the spectral gradient (see ocean_gpu_spectral_grad).
template <typename T2, int NA>
int grad_t(H* h, int nE, const double2* xs, const double2* ys, const std::vector<double>& r, double kpref,
           bool want_s, bool want_w, double* gG, double* gD, double* gS, double* Q) {
  const size_t nv = h->nvec();
  const int f = flags_of(h, 0, nullptr);
  T2 *PX = nullptr, *PY = nullptr;
  double2* v = nullptr;
  int rc = OCEAN_GPU_OK;
  if (want_w) {
    rc = dmalloc(&PX, h->nphi());
    if (rc == 0) rc = dmalloc(&PY, h->nphi());
  }
  if (rc == 0 && want_s) rc = dmalloc(&v, nv);
  const double s = h->d.inter_scale;
  for (int e = 0; e < nE && rc == 0; ++e) {
    const double2* x = xs + (size_t)e * nv;
    const double2* y = ys + (size_t)e * nv;
    const double c = 2.0 * r[e] * kpref;
    double zr = 0.0, zi = 0.0;
    if ((rc = dev_dot(h, y, x, &zr, &zi))) break;
    *gG += c * zr;   // d sigma / d Gamma = kpref Re <y|x>
    *gD -= c * zi;   // d sigma / d Delta = -kpref Im <y|x>
    if (want_s) {
      // (H_lr + H_mult) x: the action with the energy and spin-orbit terms off
      if ((rc = xact_dev(h, x, v, f & (og::F_LR | og::F_FG)))) break;
      if ((rc = dev_dot(h, y, v, &zr, &zi))) break;
      *gS -= c * zi / s;
    }
    if (want_w && (f & og::F_LR)) {
      og::ActArgs<T2> a;
      Geom<T2> g;
      const T2* u = StateOf<T2>::u(h);
      const auto* W = StateOf<T2>::W(h);
      for (int side = 0; side < 2 && rc == 0; ++side) {
        T2* P = side == 0 ? PX : PY;
        const Tiles tl = tiles_for<T2, NA>(h);
        rc = make_args<T2, NA>(h, side == 0 ? x : y, nullptr, og::F_LR, tl, u, StateOf<T2>::uT(h), P, W, a, g);
        if (rc) break;
        og::k_proj<T2, NA><<<g.nblk_proj, tl.proj, g.sm_proj, h->stream>>>(a);
        const size_t tile = (size_t)h->nk * a.G;
        const size_t smf = (3 * (size_t)g.lmax + tile + tile / 8 + 1) * sizeof(T2);
        if (g.lmax == 8) {
          og::k_fft_fwd<T2, 8><<<a.nfft, kBlock, smf, h->stream>>>(a);
        } else if (g.lmax == 16) {
          if ((rc = lds_opt_in(og::k_fft_fwd<T2, 16>, smf))) break;
          og::k_fft_fwd<T2, 16><<<a.nfft, kBlock, smf, h->stream>>>(a);
        } else {
          if ((rc = lds_opt_in(og::k_fft_fwd<T2, 24>, smf))) break;
          og::k_fft_fwd<T2, 24><<<a.nfft, kBlock, smf, h->stream>>>(a);
        }
        if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
      }
      if (rc) break;
      const size_t tot = (size_t)h->nk * h->nx;
      og::k_wgrad<T2><<<grid_for(tot), kBlock, 0, h->stream>>>(PX, PY, h->nk, h->nx, h->na, h->ldphi,
                                                              c / (double)h->nk, Q);
      if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
    }
  }
  if (rc == 0 && hipStreamSynchronize(h->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  dfree(PX);
  dfree(PY);
  dfree(v);
  return rc;
}

}  // namespace

extern "C" {

int ocean_gpu_spectral_grad(ocean_gpu_handle* h, const double* hay_re, const double* hay_im, int32_t nE,
                            const double* energies_eV, double gres_eV, double gprc_eV, double ffff,
                            int32_t depth, double kpref, const double* target, double* loss, double* sigma,
                            double* g_ptab, double* g_scale, double* g_gamma_eV, double* g_scissor_eV) {
  if (!h || !hay_re || !hay_im || nE <= 0 || !energies_eV || depth <= 0 || !target || !loss)
    return OCEAN_GPU_ERR_ARG;
  if (h->obf_on) return OCEAN_GPU_ERR_UNSUPPORTED;  // the W derivative is taken on the Bloch form
  if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  if (g_ptab && (h->nptab <= 0 || !h->w_i0)) return OCEAN_GPU_ERR_STATE;  // W must come from build_w
  const size_t nv = h->nvec();
  double2* rhs = nullptr;
  OG_TRY(gmres_rhs(h, hay_re, hay_im, &rhs));
  double2 *xs = nullptr, *ys = nullptr;
  double* Q = nullptr;
  int rc = dmalloc(&xs, (size_t)nE * nv);
  if (rc == 0) rc = dmalloc(&ys, (size_t)nE * nv);
  const size_t nw = (size_t)h->nk * h->nx;
  if (rc == 0 && g_ptab) {
    rc = dmalloc(&Q, nw);
    if (rc == 0) {
      og::k_fill_d<<<grid_for(nw), kBlock, 0, h->stream>>>(Q, 0.0, nw);
      if (hipGetLastError() != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
    }
  }
  // forward solves at E + i gres and adjoint solves at the conjugate energy
  std::vector<double> rows_f(6 * (size_t)nE), rows_b(6 * (size_t)nE);
  if (rc == 0) rc = gmres_run(h, rhs, nE, energies_eV, gres_eV, gprc_eV, ffff, depth, kpref, rows_f.data(), xs,
                              nullptr, nullptr);
  // (the adjoint's preconditioner is the conjugate one: gprc -> -gprc)
  if (rc == 0) rc = gmres_run(h, rhs, nE, energies_eV, -gres_eV, -gprc_eV, ffff, depth, kpref, rows_b.data(), ys,
                              nullptr, nullptr);
  std::vector<double> r((size_t)nE);
  double L = 0.0;
  for (int e = 0; e < nE; ++e) {
    const double sg = rows_f[6 * (size_t)e + 2];
    if (sigma) sigma[e] = sg;
    r[e] = sg - target[e];
    L += r[e] * r[e];
  }
  *loss = L;
  double gG = 0.0, gD = 0.0, gS = 0.0;
  if (rc == 0 && (g_ptab || g_scale || g_gamma_eV || g_scissor_eV)) {
    const int acnt = (h->ns == 2) ? h->na / 2 : h->na;
    const int NA = na_tmpl(acnt);
    const bool mixed = h->precision == OCEAN_GPU_MIXED;
    const bool ws = g_scale != nullptr, ww = g_ptab != nullptr;
#define OG_GRAD(NAV)                                                                                   \
    if (NA == NAV)                                                                                     \
      rc = mixed ? grad_t<float2, NAV>(h, nE, xs, ys, r, kpref, ws, ww, &gG, &gD, &gS, Q)             \
                 : grad_t<double2, NAV>(h, nE, xs, ys, r, kpref, ws, ww, &gG, &gD, &gS, Q);
    OG_GRAD(2)
    OG_GRAD(4)
    OG_GRAD(6)
    OG_GRAD(12)
#undef OG_GRAD
  }
  // dL/dptab_i = sum_(k,x) Q(k,x) dW(k,x)/dptab_i with the hat weights of build_w
  if (rc == 0 && g_ptab) {
    std::vector<double> q(nw), w0(nw), w1(nw);
    std::vector<int32_t> i0(nw);
    if (hipMemcpy(q.data(), Q, nw * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(i0.data(), h->w_i0, nw * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(w0.data(), h->w_w0, nw * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(w1.data(), h->w_w1, nw * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
    if (rc == 0) {
      for (int i = 0; i < h->nptab; ++i) g_ptab[i] = 0.0;
      for (int k = 0; k < h->nk; ++k)
        for (int x = 0; x < h->nx; ++x) {
          const size_t eq = (size_t)k * h->nx + x, ew = (size_t)x * h->nk + k;  // Q [k][x], W [x][k]
          const int j = i0[ew];
          g_ptab[j] += q[eq] * w0[ew];
          if (j + 1 < h->nptab) g_ptab[j + 1] += q[eq] * w1[ew];
        }
    }
  }
  if (g_gamma_eV) *g_gamma_eV = gG * kEV2Hartree;
  if (g_scissor_eV) *g_scissor_eV = gD * kEV2Hartree;
  if (g_scale) *g_scale = gS;
  dfree(rhs); dfree(xs); dfree(ys); dfree(Q);
  return rc;
}

int ocean_gpu_gmres(ocean_gpu_handle* h, const double* hay_re, const double* hay_im, int32_t nE,
                    const double* energies_eV, double gres_eV, double gprc_eV, double ffff, int32_t depth,
                    double kpref, double* out, double* x_re, double* x_im) {
  if (!h || !hay_re || !hay_im || nE <= 0 || !energies_eV || depth <= 0 || !out) return OCEAN_GPU_ERR_ARG;
  if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  double2* rhs = nullptr;
  OG_TRY(gmres_rhs(h, hay_re, hay_im, &rhs));
  const int rc = gmres_run(h, rhs, nE, energies_eV, gres_eV, gprc_eV, ffff, depth, kpref, out, nullptr, x_re, x_im);
  dfree(rhs);
  return rc;
}

}  // extern "C"

// ===========================================================================
// Batches: many independent problems (one loaded handle each) whose actions and
// Haydock recursions run together, one launch per stage for all problems of a
// group.  A group shares the kernel instantiation: precision, alpha count and
// the transform's edge class.
// ===========================================================================
struct ocean_gpu_batch {
  struct Group {
    int prec = OCEAN_GPU_FP64, NA = 4, lmax = 8;
    std::vector<int> idx;  // problems of the group, in batch order
    Tiles tiles;
    float tune_ms[3] = {0.f, 0.f, 0.f};  // best stage times found by the autotuner
  };
  std::vector<H*> hs;
  std::vector<Group> groups;
  hipStream_t stream = nullptr;
  int* dcount = nullptr;  // [0] Haydock iteration number, [1] its copy for the third step
  double t_last = 0.0;    // ms per batched action (all problems) of the last timed call
};

namespace {

using B = ocean_gpu_batch;

// Device tables of one group for one set of (in, out) vectors.
struct Tables {
  void* args = nullptr;  // og::ActArgs<T2>[np]
  int* off = nullptr;    // [4][np + 1] first workgroup of each problem: proj, fft, adj, matrix-core lr
  int np = 0;
  int nblk[4] = {0, 0, 0, 0};
  size_t sm[3] = {0, 0, 0};
  int threads[3] = {256, kBlock, 512};
  int lmax = 8;
  void release() {
    if (args) (void)hipFree(args);
    if (off) (void)hipFree(off);
    args = nullptr;
    off = nullptr;
  }
};

template <typename F>
int group_call(const B::Group& g, F&& f) {
  const bool mixed = g.prec == OCEAN_GPU_MIXED;
  switch (g.NA) {
    case 2: return mixed ? f(std::integral_constant<int, 2>{}, float2{}) : f(std::integral_constant<int, 2>{}, double2{});
    case 4: return mixed ? f(std::integral_constant<int, 4>{}, float2{}) : f(std::integral_constant<int, 4>{}, double2{});
    case 6: return mixed ? f(std::integral_constant<int, 6>{}, float2{}) : f(std::integral_constant<int, 6>{}, double2{});
    case 12: return mixed ? f(std::integral_constant<int, 12>{}, float2{}) : f(std::integral_constant<int, 12>{}, double2{});
    case 24: return mixed ? f(std::integral_constant<int, 24>{}, float2{}) : f(std::integral_constant<int, 24>{}, double2{});
    default: return OCEAN_GPU_ERR_ARG;
  }
}

template <typename T2, int NA>
int build_tables(const B* b, const B::Group& g, const std::vector<const double2*>& in,
                 const std::vector<double2*>& out, const Tiles& t, Tables& tb) {
  const int np = (int)g.idx.size();
  std::vector<og::ActArgs<T2>> host((size_t)np);
  std::vector<int> off((size_t)4 * (np + 1), 0);
  tb = Tables{};
  tb.np = np;
  tb.lmax = g.lmax;
  for (int q = 0; q < np; ++q) {
    const H* h = b->hs[g.idx[q]];
    const int f = flags_of(h, 0, nullptr);
    if (!(f & og::F_LR)) return OCEAN_GPU_ERR_ARG;
    OG_TRY(check_ready(h, f));
    Geom<T2> ge;
    OG_TRY((make_args<T2, NA>(h, in[q], out[q], f, t, StateOf<T2>::u(h), StateOf<T2>::uT(h),
                              StateOf<T2>::phi(h), StateOf<T2>::W(h), host[q], ge)));
    off[q + 1] = off[q] + ge.nblk_proj;
    off[(np + 1) + q + 1] = off[(np + 1) + q] + ge.nblk_fft;
    off[2 * (np + 1) + q + 1] = off[2 * (np + 1) + q] + ge.nblk_adj;
    // many alphas on one spin: the matrix-core adjoint, as in xact_t
    int nlr = 0;
    if (adj_mf<T2, NA>(h, f, t)) {
      nlr = ((h->nb + 63) / 64) * h->nk;
      off[2 * (np + 1) + q + 1] = off[2 * (np + 1) + q];
    }
    off[3 * (np + 1) + q + 1] = off[3 * (np + 1) + q] + nlr;
    tb.sm[0] = std::max(tb.sm[0], ge.sm_proj);
    tb.sm[1] = std::max(tb.sm[1], ge.sm_fft);
    tb.sm[2] = std::max(tb.sm[2], ge.sm_adj);
  }
  for (int i = 0; i < 4; ++i) tb.nblk[i] = off[(size_t)i * (np + 1) + np];
  tb.threads[0] = t.proj;
  tb.threads[2] = t.adj;
  OG_CHECK(hipMalloc(&tb.args, sizeof(og::ActArgs<T2>) * (size_t)np));
  if (hipMalloc((void**)&tb.off, sizeof(int) * off.size()) != hipSuccess) { tb.release(); return OCEAN_GPU_ERR_NOMEM; }
  if (hipMemcpy(tb.args, host.data(), sizeof(og::ActArgs<T2>) * (size_t)np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(tb.off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice) != hipSuccess) {
    tb.release();
    return OCEAN_GPU_ERR_HIP;
  }
  return OCEAN_GPU_OK;
}

// stage bits: 1 projection, 2 transform (+ multiplet reduction), 4 adjoint
template <typename T2, int NA>
int launch_tables(const Tables& tb, hipStream_t st, int stages) {
  const auto* A = static_cast<const og::ActArgs<T2>*>(tb.args);
  const int np = tb.np;
  if (stages & 1) {
    OG_TRY(lds_opt_in(og::k_proj_batch<T2, NA>, tb.sm[0]));
    og::k_proj_batch<T2, NA><<<tb.nblk[0], tb.threads[0], tb.sm[0], st>>>(A, tb.off, np);
  }
  if (stages & 2) {
    if (tb.lmax == 8) {
      og::k_fft_w_batch<T2, 8><<<tb.nblk[1], kBlock, tb.sm[1], st>>>(A, tb.off + (np + 1), np);
    } else if (tb.lmax == 16) {
      OG_TRY(lds_opt_in(og::k_fft_w_batch<T2, 16>, tb.sm[1]));
      og::k_fft_w_batch<T2, 16><<<tb.nblk[1], kBlock, tb.sm[1], st>>>(A, tb.off + (np + 1), np);
    } else {
      OG_TRY(lds_opt_in(og::k_fft_w_batch<T2, 24>, tb.sm[1]));
      og::k_fft_w_batch<T2, 24><<<tb.nblk[1], kBlock, tb.sm[1], st>>>(A, tb.off + (np + 1), np);
    }
  }
  if (stages & 4) {
    if constexpr (NA >= 6) {
      if (tb.nblk[3] > 0)
        og::k_adj_lr_batch<T2, NA><<<tb.nblk[3], 256, 0, st>>>(A, tb.off + 3 * (np + 1), np);
    }
    if constexpr (NA <= 12) {
      if (tb.nblk[2] > 0) {
        OG_TRY(lds_opt_in(og::k_adj_batch<T2, NA>, tb.sm[2]));
        og::k_adj_batch<T2, NA><<<tb.nblk[2], tb.threads[2], tb.sm[2], st>>>(A, tb.off + 2 * (np + 1), np);
      }
    } else if (tb.nblk[2] > 0) {
      return OCEAN_GPU_ERR_UNSUPPORTED;  // 24 alphas: the matrix-core adjoint only
    }
  }
  OG_CHECK(hipGetLastError());
  return OCEAN_GPU_OK;
}

int group_build(const B* b, const B::Group& g, const std::vector<const double2*>& in,
                const std::vector<double2*>& out, const Tiles& t, Tables& tb) {
  return group_call(g, [&](auto na, auto t2) {
    return build_tables<decltype(t2), decltype(na)::value>(b, g, in, out, t, tb);
  });
}

int group_launch(const B::Group& g, const Tables& tb, hipStream_t st, int stages) {
  return group_call(g, [&](auto na, auto t2) {
    return launch_tables<decltype(t2), decltype(na)::value>(tb, st, stages);
  });
}

// the (in, out) vectors of group g: per problem, in[q] and out[q] of the batch order
void group_vectors(const B::Group& g, const std::vector<const double2*>& in_all,
                   const std::vector<double2*>& out_all, std::vector<const double2*>& in,
                   std::vector<double2*>& out) {
  in.clear();
  out.clear();
  for (int q : g.idx) {
    in.push_back(in_all[q]);
    out.push_back(out_all[q]);
  }
}

// Time `stages` of group g over reps launches (ms per launch), after one warm-up.
// Average ms per launch of the given stages: one untimed launch, then reps timed
// ones, repeated with more launches when reps of them take under 2 ms (a few
// launches of a short stage measure clock ramps and launch gaps, not the kernel).
int time_group(B* b, const B::Group& g, const Tables& tb, int stages, int reps, float* ms) {
  hipEvent_t e0, e1;
  OG_CHECK(hipEventCreate(&e0));
  OG_CHECK(hipEventCreate(&e1));
  int rc = group_launch(g, tb, b->stream, stages);
  for (int pass = 0; pass < 2 && rc == 0; ++pass) {
    if (rc == 0) rc = hipEventRecord(e0, b->stream) == hipSuccess ? 0 : OCEAN_GPU_ERR_HIP;
    for (int r = 0; r < reps && rc == 0; ++r) rc = group_launch(g, tb, b->stream, stages);
    if (rc == 0 && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(ms, e0, e1) != hipSuccess))
      rc = OCEAN_GPU_ERR_HIP;
    if (rc != 0 || *ms >= 2.0f) break;
    reps = std::min(1000, (int)(2.0f * reps / std::max(*ms, 1e-3f)) + 1);
  }
  if (rc == 0) *ms /= (float)reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

int ocean_gpu_batch_create(ocean_gpu_handle* const* hs, int32_t n, ocean_gpu_batch** out) {
  if (!hs || n <= 0 || !out) return OCEAN_GPU_ERR_ARG;
  *out = nullptr;
  B* b = new B();
  for (int i = 0; i < n; ++i) {
    // each problem owns its multiplet scratch and counter: no handle twice
    if (!hs[i] || hs[i]->device != hs[0]->device ||
        std::find(b->hs.begin(), b->hs.end(), hs[i]) != b->hs.end()) {
      delete b;
      return OCEAN_GPU_ERR_ARG;
    }
    if (hs[i]->obf_on) {  // batched groups run the Bloch-form kernels
      delete b;
      return OCEAN_GPU_ERR_UNSUPPORTED;
    }
    b->hs.push_back(hs[i]);
  }
  if (hipStreamCreate(&b->stream) != hipSuccess || hipMalloc((void**)&b->dcount, 2 * sizeof(int)) != hipSuccess) {
    ocean_gpu_batch_destroy(b);
    return OCEAN_GPU_ERR_HIP;
  }
  const int zero[2] = {1, 0};
  if (hipMemcpy(b->dcount, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
    ocean_gpu_batch_destroy(b);
    return OCEAN_GPU_ERR_HIP;
  }
  for (int i = 0; i < n; ++i) {
    const H* h = hs[i];
    const int acnt = (h->ns == 2) ? h->na / 2 : h->na;
    B::Group key;
    key.prec = h->precision;
    key.NA = na_tmpl(acnt);
    key.lmax = edge_class(h->km);
    auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const B::Group& g) {
      return g.prec == key.prec && g.NA == key.NA && g.lmax == key.lmax;
    });
    if (it == b->groups.end()) {
      if (key.prec != OCEAN_GPU_FP64 && key.NA < 6) key.tiles.proj = Tiles{}.proj / 2;  // same x per workgroup
      b->groups.push_back(key);
      it = b->groups.end() - 1;
    }
    it->idx.push_back(i);
    // the largest band count of the group bounds the adjoint thread budget
    int nbmax = 0;
    for (int q : it->idx) nbmax = std::max(nbmax, b->hs[q]->nb);
    it->tiles.adj = nbmax > 256 ? 512 : it->tiles.adj;
  }
  *out = b;
  return OCEAN_GPU_OK;
}

int ocean_gpu_batch_destroy(ocean_gpu_batch* b) {
  if (!b) return OCEAN_GPU_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->dcount) (void)hipFree(b->dcount);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return OCEAN_GPU_OK;
}

int ocean_gpu_batch_ngroups(ocean_gpu_batch* b, int32_t* ngroups) {
  if (!b || !ngroups) return OCEAN_GPU_ERR_ARG;
  *ngroups = (int32_t)b->groups.size();
  return OCEAN_GPU_OK;
}

int ocean_gpu_batch_group_info(ocean_gpu_batch* b, int32_t group, int32_t* info, double* tune_ms) {
  if (!b || !info || group < 0 || group >= (int)b->groups.size()) return OCEAN_GPU_ERR_ARG;
  const B::Group& g = b->groups[group];
  info[0] = (int32_t)g.idx.size();
  info[1] = g.prec;
  info[2] = g.NA;
  info[3] = g.tiles.proj;
  info[4] = g.tiles.adj;
  info[5] = g.tiles.G;
  info[6] = g.tiles.adj_mf;
  if (tune_ms)
    for (int i = 0; i < 3; ++i) tune_ms[i] = g.tune_ms[i];
  return OCEAN_GPU_OK;
}

int ocean_gpu_batch_autotune(ocean_gpu_batch* b, int32_t reps) {
  if (!b || reps <= 0) return OCEAN_GPU_ERR_ARG;
  const int n = (int)b->hs.size();
  std::vector<const double2*> in_all(n);
  std::vector<double2*> out_all(n);
  for (int q = 0; q < n; ++q) {
    H* h = b->hs[q];
    OG_CHECK(hipStreamSynchronize(h->stream));
    og::k_fill<<<grid_for(h->nvec()), kBlock, 0, b->stream>>>(h->tmp_in, make_double2(1e-3, -2e-3), h->nvec());
    OG_CHECK(hipGetLastError());
    in_all[q] = h->tmp_in;
    out_all[q] = h->tmp_out;
  }
  for (auto& g : b->groups) {
    std::vector<const double2*> in;
    std::vector<double2*> out;
    group_vectors(g, in_all, out_all, in, out);
    int nbmax = 0;
    for (int q : g.idx) nbmax = std::max(nbmax, b->hs[q]->nb);
    // each stage's tile only changes that stage, so they are tuned one at a time
    const int proj_c[3] = {128, 256, 512};
    const int adj_c[2] = {256, 512};
    const int G_c[5] = {0, 1, 2, 4, 8};
    // the adjoint's form: VALU at either thread budget, or the matrix-core form
    // where the group holds problems it applies to (many alphas on one spin)
    bool mf_ok = false;
    for (int q : g.idx) mf_ok = mf_ok || (g.NA >= 6 && b->hs[q]->ns == 1 && b->hs[q]->na == g.NA);
    Tiles best = g.tiles;
    float bms[3] = {1e30f, 1e30f, 1e30f};
    for (int stage = 0; stage < 3; ++stage) {
      const int nc = stage == 0 ? 3 : (stage == 1 ? 5 : (mf_ok ? 3 : 2));
      for (int c = 0; c < nc; ++c) {
        Tiles t = best;
        if (stage == 0) t.proj = proj_c[c];
        if (stage == 1) t.G = G_c[c];
        if (stage == 2) {
          if (c == 2) {
            t.adj_mf = 1;
          } else {
            if (adj_c[c] < nbmax) continue;
            t.adj = adj_c[c];
            if (mf_ok) t.adj_mf = 0;
          }
        }
        Tables tb;
        if (group_build(b, g, in, out, t, tb) != OCEAN_GPU_OK) continue;  // shape not possible
        float ms = 0.f;
        const int rc = time_group(b, g, tb, 1 << stage, reps, &ms);
        tb.release();
        if (rc) return rc;
        if (ms < bms[stage]) {
          bms[stage] = ms;
          if (stage == 0) best.proj = t.proj;
          if (stage == 1) best.G = t.G;
          if (stage == 2) {
            best.adj = t.adj;
            best.adj_mf = t.adj_mf;
          }
        }
      }
    }
    g.tiles = best;
    for (int i = 0; i < 3; ++i) g.tune_ms[i] = bms[i];
  }
  return OCEAN_GPU_OK;
}

// One problem's tiles by the batch autotuner on a batch of that problem alone (the
// batched kernels run the single-problem bodies).  The single-problem launch derives
// its FP32 projection workgroup from the handle's tile (tiles_for), so the pick is
// stored in the handle's convention.
int ocean_gpu_autotune(ocean_gpu_handle* h, int32_t reps) {
  if (!h || reps <= 0) return OCEAN_GPU_ERR_ARG;
  ocean_gpu_batch* b = nullptr;
  OG_TRY(ocean_gpu_batch_create(&h, 1, &b));
  int rc = ocean_gpu_batch_autotune(b, reps);
  if (rc == 0 && b->groups.size() == 1) {
    const B::Group& gr = b->groups[0];
    Tiles t = gr.tiles;
    if (gr.prec == OCEAN_GPU_MIXED) t.proj = gr.NA < 6 ? 2 * t.proj : std::max(64, t.proj / 2);
    h->tiles = t;
  }
  ocean_gpu_batch_destroy(b);
  return rc;
}

int ocean_gpu_set_tiles(ocean_gpu_handle* h, const int32_t* tiles) {
  if (!h || !tiles) return OCEAN_GPU_ERR_ARG;
  const int proj = tiles[0], adj = tiles[1], G = tiles[2], mf = tiles[3];
  if (proj < 64 || proj > 1024 || (proj & 63) || adj < 64 || adj > 1024 || (adj & 63) || G < 0 || G > 64 ||
      mf < -1 || mf > 1)
    return OCEAN_GPU_ERR_ARG;
  h->tiles.proj = proj;
  h->tiles.adj = adj;
  h->tiles.G = G;
  h->tiles.adj_mf = mf;
  return OCEAN_GPU_OK;
}

int ocean_gpu_get_tiles(ocean_gpu_handle* h, int32_t* tiles) {
  if (!h || !tiles) return OCEAN_GPU_ERR_ARG;
  tiles[0] = h->tiles.proj;
  tiles[1] = h->tiles.adj;
  tiles[2] = h->tiles.G;
  tiles[3] = h->tiles.adj_mf;
  return OCEAN_GPU_OK;
}

int ocean_gpu_batch_xact(ocean_gpu_batch* b, const int32_t* in_ids, const int32_t* out_ids, int32_t nrep) {
  if (!b || !in_ids || !out_ids || nrep <= 0) return OCEAN_GPU_ERR_ARG;
  const int n = (int)b->hs.size();
  std::vector<const double2*> in_all(n);
  std::vector<double2*> out_all(n);
  for (int q = 0; q < n; ++q) {
    H* h = b->hs[q];
    in_all[q] = vec_ptr(h, in_ids[q]);
    out_all[q] = vec_ptr(h, out_ids[q]);
    if (!in_all[q] || !out_all[q] || in_ids[q] == out_ids[q]) return OCEAN_GPU_ERR_ARG;
    OG_CHECK(hipStreamSynchronize(h->stream));
  }
  std::vector<Tables> tabs(b->groups.size());
  int rc = OCEAN_GPU_OK;
  for (size_t gi = 0; gi < b->groups.size() && rc == 0; ++gi) {
    std::vector<const double2*> in;
    std::vector<double2*> out;
    group_vectors(b->groups[gi], in_all, out_all, in, out);
    rc = group_build(b, b->groups[gi], in, out, b->groups[gi].tiles, tabs[gi]);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0 && hipEventRecord(e0, b->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  for (int r = 0; r < nrep && rc == 0; ++r)
    for (size_t gi = 0; gi < b->groups.size() && rc == 0; ++gi) rc = group_launch(b->groups[gi], tabs[gi], b->stream, 7);
  if (rc == 0 && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
    b->t_last = ms / nrep;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (auto& t : tabs) t.release();
  return rc;
}

int ocean_gpu_batch_last_ms(ocean_gpu_batch* b, double* ms) {
  if (!b || !ms) return OCEAN_GPU_ERR_ARG;
  *ms = b->t_last;
  return OCEAN_GPU_OK;
}

int ocean_gpu_batch_haydock(ocean_gpu_batch* b, const double* const* psi_re, const double* const* psi_im,
                            int32_t niter, double* const* real_a, double* const* real_b,
                            double* const* imag_a, int32_t* niter_done) {
  if (!b || !psi_re || !psi_im || niter <= 0 || !real_a || !real_b) return OCEAN_GPU_ERR_ARG;
  const int n = (int)b->hs.size();
  // one recursion per problem: a problem with several right-hand sides runs
  // ocean_gpu_haydock_multi on its own
  for (H* h : b->hs)
    if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  const size_t n1 = (size_t)niter + 1;
  const size_t nscal = 3 * n1 + 3 * (size_t)kRedBlocks;
  std::vector<double2*> V((size_t)3 * n, nullptr);
  std::vector<double*> Sd((size_t)n, nullptr);
  int rc = OCEAN_GPU_OK;
  for (int q = 0; q < n && rc == 0; ++q) {
    H* h = b->hs[q];
    if (!psi_re[q] || !psi_im[q] || !real_a[q] || !real_b[q]) rc = OCEAN_GPU_ERR_ARG;
    for (int i = 0; i < 3 && rc == 0; ++i) rc = dmalloc(&V[(size_t)3 * q + i], h->nvec());
    if (rc == 0) rc = dmalloc(&Sd[q], nscal);
    if (rc == 0) rc = put_planes(h, psi_re[q], psi_im[q], V[(size_t)3 * q], h->nvec());
  }
  // rotation r of an iteration: (psi, old, hv) = (V0, V1, V2), (V2, V0, V1), (V1, V2, V0)
  const int rot[3][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}};
  std::vector<std::vector<Tables>> tabs(3, std::vector<Tables>(b->groups.size()));
  og::HayArgs* dhay[3] = {nullptr, nullptr, nullptr};
  for (int r = 0; r < 3 && rc == 0; ++r) {
    std::vector<const double2*> in_all(n);
    std::vector<double2*> out_all(n);
    std::vector<og::HayArgs> ha((size_t)n);
    for (int q = 0; q < n; ++q) {
      double2* psi = V[(size_t)3 * q + rot[r][0]];
      double2* old = V[(size_t)3 * q + rot[r][1]];
      double2* hv = V[(size_t)3 * q + rot[r][2]];
      in_all[q] = psi;
      out_all[q] = hv;
      double* s = Sd[q];
      ha[q] = og::HayArgs{hv, old, psi, s + 3 * n1, s + 3 * n1 + 2 * kRedBlocks, s, s + n1, s + 2 * n1,
                          b->hs[q]->nvec()};
    }
    for (size_t gi = 0; gi < b->groups.size() && rc == 0; ++gi) {
      std::vector<const double2*> in;
      std::vector<double2*> out;
      group_vectors(b->groups[gi], in_all, out_all, in, out);
      rc = group_build(b, b->groups[gi], in, out, b->groups[gi].tiles, tabs[r][gi]);
    }
    if (rc == 0) rc = dmalloc(&dhay[r], (size_t)n);
    if (rc == 0 && hipMemcpy(dhay[r], ha.data(), sizeof(og::HayArgs) * n, hipMemcpyHostToDevice) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  // initial state: psi_old = 0, coefficients 0 (b_0 = 0), iteration 1
  if (rc == 0) {
    for (int q = 0; q < n; ++q) {
      og::k_fill<<<grid_for(b->hs[q]->nvec()), kBlock, 0, b->stream>>>(V[(size_t)3 * q + 1], make_double2(0.0, 0.0),
                                                                       b->hs[q]->nvec());
      og::k_fill_d<<<grid_for(3 * n1), kBlock, 0, b->stream>>>(Sd[q], 0.0, 3 * n1);
    }
    const int init[2] = {1, 0};
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(b->dcount, init, sizeof(init), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      rc = OCEAN_GPU_ERR_HIP;
  }
  auto enqueue_iter = [&](int r) -> int {
    for (size_t gi = 0; gi < b->groups.size(); ++gi) OG_TRY(group_launch(b->groups[gi], tabs[r][gi], b->stream, 7));
    og::k_hay1_batch<<<n * kRedBlocks, kBlock, 0, b->stream>>>(dhay[r], b->dcount);
    og::k_hay2_batch<<<n * kRedBlocks, kBlock, 0, b->stream>>>(dhay[r], b->dcount);
    og::k_hay3_batch<<<n * kRedBlocks, kBlock, 0, b->stream>>>(dhay[r], b->dcount);
    OG_CHECK(hipGetLastError());
    return OCEAN_GPU_OK;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = OCEAN_GPU_ERR_HIP;
  // one captured graph holds a full rotation cycle (three iterations)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const int cycles = niter / 3, rem = niter % 3;
  if (rc == 0 && cycles > 0) {
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
    int rc_cap = OCEAN_GPU_OK;
    for (int r = 0; r < 3 && rc == 0 && rc_cap == 0; ++r) rc_cap = enqueue_iter(r);
    if (rc == 0) {
      if (hipStreamEndCapture(b->stream, &graph) != hipSuccess || rc_cap != 0) rc = rc_cap ? rc_cap : OCEAN_GPU_ERR_HIP;
    }
    if (rc == 0 && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  }
  if (rc == 0 && hipEventRecord(e0, b->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  for (int c = 0; c < cycles && rc == 0; ++c)
    if (hipGraphLaunch(exec, b->stream) != hipSuccess) rc = OCEAN_GPU_ERR_HIP;
  for (int r = 0; r < rem && rc == 0; ++r) rc = enqueue_iter(r);
  if (rc == 0 && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
    rc = OCEAN_GPU_ERR_HIP;
  if (rc == 0) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) b->t_last = ms;
  }
  // coefficients back to the host, NaN check per problem as the single recursion
  int rc_nan = OCEAN_GPU_OK;
  std::vector<double> hb(3 * n1);
  for (int q = 0; q < n && rc == 0; ++q) {
    if (hipMemcpy(hb.data(), Sd[q], hb.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = OCEAN_GPU_ERR_HIP;
      break;
    }
    int done = niter;
    for (int i = 0; i < niter; ++i)
      if (hb[i] != hb[i]) { done = i + 1; rc_nan = OCEAN_GPU_ERR_NAN; break; }
    for (size_t i = 0; i < n1; ++i) {
      real_a[q][i] = hb[i];
      real_b[q][i] = hb[2 * n1 + i];
      if (imag_a && imag_a[q]) imag_a[q][i] = hb[n1 + i];
    }
    if (niter_done) niter_done[q] = done;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (auto& tr : tabs)
    for (auto& t : tr) t.release();
  for (int r = 0; r < 3; ++r) dfree(dhay[r]);
  for (auto& v : V) dfree(v);
  for (auto& s : Sd) dfree(s);
  return rc ? rc : rc_nan;
}

// GMRES of every problem of the batch at once: each problem's recycled solve runs on
// its own handle's stream, driven by one of up to kGmresWorkers host threads, so the
// solves of small problems overlap on the device and each solve's host read-backs
// (one per block of iterations) overlap the others' kernels.
int ocean_gpu_batch_gmres(ocean_gpu_batch* b, const double* const* hay_re, const double* const* hay_im,
                          const int32_t* nE, const double* const* energies_eV, double gres_eV, double gprc_eV,
                          double ffff, int32_t depth, const double* kpref, double* const* out, int32_t* rcs) {
  if (!b || !hay_re || !hay_im || !nE || !energies_eV || !kpref || !out || depth <= 0) return OCEAN_GPU_ERR_ARG;
  const int n = (int)b->hs.size();
  for (int q = 0; q < n; ++q)
    if (!hay_re[q] || !hay_im[q] || nE[q] <= 0 || !energies_eV[q] || !out[q]) return OCEAN_GPU_ERR_ARG;
  for (H* h : b->hs)
    if (h->nrhs != 1) return OCEAN_GPU_ERR_UNSUPPORTED;
  constexpr int kGmresWorkers = 16;
  std::vector<int> rc(n, OCEAN_GPU_OK);
  std::atomic<int> next{0};
  const int dev = b->hs[0]->device;
  auto worker = [&]() {
    if (hipSetDevice(dev) != hipSuccess) return;
    for (int q = next++; q < n; q = next++) {
      H* h = b->hs[q];
      double2* rhs = nullptr;
      int r = gmres_rhs(h, hay_re[q], hay_im[q], &rhs);
      if (r == 0) r = gmres_run(h, rhs, nE[q], energies_eV[q], gres_eV, gprc_eV, ffff, depth, kpref[q], out[q],
                                nullptr, nullptr, nullptr);
      dfree(rhs);
      rc[q] = r;
    }
  };
  for (int q = 0; q < n; ++q) rc[q] = OCEAN_GPU_ERR_HIP;  // until its worker reports
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  const int nw = std::min(n, kGmresWorkers);
  for (int w = 0; w < nw; ++w) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  b->t_last = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  int first = OCEAN_GPU_OK;
  for (int q = 0; q < n; ++q) {
    if (rcs) rcs[q] = rc[q];
    if (first == OCEAN_GPU_OK && rc[q] != OCEAN_GPU_OK) first = rc[q];
  }
  return first;
}

