/*
 * ocean_oracle.c -- C/OpenMP restatement of OCEAN2's core-level H.psi and Haydock
 * loop, used as the host-core CPU baseline and as a second check of the numpy
 * oracle.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's cpu_baseline leg.
 * The GPU product never links or calls it.
 *
 * Follows, loop for loop (OCEAN 3.2.1 tree):
 *   lr_kernel            OCEAN2/long_range.f90:873-963  (DDOT over bands per k,
 *                        3-D FFT BACKWARD unnormalised, x W, FORWARD, x 1/N_k,
 *                        DAXPY of conj(u) phi')
 *   lr_act_traditional_x OCEAN2/long_range.f90:758-870  (here threads split the x
 *                        points, as OCEAN2's MPI ranks do, each with a private
 *                        partial h.psi summed at the end)
 *   OCEAN_fg_combo       OCEAN2/OCEAN_multiplet.f90:1617-2005
 *   OCEAN_new_soact      OCEAN2/OCEAN_multiplet.f90:1543-1612
 *   OCEAN_energies_act   OCEAN2/OCEAN_energies.f90:890-906
 *   OCEAN_xact order     OCEAN2/OCEAN_action.f90:286-618
 *   OCEAN_hay_ab_twoterm OCEAN2/OCEAN_haydock.f90:824-938
 * The FFT is a mixed-radix (2, 3, 5, generic) Cooley-Tukey transform with the
 * FFTW sign conventions of Modules/FFT_wrapper.f90:67-72.
 *
 * Array layouts are the reference's column-major ones with split re/im planes.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ---------------- 1-D mixed-radix FFT (recursive, out of place) ---------------- */
static void fft1(const cplx* in, cplx* out, int n, int stride, int sign, cplx* scratch) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = 0;
  for (int f = 2; f * f <= n; ++f)
    if (n % f == 0) { p = f; break; }
  if (p == 0) { /* prime length: direct DFT */
    for (int k = 0; k < n; ++k) {
      cplx s = 0;
      for (int j = 0; j < n; ++j) {
        double ang = sign * 2.0 * M_PI * (double)((long)j * k % n) / (double)n;
        s += in[(size_t)j * stride] * (cos(ang) + I * sin(ang));
      }
      out[k] = s;
    }
    return;
  }
  int m = n / p;
  /* p sub-transforms of length m over in[q*stride + j*p*stride] */
  cplx* sub = scratch;
  cplx* scr2 = scratch + n;
  for (int q = 0; q < p; ++q) fft1(in + (size_t)q * stride, sub + (size_t)q * m, m, stride * p, sign, scr2);
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      cplx s = 0;
      int kk = k + r * m;
      for (int q = 0; q < p; ++q) {
        double ang = sign * 2.0 * M_PI * (double)((long)q * kk % n) / (double)n;
        s += sub[(size_t)q * m + k] * (cos(ang) + I * sin(ang));
      }
      out[kk] = s;
    }
  }
}

/* 3-D transform over (k1, k2, k3), k3 fastest, in place on v[nk]. */
static void fft3(cplx* v, const int km[3], int sign, cplx* buf, cplx* scr) {
  const int K1 = km[0], K2 = km[1], K3 = km[2];
  /* along k3 */
  for (int a = 0; a < K1 * K2; ++a) {
    fft1(v + (size_t)a * K3, buf, K3, 1, sign, scr);
    memcpy(v + (size_t)a * K3, buf, sizeof(cplx) * K3);
  }
  /* along k2 */
  for (int a = 0; a < K1; ++a)
    for (int c = 0; c < K3; ++c) {
      cplx* base = v + (size_t)a * K2 * K3 + c;
      fft1(base, buf, K2, K3, sign, scr);
      for (int j = 0; j < K2; ++j) base[(size_t)j * K3] = buf[j];
    }
  /* along k1 */
  for (int b = 0; b < K2 * K3; ++b) {
    cplx* base = v + b;
    fft1(base, buf, K1, K2 * K3, sign, scr);
    for (int j = 0; j < K1; ++j) base[(size_t)j * K2 * K3] = buf[j];
  }
}

void oracle_fft3(double* re, double* im, const int* km, int sign) {
  int nk = km[0] * km[1] * km[2];
  cplx* v = malloc(sizeof(cplx) * nk);
  cplx* buf = malloc(sizeof(cplx) * 64);
  cplx* scr = malloc(sizeof(cplx) * 256);
  for (int i = 0; i < nk; ++i) v[i] = re[i] + I * im[i];
  fft3(v, km, sign, buf, scr);
  for (int i = 0; i < nk; ++i) { re[i] = creal(v[i]); im[i] = cimag(v[i]); }
  free(v); free(buf); free(scr);
}

/* ---------------- the action ---------------- */
typedef struct {
  int nb, nk, nx, na, ns, lc;
  int km[3];
  const double *ur, *ui;     /* (nb, nk, nx, ns) */
  const double* W;           /* (nk, nx) */
  const double *er, *ei;     /* (nb, nk, na) */
  const double* allow;       /* (nb, nk, na) */
  /* multiplet */
  int lmin, lmax, npmax, lvl, lvh;
  const int* nproj;          /* lmin..lmax */
  const double *mpcr, *mpci; /* (nb, nk, npmax, 2lmax+1, nl, ns) */
  const double* mpm;         /* (npmax, npmax, nl) */
  const double *mhr, *mhi;   /* packed per l */
  const double *sr, *si;     /* (na, na) */
  int have_so;
  double celvol, inter;
  int e0, mult, lr;
} oracle_sys;

static int vspin(const oracle_sys* s, int a) { return s->ns == 1 ? 0 : (a & 1); }

/* lr_act: threads split x; each thread: for its x, for each alpha, lr_kernel. */
static void lr_act(const oracle_sys* s, const double* pr, const double* pi, double* hr, double* hi, int nthr) {
  const int nb = s->nb, nk = s->nk, nx = s->nx, na = s->na;
  const size_t N = (size_t)nb * nk * na;
#pragma omp parallel num_threads(nthr)
  {
    int t = 0, T = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    T = omp_get_num_threads();
#endif
    double* lhr = calloc(N, sizeof(double));
    double* lhi = calloc(N, sizeof(double));
    double* xr = malloc(sizeof(double) * nk);
    double* xi = malloc(sizeof(double) * nk);
    cplx* sc = malloc(sizeof(cplx) * nk);
    cplx* buf = malloc(sizeof(cplx) * 64);
    cplx* scr = malloc(sizeof(cplx) * 256);
    const int x0 = (int)((long)nx * t / T), x1 = (int)((long)nx * (t + 1) / T);
    for (int x = x0; x < x1; ++x) {
      for (int a = 0; a < na; ++a) {
        const int sp = vspin(s, a);
        const double* ur = s->ur + (((size_t)sp * nx + x) * nk) * nb;
        const double* ui = s->ui + (((size_t)sp * nx + x) * nk) * nb;
        const double* par = pr + (size_t)a * nk * nb;
        const double* pai = pi + (size_t)a * nk * nb;
        for (int k = 0; k < nk; ++k) {
          double d1 = 0, d2 = 0, d3 = 0, d4 = 0;
          const double* a_r = par + (size_t)k * nb;
          const double* a_i = pai + (size_t)k * nb;
          const double* b_r = ur + (size_t)k * nb;
          const double* b_i = ui + (size_t)k * nb;
          for (int n = 0; n < nb; ++n) {
            d1 += a_r[n] * b_r[n];
            d2 += a_i[n] * b_i[n];
            d3 += a_r[n] * b_i[n];
            d4 += a_i[n] * b_r[n];
          }
          xr[k] = d1 - d2;
          xi[k] = d3 + d4;
        }
        for (int k = 0; k < nk; ++k) sc[k] = xr[k] + I * xi[k];
        fft3(sc, s->km, +1, buf, scr);
        const double* Wx = s->W + (size_t)x * nk;
        for (int k = 0; k < nk; ++k) sc[k] *= Wx[k];
        fft3(sc, s->km, -1, buf, scr);
        const double norm = 1.0 / (double)nk;
        for (int k = 0; k < nk; ++k) { xr[k] = creal(sc[k]) * norm; xi[k] = cimag(sc[k]) * norm; }
        double* ohr = lhr + (size_t)a * nk * nb;
        double* ohi = lhi + (size_t)a * nk * nb;
        for (int k = 0; k < nk; ++k) {
          const double* b_r = ur + (size_t)k * nb;
          const double* b_i = ui + (size_t)k * nb;
          double* yr = ohr + (size_t)k * nb;
          double* yi = ohi + (size_t)k * nb;
          const double fr = xr[k], fi = xi[k];
          for (int n = 0; n < nb; ++n) {
            yr[n] += -fr * b_r[n];
            yr[n] += -fi * b_i[n];
            yi[n] += -fi * b_r[n];
            yi[n] += fr * b_i[n];
          }
        }
      }
    }
#pragma omp critical
    {
      for (size_t i = 0; i < N; ++i) { hr[i] += lhr[i]; hi[i] += lhi[i]; }
    }
    free(lhr); free(lhi); free(xr); free(xi); free(sc); free(buf); free(scr);
  }
}

static void fg_combo(const oracle_sys* s, const double* pr, const double* pi, double* hr, double* hi) {
  const int nb = s->nb, nk = s->nk, na = s->na, npmax = s->npmax, lmin = s->lmin, lmax = s->lmax;
  const int nm = 2 * lmax + 1, nl = lmax - lmin + 1;
  const double mul = s->inter / ((double)nk * s->celvol);
  const int LandM = (lmax + 1) * (lmax + 1) - lmin * lmin;
  const size_t asz = (size_t)npmax * LandM * na;
  double* ampr = calloc(asz, sizeof(double));
  double* ampi = calloc(asz, sizeof(double));
  double* hampr = calloc(asz, sizeof(double));
  double* hampi = calloc(asz, sizeof(double));
#define CIDX(n, k, nu, m, li, sp) ((((((size_t)(sp) * nl + (li)) * nm + ((m) + lmax)) * npmax + (nu)) * nk + (k)) * nb + (n))
#define AIDX(nu, ihd, a) (((size_t)(a) * LandM + (ihd)) * npmax + (nu))
  for (int a = 0; a < na; ++a) {
    const int sp = vspin(s, a);
    for (int l = lmin; l <= lmax; ++l)
      for (int m = -l; m <= l; ++m) {
        const int ihd = l * l + m + l - lmin * lmin;
        for (int nu = 0; nu < s->nproj[l - lmin]; ++nu) {
          double ar = 0, ai = 0;
          for (int k = 0; k < nk; ++k)
            for (int n = 0; n < nb; ++n) {
              const size_t pidx = ((size_t)a * nk + k) * nb + n;
              const size_t c = CIDX(n, k, nu, m, l - lmin, sp);
              ar += pr[pidx] * s->mpcr[c] - pi[pidx] * s->mpci[c];
              ai += pr[pidx] * s->mpci[c] + pi[pidx] * s->mpcr[c];
            }
          ampr[AIDX(nu, ihd, a)] = ar;
          ampi[AIDX(nu, ihd, a)] = ai;
        }
      }
  }
  for (int a = 0; a < na; ++a)
    for (int l = lmin; l <= lmax; ++l) {
      const int np = s->nproj[l - lmin];
      for (int m = -l; m <= l; ++m) {
        const int ihd = l * l + m + l - lmin * lmin;
        for (int mu = 0; mu < np; ++mu) {
          double hrr = 0, hii = 0;
          for (int nu = 0; nu < np; ++nu) {
            const double v = s->mpm[((size_t)(l - lmin) * npmax + nu) * npmax + mu];
            hrr -= v * ampr[AIDX(nu, ihd, a)];
            hii -= v * ampi[AIDX(nu, ihd, a)];
          }
          hampr[AIDX(mu, ihd, a)] = hrr * mul;
          hampi[AIDX(mu, ihd, a)] = hii * mul;
        }
      }
    }
  /* exchange */
  size_t jbeg = 0;
  for (int l = s->lvl; l <= s->lvh; ++l) {
    const int np = (l >= lmin && l <= lmax) ? s->nproj[l - lmin] : 0;
    const int mham = na * (2 * l + 1) * np;
    if (np > 0) {
      for (int ii = 0; ii < mham; ++ii) {
        const int nu = ii % np, m = (ii / np) % (2 * l + 1) - l, a = ii / (np * (2 * l + 1));
        const int ihd = l * l + m + l - lmin * lmin;
        double sr = 0, si = 0;
        for (int jj = 0; jj < mham; ++jj) {
          const int nu2 = jj % np, m2 = (jj / np) % (2 * l + 1) - l, a2 = jj / (np * (2 * l + 1));
          const int ihd2 = l * l + m2 + l - lmin * lmin;
          const size_t j1 = jbeg + (size_t)jj + (size_t)ii * mham;
          const double pwr = ampr[AIDX(nu2, ihd2, a2)], pwi = ampi[AIDX(nu2, ihd2, a2)];
          sr += s->mhr[j1] * pwr - s->mhi[j1] * pwi;
          si += s->mhr[j1] * pwi + s->mhi[j1] * pwr;
        }
        hampr[AIDX(nu, ihd, a)] += sr * mul;
        hampi[AIDX(nu, ihd, a)] += si * mul;
      }
    }
    jbeg += (size_t)mham * mham;
  }
  for (int a = 0; a < na; ++a) {
    const int sp = vspin(s, a);
    for (int l = lmin; l <= lmax; ++l)
      for (int m = -l; m <= l; ++m) {
        const int ihd = l * l + m + l - lmin * lmin;
        for (int nu = 0; nu < s->nproj[l - lmin]; ++nu) {
          const double h_r = hampr[AIDX(nu, ihd, a)], h_i = hampi[AIDX(nu, ihd, a)];
          for (int k = 0; k < nk; ++k)
            for (int n = 0; n < nb; ++n) {
              const size_t pidx = ((size_t)a * nk + k) * nb + n;
              const size_t c = CIDX(n, k, nu, m, l - lmin, sp);
              hr[pidx] += s->mpcr[c] * h_r + s->mpci[c] * h_i;
              hi[pidx] += s->mpcr[c] * h_i - s->mpci[c] * h_r;
            }
        }
      }
  }
#undef CIDX
#undef AIDX
  free(ampr); free(ampi); free(hampr); free(hampi);
}

/* OCEAN_xact, core branch */
static void xact(const oracle_sys* s, const double* pr, const double* pi, double* hr, double* hi,
                 int backwards, int nthr, double* wr, double* wi, double* fr, double* fi) {
  const size_t N = (size_t)s->nb * s->nk * s->na;
  for (size_t i = 0; i < N; ++i) {
    wr[i] = pr[i] * s->allow[i];
    wi[i] = pi[i] * s->allow[i];
    fr[i] = 0.0;
    fi[i] = 0.0;
  }
  if (s->mult) fg_combo(s, wr, wi, fr, fi);
  if (s->lr) lr_act(s, wr, wi, fr, fi, nthr);
  for (size_t i = 0; i < N; ++i) { fr[i] *= s->allow[i]; fi[i] *= s->allow[i]; }
  if (s->mult && s->have_so) {
    const int na = s->na;
    const size_t nkb = (size_t)s->nk * s->nb;
    for (int ic = 0; ic < na; ++ic)
      for (int jc = 0; jc < na; ++jc) {
        const double Sr = s->sr[(size_t)jc * na + ic];
        const double Si = backwards ? -s->si[(size_t)jc * na + ic] : s->si[(size_t)jc * na + ic];
        if (Sr == 0.0 && Si == 0.0) continue;
        for (size_t q = 0; q < nkb; ++q) {
          fr[ic * nkb + q] += Sr * pr[jc * nkb + q] - Si * pi[jc * nkb + q];
          fi[ic * nkb + q] += Sr * pi[jc * nkb + q] + Si * pr[jc * nkb + q];
        }
      }
  }
  for (size_t i = 0; i < N; ++i) {
    double er = 0.0, ei = 0.0;
    if (s->e0) {
      er = s->er[i];
      ei = s->ei ? s->ei[i] : 0.0;
      if (backwards) ei = -ei;
    }
    hr[i] = er * pr[i] - ei * pi[i] + fr[i];
    hi[i] = er * pi[i] + ei * pr[i] + fi[i];
  }
}

/* --------------- exported --------------- */
static oracle_sys make_sys(const int* dims, const double* scal, const double* ur, const double* ui,
                           const double* W, const double* er, const double* ei, const double* allow,
                           const int* nproj, const double* mpcr, const double* mpci, const double* mpm,
                           const double* mhr, const double* mhi, const double* sr, const double* si) {
  /* dims: nb, k1, k2, k3, nx, lc, ns, lmin, lmax, npmax, lvl, lvh, have_so, e0, mult, lr */
  oracle_sys s;
  memset(&s, 0, sizeof(s));
  s.nb = dims[0];
  s.km[0] = dims[1]; s.km[1] = dims[2]; s.km[2] = dims[3];
  s.nk = dims[1] * dims[2] * dims[3];
  s.nx = dims[4];
  s.lc = dims[5];
  s.na = 4 * (2 * dims[5] + 1);
  s.ns = dims[6];
  s.lmin = dims[7]; s.lmax = dims[8]; s.npmax = dims[9]; s.lvl = dims[10]; s.lvh = dims[11];
  s.have_so = dims[12]; s.e0 = dims[13]; s.mult = dims[14]; s.lr = dims[15];
  s.celvol = scal[0];
  s.inter = scal[1];
  s.ur = ur; s.ui = ui; s.W = W; s.er = er; s.ei = ei; s.allow = allow;
  s.nproj = nproj; s.mpcr = mpcr; s.mpci = mpci; s.mpm = mpm; s.mhr = mhr; s.mhi = mhi;
  s.sr = sr; s.si = si;
  return s;
}

int oracle_xact(const int* dims, const double* scal, const double* ur, const double* ui, const double* W,
                const double* er, const double* ei, const double* allow, const int* nproj,
                const double* mpcr, const double* mpci, const double* mpm, const double* mhr,
                const double* mhi, const double* sr, const double* si, const double* pr, const double* pi,
                double* hr, double* hi, int backwards, int nthr, int nrep) {
  oracle_sys s = make_sys(dims, scal, ur, ui, W, er, ei, allow, nproj, mpcr, mpci, mpm, mhr, mhi, sr, si);
  const size_t N = (size_t)s.nb * s.nk * s.na;
  double* w = malloc(sizeof(double) * 4 * N);
  if (!w) return 1;
  for (int r = 0; r < (nrep > 0 ? nrep : 1); ++r) xact(&s, pr, pi, hr, hi, backwards, nthr, w, w + N, w + 2 * N, w + 3 * N);
  free(w);
  return 0;
}

int oracle_haydock(const int* dims, const double* scal, const double* ur, const double* ui, const double* W,
                   const double* er, const double* ei, const double* allow, const int* nproj,
                   const double* mpcr, const double* mpci, const double* mpm, const double* mhr,
                   const double* mhi, const double* sr, const double* si, const double* p0r,
                   const double* p0i, int niter, double* ra, double* rb, double* ia, int nthr) {
  oracle_sys s = make_sys(dims, scal, ur, ui, W, er, ei, allow, nproj, mpcr, mpci, mpm, mhr, mhi, sr, si);
  const size_t N = (size_t)s.nb * s.nk * s.na;
  double* buf = malloc(sizeof(double) * 10 * N);
  if (!buf) return 1;
  double *psr = buf, *psi = buf + N, *olr = buf + 2 * N, *oli = buf + 3 * N, *hr = buf + 4 * N, *hi = buf + 5 * N;
  double* w = buf + 6 * N;
  memcpy(psr, p0r, sizeof(double) * N);
  memcpy(psi, p0i, sizeof(double) * N);
  memset(olr, 0, sizeof(double) * 2 * N);
  for (int i = 0; i <= niter; ++i) { ra[i] = 0; rb[i] = 0; ia[i] = 0; }
  for (int it = 1; it <= niter; ++it) {
    xact(&s, psr, psi, hr, hi, 0, nthr, w, w + N, w + 2 * N, w + 3 * N);
    const double mb = -rb[it - 1];
    for (size_t i = 0; i < N; ++i) { hr[i] += mb * olr[i]; hi[i] += mb * oli[i]; }
    double ar = 0, ai = 0;
    for (size_t i = 0; i < N; ++i) { ar += hr[i] * psr[i] + hi[i] * psi[i]; ai += hr[i] * psi[i] - hi[i] * psr[i]; }
    ra[it - 1] = ar;
    ia[it - 1] = ai;
    for (size_t i = 0; i < N; ++i) { hr[i] += -ar * psr[i]; hi[i] += -ar * psi[i]; }
    double bb = 0;
    for (size_t i = 0; i < N; ++i) bb += hr[i] * hr[i] + hi[i] * hi[i];
    memcpy(olr, psr, sizeof(double) * N);
    memcpy(oli, psi, sizeof(double) * N);
    rb[it] = sqrt(bb);
    const double inv = 1.0 / rb[it];
    for (size_t i = 0; i < N; ++i) { psr[i] = hr[i] * inv; psi[i] = hi[i] * inv; }
    if (ar != ar) { free(buf); return -1; }
  }
  free(buf);
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
