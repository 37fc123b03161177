/*
 * ocean_gpu.h -- C-ABI of the MI355X (gfx950) core-level BSE action and Krylov
 * solvers.  Drop-in boundary for OCEAN2's hot path (OCEAN 3.2.1 tree):
 *
 *   OCEAN_xact            OCEAN2/OCEAN_action.f90:286   -> ocean_gpu_xact, ocean_gpu_xact_vec
 *   OCEAN_haydock_do      OCEAN2/OCEAN_haydock.f90:167  -> ocean_gpu_haydock
 *     (complex_haydock branch, :53-160, haydock_abc_1 :1250) -> ocean_gpu_haydock_abc
 *   OCEAN_gmres_do_recycle OCEAN2/OCEAN_gmres.f90:508   -> ocean_gpu_gmres
 *   write_core            OCEAN2/OCEAN_haydock.f90:1801 -> ocean_gpu_spectrum,
 *                                                          ocean_gpu_spectrum_abc
 *   OCEAN_psi_* vector ops OCEAN2/OCEAN_psi.f90         -> ocean_gpu_vec_*
 *   module state loaded by lr_init / OCEAN_create_central / OCEAN_energies_load
 *                          (long_range.f90:108-260, OCEAN_multiplet.f90:248-570,
 *                           OCEAN_energies.f90:303-368) -> ocean_gpu_load_*
 *
 * Conventions (those of the reference):
 *   - every entry point returns an int error code; 0 is success (the reference's
 *     ierr, OCEAN2/OCEAN.f90:92-93);
 *   - host arrays are caller-owned, column-major, with separate real and
 *     imaginary FP64 planes (OCEAN2/OCEAN_psi.f90:64-65); they are borrowed for
 *     the duration of the call only;
 *   - energies are in Hartree unless a name says eV;
 *   - one handle per host thread; calls on one handle are serialised by the
 *     caller (the reference runs MPI_THREAD_SERIALIZED, OCEAN2/OCEAN_mpi.f90:40).
 *
 * Callable from Fortran through iso_c_binding (ocean_amd/fortran/ocean_gpu_iface.f90)
 * and from Python through ctypes (ocean_amd/core.py).
 */
#ifndef OCEAN_GPU_H
#define OCEAN_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes mirror the reference where it has one. */
#define OCEAN_GPU_OK              0
#define OCEAN_GPU_ERR_NAN        -1     /* NaN in Haydock a: OCEAN2/OCEAN_haydock.f90:912-920 */
#define OCEAN_GPU_ERR_ARG      1001     /* inconsistent sizes or a null pointer */
#define OCEAN_GPU_ERR_STATE    1002     /* required module state not loaded */
#define OCEAN_GPU_ERR_HIP      1003     /* HIP runtime failure */
#define OCEAN_GPU_ERR_NOMEM    1004     /* device allocation failed */
#define OCEAN_GPU_ERR_NOCONV   1005     /* GMRES did not reach ffff within 50 restarts at one
                                           energy (the reference would loop on; an FP32 action
                                           cannot reach FP64-sized thresholds) */
#define OCEAN_GPU_ERR_UNSUPPORTED 9963  /* as OCEAN_gmres.f90:538-542 for backf */
/* GMRES least-squares failure returns the zhptrf-style pivot index (> 0),
 * as OCEAN2/OCEAN_gmres.f90:312-329 returns LAPACK's info. */

/* Term flags of OCEAN_xact (OCEAN2/OCEAN_system.f90:465-471): hflag[0] lr,
 * hflag[4] fg (multiplet), hflag[5] spin-orbit.  Ladder/bubble/ct are valence
 * or legacy terms and are not part of the core path. */

/* Precision forms */
#define OCEAN_GPU_FP64   0   /* the drop-in: everything in FP64 */
#define OCEAN_GPU_MIXED  1   /* u.psi, FFT and W in FP32; recurrences in FP64 */

/* System description: the o_system / o_run fields the core action reads
 * (OCEAN2/OCEAN_system.f90:14-124). */
typedef struct ocean_bse_desc {
  int32_t num_bands;     /* sys%num_bands */
  int32_t kmesh[3];      /* sys%kmesh */
  int32_t xmesh[3];      /* sys%xmesh */
  int32_t nspn;          /* sys%nspn (1 or 2) */
  int32_t lc;            /* core l, sys%cur_run%ZNL(3); nalpha = 4 (2 lc + 1) */
  int32_t lmin, lmax;    /* projector l range (prjfile) */
  int32_t nproj[8];      /* nproj(lmin..lmax) */
  int32_t lvl, lvh;      /* exchange l range (bse.in line 1); must cover lmin..lmax */
  double  celvol;        /* sys%celvol (bohr^3) */
  double  inter_scale;   /* sys%interactionScale */
  int32_t e0;            /* sys%e0: include the band-energy term */
  int32_t mult;          /* sys%mult: include the multiplet/spin-orbit terms */
  int32_t long_range;    /* sys%long_range: include the screened direct term */
  int32_t precision;     /* OCEAN_GPU_FP64 or OCEAN_GPU_MIXED */
} ocean_bse_desc;

typedef struct ocean_gpu_handle ocean_gpu_handle;

int  ocean_gpu_create(const ocean_bse_desc* desc, ocean_gpu_handle** out);
/* A handle whose vectors carry nrhs right-hand sides that share one Hamiltonian: the
 * photon runs of one site and edge, which the reference solves one after another
 * over the same loaded state (run list OCEAN2/OCEAN_system.f90:879-913, loop
 * OCEAN2/OCEAN.f90:52-84).  Every vector is (nb, nk, nrhs * nalpha): right-hand side
 * r holds alphas r * nalpha .. (r + 1) * nalpha - 1, and each u stream of the action
 * serves all of them.  ocean_gpu_load_energies, ocean_gpu_load_mult and
 * ocean_gpu_mult_setup take one right-hand side's arrays (nalpha alphas), which
 * every right-hand side shares.  The alphas per spin, nrhs * nalpha / nspn, must be
 * 2, 4, 6 or 12, or 24 with nspn 1 (K edge: nrhs 1, 3 or 6, or 6 with nspn 2; L2,3:
 * nrhs 2); OCEAN_GPU_ERR_UNSUPPORTED otherwise.  Such a handle runs xact, the vector
 * operations, ocean_gpu_autotune and ocean_gpu_haydock_multi, and batched actions;
 * the single-right-hand-side solvers (Haydock, GMRES, gradient, batched Haydock and
 * GMRES) and the OBF form refuse it (OCEAN_GPU_ERR_UNSUPPORTED).  No reference
 * counterpart: it replaces nrhs separate runs. */
int  ocean_gpu_create_multi(const ocean_bse_desc* desc, int32_t nrhs, ocean_gpu_handle** out);
int  ocean_gpu_nrhs(ocean_gpu_handle* h, int32_t* nrhs);
int  ocean_gpu_destroy(ocean_gpu_handle* h);
const char* ocean_gpu_strerror(int code);
int  ocean_gpu_set_precision(ocean_gpu_handle* h, int32_t precision);

/* Phased Bloch functions re/im_bloch_state(nb, nk, nx, nspn) after
 * OCEAN_bloch_lrLOAD (OCEAN2/OCEAN_bloch.f90:54-219; long_range.f90:186-187). */
int  ocean_gpu_load_bloch(ocean_gpu_handle* h, const double* u_re, const double* u_im);

/* As ocean_gpu_load_bloch with the two planes already in device memory (same
 * layout); for callers that build or keep u on the GPU. */
int  ocean_gpu_load_bloch_dev(ocean_gpu_handle* h, const double* u_re_dev, const double* u_im_dev);
/* Raw Bloch functions (con.u2.dat order: OCEAN2/OCEAN_bloch.f90:998-1115) phased
 * on the device as OCEAN_bloch_lrLOAD does; xoff(3, nx) are the wrapped integer x
 * offsets (xph + ix - 1, ...) of OCEAN2/OCEAN_bloch.f90:133-203. */
int  ocean_gpu_load_bloch_raw(ocean_gpu_handle* h, const double* u_re, const double* u_im,
                              const int32_t* xoff);

/* Orbital-basis-function (OBF) form of the Bloch functions, and from then on the
 * OBF long-range term lr_act_obf2 (OCEAN2/long_range.f90:262-455), which the
 * reference selects with bloch_control's use_obf (:81-82, :128-146) and loads in
 * OCEAN_obf_load (OCEAN2/OCEAN_obf.f90:510-732):
 *   obf_re/im   re/im_obf(num_obf, nx), x in the Bloch x order (:551-600)
 *   obf2u_re/im re/im_obf2u(nb, num_obf, nk) (:704-715)
 *   xoff(3, nx) the wrapped x offsets of OCEAN_obf_make_phase (:267-395), whose
 *               phase the device forms; for an odd kmesh(1) its x shift along the
 *               first axis uses xmesh(2) (:300-301).
 * As in the reference, the OBF term replaces the interaction vector, so the
 * multiplet term is dropped whenever the long-range term runs (:445-446).  Every
 * action entry point of the handle then uses it.  FP64 and nspn = 1 only
 * (OCEAN_GPU_ERR_STATE otherwise, and from ocean_gpu_set_precision while it is
 * on); ocean_gpu_spectral_grad and ocean_gpu_batch_create return
 * OCEAN_GPU_ERR_UNSUPPORTED for such a handle.  A later ocean_gpu_load_bloch*
 * switches the handle back to the Bloch form.  In ocean_gpu_last_timing, "proj"
 * is the band-side projection, "fft" the x side (both products with obf and the
 * transform) and "adj" the adjoint. */
int  ocean_gpu_load_obf(ocean_gpu_handle* h, int32_t num_obf, const double* obf_re, const double* obf_im,
                        const double* obf2u_re, const double* obf2u_im, const int32_t* xoff);

/* W(nk, nx) as filled by lr_populate_W2 (OCEAN2/long_range.f90:1659-1853). */
int  ocean_gpu_load_w(ocean_gpu_handle* h, const double* W);

/* Build W on the device from the radial table (rtab, ptab: rpottrim,
 * OCEAN2/long_range.f90:1688-1696) and the geometry: rdist(nk, nx) is r = |A(R+x-tau)|
 * of long_range.f90:1814.  W = s * {intval cap/cap for r <= rtab(n); 1/(eps r) beyond}.
 * Also keeps the hat weights that the differentiation form contracts with. */
int  ocean_gpu_build_w(ocean_gpu_handle* h, const double* rdist, const double* rtab,
                       const double* ptab, int32_t nptab, double epsilon0, double scale);

/* Multiplet state of OCEAN_create_central (OCEAN2/OCEAN_multiplet.f90:248-570):
 * mpcr/mpci(nb, nk, npmax, -lmax:lmax, lmin:lmax, nspn), mpm(npmax, npmax, lmin:lmax),
 * mhr/mhi packed per l = lvl..lvh at jbeg with mham = 4(2lc+1)(2l+1)nproj(l)
 * (OCEAN_multiplet.f90:318-327), somelr/someli(nalpha, nalpha) (:572-690).
 * have_so = 0 reproduces have_spin_orbit = .false. */
/* The interaction blocks of OCEAN_create_central formed on the device: mh_l of
 * nbsemhsetup2 (OCEAN2/OCEAN_multiplet.f90:2007-2181, converted to Hartree as
 * :388-389 does) for l = lvl..lvh, packed as ocean_gpu_load_mult takes mhr/mhi, and
 * somelr/someli(nalpha, nalpha) of soprep (:572-690); life_time_ev (two values, the
 * content of core_broaden.ipt, :613-627) adds the spin-orbit-dependent lifetimes, NULL
 * when the file is absent; have_so = 0 when xi < 1e-12 Hartree and no lifetimes (:612-615).
 *   slater  per l = lvl..lvh with nproj(l) > 0: the F^k matrices, k = 2, 4, ..,
 *           min(2 lc, 2 l) (with tdlda: the one kk matrix of k = 0, :2082-2094), then
 *           the G^k, k = |lc - l|, .., lc + l step 2; each nproj(l) x nproj(l)
 *           column-major and multiplied by its file's scale factor (fk/kk/gk files,
 *           :2064-2114)
 *   scfac   the overall scale (scfac file, default 1); tdlda the tdlda file's switch;
 *           xi_ev from bse.in (eV). */
int  ocean_gpu_mult_setup(ocean_gpu_handle* h, const double* slater, double scfac, int32_t tdlda,
                          double xi_ev, const double* life_time_ev, double* mhr, double* mhi,
                          double* somelr, double* someli, int32_t* have_so);
int  ocean_gpu_load_mult(ocean_gpu_handle* h, const double* mpcr, const double* mpci,
                         const double* mpm, const double* mhr, const double* mhi,
                         const double* somelr, const double* someli, int32_t have_so);

/* p_energy%r/%i (nb, nk, nalpha) and allow%r (nb, nk, nalpha)
 * (OCEAN2/OCEAN_energies.f90:829-906, :68-112). e_im may be NULL (zero). */
int  ocean_gpu_load_energies(ocean_gpu_handle* h, const double* e_re, const double* e_im,
                             const double* allow);

/* OCEAN_xact (OCEAN2/OCEAN_action.f90:286-618), core branch, on host planes
 * (nb, nk, nalpha).  hflag may be NULL (all terms on). */
int  ocean_gpu_xact(ocean_gpu_handle* h, const double* psi_re, const double* psi_im,
                    double* hpsi_re, double* hpsi_im, int32_t backwards,
                    const int32_t* hflag);

/* ---- device-resident vectors: the OCEAN_vector of OCEAN2/OCEAN_psi.f90:63-133 ---- */
int  ocean_gpu_vec_new(ocean_gpu_handle* h, int32_t* id);
int  ocean_gpu_vec_free(ocean_gpu_handle* h, int32_t id);
int  ocean_gpu_vec_put(ocean_gpu_handle* h, int32_t id, const double* re, const double* im);
int  ocean_gpu_vec_get(ocean_gpu_handle* h, int32_t id, double* re, double* im);
/* <p|q> = sum p_r q_r + p_i q_i + i sum p_r q_i - p_i q_r (OCEAN_psi.f90:1930-2126) */
int  ocean_gpu_vec_dot(ocean_gpu_handle* h, int32_t p, int32_t q, double* rval, double* ival);
/* squared norm (OCEAN_psi.f90:1673-1760) */
int  ocean_gpu_vec_nrm(ocean_gpu_handle* h, int32_t p, double* val);
/* y += (rval + i ival) x (OCEAN_psi.f90:1386-1458) */
int  ocean_gpu_vec_axpy(ocean_gpu_handle* h, double rval, double ival, int32_t x, int32_t y);
/* x *= rval (OCEAN_psi.f90:1876-1917) */
int  ocean_gpu_vec_scal(ocean_gpu_handle* h, double rval, int32_t x);
int  ocean_gpu_vec_copy(ocean_gpu_handle* h, int32_t dst, int32_t src);
/* y = (rval + i ival) x - y  (OCEAN_psi_axmy, OCEAN2/OCEAN_psi.f90:1468-1562) */
int  ocean_gpu_vec_axmy(ocean_gpu_handle* h, double rval, double ival, int32_t x, int32_t y);
/* y = (rval + i ival) x - z  (OCEAN_psi_axmz, OCEAN2/OCEAN_psi.f90:1571-1670) */
int  ocean_gpu_vec_axmz(ocean_gpu_handle* h, double rval, double ival, int32_t x, int32_t y, int32_t z);
/* x = x / (rval + i ival)  (OCEAN_psi_divide, OCEAN2/OCEAN_psi.f90:1764-1874); which
 * selects the reference's optional arguments: 1 rval only, 2 ival only, 3 both. */
int  ocean_gpu_vec_divide(ocean_gpu_handle* h, int32_t x, double rval, double ival, int32_t which);
/* xact between device vectors; the form bench.py times (inputs resident in HBM). */
int  ocean_gpu_xact_vec(ocean_gpu_handle* h, int32_t in, int32_t out, int32_t backwards,
                        const int32_t* hflag);
/* Repeat xact_vec nrep times on one stream with no host synchronisation in
 * between (returns after the last one completes). */
int  ocean_gpu_xact_vec_repeat(ocean_gpu_handle* h, int32_t in, int32_t out, int32_t nrep);
/* nrep actions with HIP events between the kernel families, for per-kernel
 * average durations (read back with ocean_gpu_last_timing). */
int  ocean_gpu_profile_xact(ocean_gpu_handle* h, int32_t in, int32_t out, int32_t nrep);

/* ---- solvers ---- */
/* Hermitian Haydock recursion (OCEAN2/OCEAN_haydock.f90:190-320, :824-938).
 * psi0 is the normalised starting vector.  real_a, real_b, imag_a have niter+1
 * entries (index 0..niter, as the reference's real_a(0:niter)).  niter_done
 * returns the iterations completed (less than niter only on NaN). */
int  ocean_gpu_haydock(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im,
                       int32_t niter, double* real_a, double* real_b, double* imag_a,
                       int32_t* niter_done);

/* The same recursion with the reference's early exit (sys%earlyExit,
 * OCEAN2/OCEAN_haydock.f90:272-291): after each iteration past `spacing`, relArea
 * between the spectra of iterations i and i - spacing (check_convergence,
 * calc_spect_core, :1608-1661, :1991-2027: ne midpoint energies over [el, eh],
 * broadening gam, all in Hartree) is compared with `threshold`; two successes in a
 * row stop the loop.  niter_done returns the iteration it stopped at; entries past
 * it are zero. */
int  ocean_gpu_haydock_early(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im,
                             int32_t niter, int32_t spacing, double threshold, double kpref,
                             double gam, double el, double eh, int32_t ne, double* real_a,
                             double* real_b, double* imag_a, int32_t* niter_done);

/* ocean_gpu_haydock for every right-hand side of an ocean_gpu_create_multi handle
 * at once: psi0 is the stacked starting vector (nb, nk, nrhs * nalpha), each
 * right-hand side normalised on its own; real_a, real_b, imag_a hold nrhs runs of
 * niter+1 entries (right-hand side r at r * (niter + 1)); niter_done[nrhs].  Each
 * right-hand side's coefficients are those of ocean_gpu_haydock on its own vector
 * (OCEAN2/OCEAN_haydock.f90:190-320, :824-938), to rounding.  Returns -1 if any
 * recursion met a NaN. */
int  ocean_gpu_haydock_multi(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im,
                             int32_t niter, double* real_a, double* real_b, double* imag_a,
                             int32_t* niter_done);

/* Non-Hermitian (bi-Lanczos) Haydock recursion: the reference's complex_haydock
 * branch, OCEAN_haydock_nonHerm_do (OCEAN2/OCEAN_haydock.f90:53-160) with
 * haydock_abc_1 (:1250-1433).  psi and back_psi start from psi0 under the allow
 * mask; each iteration applies the action to psi and the backwards action (the
 * conjugated spin-orbit and energy terms, ocean_gpu_xact with backwards = 1) to
 * back_psi.  real_a .. imag_c have niter+1 entries (index 0..niter); b(0) = c(0) = 0.
 * niter_done and the NaN return (-1) as ocean_gpu_haydock.  Refuses a
 * multi-right-hand-side handle (OCEAN_GPU_ERR_UNSUPPORTED). */
int  ocean_gpu_haydock_abc(ocean_gpu_handle* h, const double* psi0_re, const double* psi0_im,
                           int32_t niter, double* real_a, double* imag_a, double* real_b,
                           double* imag_b, double* real_c, double* imag_c, int32_t* niter_done);

/* Recycled preconditioned GMRES over energies (OCEAN2/OCEAN_gmres.f90:508-908).
 * energies_eV[nE]; gres_eV, gprc_eV as bse.in; ffff the |g|^2 threshold; depth the
 * restart length.  out[6*nE]: per energy (E_eV, 1 - Re<b|x> kpref, -Im<b|x> kpref,
 * rel_error, complete_iter, global_iter) as written to absspct (:864-869).
 * x_re/x_im (nb, nk, nalpha, nE) may be NULL.  1 <= depth <= 256 (OCEAN_GPU_ERR_ARG
 * otherwise; the reference default cnbse.gmres.nloop is 80, Common/oparse.json:16). */
int  ocean_gpu_gmres(ocean_gpu_handle* h, const double* hay_re, const double* hay_im,
                     int32_t nE, const double* energies_eV, double gres_eV, double gprc_eV,
                     double ffff, int32_t depth, double kpref, double* out,
                     double* x_re, double* x_im);

/* Gradient of a spectral misfit (no reference counterpart: OCEAN2 computes no
 * gradients).  Over the energy list, sigma(E) = -kpref Im<b|x_E> with
 * (E + i gres - H) x_E = b, b = allow . hay (the spectrum ocean_gpu_gmres
 * reports), and L = sum_E (sigma(E) - target[E])^2.  The adjoint vectors y_E
 * solve the conjugate-energy systems (E - i gres - H) y_E = b with the same
 * recycled GMRES; with r_E = sigma(E) - target[E]:
 *   dL/dgres      =  sum_E 2 r_E kpref Re<y_E|x_E>
 *   dL/dscissor   = -sum_E 2 r_E kpref Im<y_E|x_E>       (every band energy + Delta)
 *   dL/dscale     = -sum_E 2 r_E kpref Im<y_E|(H_lr + H_mult) x_E> / scale
 *                   (W built with scale = inter_scale, mul = inter_scale / (N_k Omega))
 *   dL/dptab[i]   =  sum_(k,x) Q(k,x) dW(k,x)/dptab[i],
 *                   Q = sum_E 2 r_E kpref / N_k Im sum_a conj(P_y) P_x,
 *                   P_v = F_+ U A v (the projection and forward transform of v),
 *                   dW/dptab the hat weights kept by ocean_gpu_build_w.
 * gres, gprc and the gamma / scissor derivatives are per eV.  sigma[nE],
 * g_ptab[nptab] and the scalar outputs may be NULL (not computed); g_ptab needs
 * W from ocean_gpu_build_w. */
int  ocean_gpu_spectral_grad(ocean_gpu_handle* h, const double* hay_re, const double* hay_im,
                             int32_t nE, const double* energies_eV, double gres_eV, double gprc_eV,
                             double ffff, int32_t depth, double kpref, const double* target,
                             double* loss, double* sigma, double* g_ptab, double* g_scale,
                             double* g_gamma_eV, double* g_scissor_eV);

/* Continued fraction with the square-root terminator (write_core,
 * OCEAN2/OCEAN_haydock.f90:1801-1894).  gam, el, eh in Hartree; ne+1 energies.
 * out[5*(ne+1)], row-major per energy: E (eV), Spect, Spect(0), SPKK, SPKK(0).
 * literal_cmplx = 1 rounds the kind-less cmplx() sites (:1846-1847, :1870-1871)
 * to single precision as OCEAN2's build does; 0 keeps everything in FP64. */
int  ocean_gpu_spectrum(const double* real_a, const double* real_b, const double* imag_a,
                        int32_t niter, double kpref, double gam, double el, double eh,
                        int32_t ne, int32_t literal_cmplx, double* out);

/* ocean_gpu_spectrum with complex b and c, the bi-Lanczos coefficients of
 * ocean_gpu_haydock_abc (write_core with the products cmplx(b) cmplx(c),
 * OCEAN2/OCEAN_haydock.f90:1846-1847, :1870).  With imag_b = imag_c = 0 and
 * real_c = real_b it equals ocean_gpu_spectrum. */
int  ocean_gpu_spectrum_abc(const double* real_a, const double* imag_a, const double* real_b,
                            const double* imag_b, const double* real_c, const double* imag_c,
                            int32_t niter, double kpref, double gam, double el, double eh,
                            int32_t ne, int32_t literal_cmplx, double* out);

/* ocean_gpu_spectrum without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and output as
 * ocean_gpu_spectrum; the coefficients and the output are staged in a fixed device
 * array of 2^19 doubles (4 MiB) per GPU, held until the process ends, so
 * 3 (niter + 1) + 5 (ne + 1) must not exceed 2^19 (OCEAN_GPU_ERR_ARG otherwise).
 * Every call on one GPU uses that GPU's array: calls that overlap on one GPU race. */
int  ocean_gpu_spectrum_noalloc(const double* real_a, const double* real_b, const double* imag_a,
                                int32_t niter, double kpref, double gam, double el, double eh,
                                int32_t ne, int32_t literal_cmplx, double* out);

/* ---- batches (no reference counterpart: OCEAN2 runs one site per process) ----
 * A batch runs many independent problems -- each a handle loaded as above, of
 * any size (nb, k-mesh, x-mesh, edge, precision) -- together: one launch per
 * kernel stage for every group of problems that shares a kernel instantiation
 * (precision, alpha count, k-mesh edge class).  The handles stay owned by the
 * caller and must outlive the batch; only the long-range-on action is batched. */
typedef struct ocean_gpu_batch ocean_gpu_batch;
int  ocean_gpu_batch_create(ocean_gpu_handle* const* hs, int32_t n, ocean_gpu_batch** out);
int  ocean_gpu_batch_destroy(ocean_gpu_batch* b);
int  ocean_gpu_batch_ngroups(ocean_gpu_batch* b, int32_t* ngroups);
/* info[7] = (problems, precision, alpha template, projection workgroup, adjoint
 * thread budget, transform strip width (0 = widest), adjoint form (1 matrix cores,
 * 0 VALU, -1 the default, matrix cores)); tune_ms[3] (nullable) = the autotuner's best
 * time of each stage (ms per launch). */
int  ocean_gpu_batch_group_info(ocean_gpu_batch* b, int32_t group, int32_t* info, double* tune_ms);
/* Times every candidate tile shape of each stage on the batch's own problems
 * (reps launches each) and keeps the fastest per group. */
int  ocean_gpu_batch_autotune(ocean_gpu_batch* b, int32_t reps);
/* nrep batched actions: problem q maps its vector in_ids[q] to out_ids[q] (ids in
 * problem q's handle); the time per batched action is read back with
 * ocean_gpu_batch_last_ms. */
int  ocean_gpu_batch_xact(ocean_gpu_batch* b, const int32_t* in_ids, const int32_t* out_ids, int32_t nrep);
int  ocean_gpu_batch_last_ms(ocean_gpu_batch* b, double* ms);
/* Haydock recursions of all problems at once (as ocean_gpu_haydock per problem:
 * pointer arrays of n host planes / coefficient arrays of niter+1 entries).
 * Iterations replay a captured hipGraph of one three-iteration cycle.  Returns
 * -1 if any problem hit a NaN (niter_done[q] says where). */
int  ocean_gpu_batch_haydock(ocean_gpu_batch* b, const double* const* psi_re, const double* const* psi_im,
                             int32_t niter, double* const* real_a, double* const* real_b,
                             double* const* imag_a, int32_t* niter_done);

/* GMRES of all problems at once (as ocean_gpu_gmres per problem; energies, starting
 * vectors and kpref per problem, the solver settings shared): each problem's
 * recycled solve runs on its own handle's stream, driven by a pool of host threads,
 * so the solves overlap on the device.  rcs[n] (nullable) receives each problem's
 * code; the return value is the first nonzero one in problem order.  The wall time
 * of the whole call is read back with ocean_gpu_batch_last_ms. */
int  ocean_gpu_batch_gmres(ocean_gpu_batch* b, const double* const* hay_re, const double* const* hay_im,
                           const int32_t* nE, const double* const* energies_eV, double gres_eV,
                           double gprc_eV, double ffff, int32_t depth, const double* kpref,
                           double* const* out, int32_t* rcs);

/* write_core as OCEAN 3.2.1 writes absspct (OCEAN2/OCEAN_haydock.f90:1801-1894) and,
 * when gam_gauss > gam0/10, write_core_Voigt (:1896-1989).  Energies in Hartree except
 * ebase (eV, added to the printed energy as the reference does). */
typedef struct ocean_spect_desc {
  double  gam0;          /* Lorentzian half width (Hartree), bse.in line 4 */
  double  el, eh;        /* energy window (Hartree) */
  int32_t ne;            /* ne + 1 output energies */
  double  ebase_ev;      /* energy offset of the printed energy (eV) */
  double  gam_gauss;     /* Gaussian width sys%gaussBroaden (Hartree); 0 = off */
  int32_t old_broaden;   /* sys%oldXASbroaden: Spect/SPKK with gamfcn broadening */
  double  nval, eps;     /* gamfcn inputs: nval.h (bohr^-3) and epsilon */
  double  ucvol;         /* cell volume (bohr^3) for the cross-section columns */
  int32_t literal_cmplx; /* 1: the kind-less cmplx sites rounded to single precision */
} ocean_spect_desc;
/* out: (ne+1) rows of *ncols values: 9 columns (E eV, Spect, Spect(0), SPKK, SPKK(0),
 * sigma a.u.^2, sigma cm^2, sigma^2 a.u.^5, sigma^2 cm^4 s), or 5 for the Voigt form
 * (E, Spect, Spect(0), SPKK, SPKK(0), Spect and SPKK with gamfcn).  out holds 9*(ne+1).
 * The Voigt form needs the output step (eh-el)/ne to be at most six Gaussian widths
 * (the reference's fine-grid reuse assumes it); a coarser grid returns
 * OCEAN_GPU_ERR_ARG. */
int  ocean_gpu_write_core(const double* real_a, const double* real_b, const double* imag_a,
                          int32_t niter, double kpref, const ocean_spect_desc* sd, double* out,
                          int32_t* ncols);

/* o_spect.x (OCEAN2/o_spect.f90:26-266) for one abslanc run: spect.in's values (eV),
 * epsilon, nval.h and the cell volume.  Core runs (is_val = 0) give (ne+1) rows of
 * (E, Spect, Spect(0), SPKK, SPKK(0)) as :228-266 writes them; valence runs give ne
 * rows of (omega eV, eps1, eps2, n, kappa, mu, R, loss) as :166-222.  real_a/real_b
 * hold a(0:iter), b(0:iter) (b(0) unused); out holds 8*(ne+1) values. */
typedef struct ocean_ospect_desc {
  int32_t ne;
  double  el_ev, eh_ev, gam0_ev, ebase_ev;
  double  eps0;          /* epsilon file */
  double  nval;          /* nval.h */
  double  ucvol;         /* |det avecsinbohr| */
  int32_t is_val;        /* 1 for VAL/RXS runs (opcons / rxsspct) */
} ocean_ospect_desc;
int  ocean_gpu_ospect(const double* real_a, const double* real_b, int32_t iter, double kpref,
                      const ocean_ospect_desc* od, double* out, int32_t* nrows, int32_t* ncols);

/* Eigenvalues, ascending, of the (n+1) x (n+1) tridiagonal Lanczos matrix with
 * diagonal real_a(0:n) and off-diagonals real_b(1:n): the block redtrid appends to
 * abslanc (OCEAN2/OCEAN_haydock.f90:2273-2316, :2354-2356).  w has n+1 entries. */
int  ocean_gpu_tridiag_eigs(const double* real_a, const double* real_b, int32_t n, double* w);

/* Tile shapes of one problem's action (no reference counterpart): the batch
 * autotuner run on that problem alone (ocean_gpu_batch_autotune, reps launches per
 * candidate) and its picks kept by the handle; tiles[4] = (projection workgroup,
 * adjoint thread budget, transform strip width (0 = widest), adjoint form (1 the
 * matrix-core adjoint with its fused epilogue, 0 the VALU adjoint, -1 the default,
 * the matrix cores; it applies to many alphas on one spin)) as the handle stores
 * them (the FP32 projection runs half as many threads per x at four alphas or fewer
 * and twice as many at more). */
int  ocean_gpu_autotune(ocean_gpu_handle* h, int32_t reps);
int  ocean_gpu_get_tiles(ocean_gpu_handle* h, int32_t* tiles);
/* Sets the handle's tiles[4] as ocean_gpu_get_tiles reports them (OCEAN_GPU_ERR_ARG
 * for values outside projection / adjoint 64..1024 in steps of 64, strip 0..64,
 * form -1..1); a shape the problem cannot take is refused by the next action. */
int  ocean_gpu_set_tiles(ocean_gpu_handle* h, const int32_t* tiles);

/* Average durations (ms) measured with HIP events on the library's stream:
 * "total" from the last ocean_gpu_xact_vec_repeat, "proj", "fft", "mult", "adj"
 * from the last ocean_gpu_profile_xact. */
int  ocean_gpu_last_timing(ocean_gpu_handle* h, const char* name, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* OCEAN_GPU_H */
