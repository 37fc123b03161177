! iso_c_binding interface to libocean_gpu.so (include/ocean_gpu.h).
!
! The C-ABI takes the reference's own arrays: column-major, separate real and
! imaginary FP64 planes (OCEAN2/OCEAN_psi.f90:64-65), so OCEAN2's module state
! (re/im_bloch_state, W, mpcr/mpci, mpm, mhr/mhi, somelr/someli, p_energy,
! allow) and its OCEAN_vector %r/%i planes are passed without copies.
module ocean_gpu_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: OCEAN_GPU_OK = 0
  integer(c_int), parameter :: OCEAN_GPU_ERR_NAN = -1
  integer(c_int32_t), parameter :: OCEAN_GPU_FP64 = 0
  integer(c_int32_t), parameter :: OCEAN_GPU_MIXED = 1

  ! o_system / o_run fields the core action reads (OCEAN2/OCEAN_system.f90:14-124)
  type, bind(C) :: ocean_bse_desc
    integer(c_int32_t) :: num_bands
    integer(c_int32_t) :: kmesh(3)
    integer(c_int32_t) :: xmesh(3)
    integer(c_int32_t) :: nspn
    integer(c_int32_t) :: lc
    integer(c_int32_t) :: lmin
    integer(c_int32_t) :: lmax
    integer(c_int32_t) :: nproj(8)
    integer(c_int32_t) :: lvl
    integer(c_int32_t) :: lvh
    real(c_double) :: celvol
    real(c_double) :: inter_scale
    integer(c_int32_t) :: e0
    integer(c_int32_t) :: mult
    integer(c_int32_t) :: long_range
    integer(c_int32_t) :: precision
  end type ocean_bse_desc

  ! ocean_spect_desc of ocean_gpu.h: write_core's inputs beyond the coefficients
  type, bind(C) :: ocean_spect_desc
    real(c_double) :: gam0
    real(c_double) :: el
    real(c_double) :: eh
    integer(c_int32_t) :: ne
    real(c_double) :: ebase_ev
    real(c_double) :: gam_gauss
    integer(c_int32_t) :: old_broaden
    real(c_double) :: nval
    real(c_double) :: eps
    real(c_double) :: ucvol
    integer(c_int32_t) :: literal_cmplx
  end type ocean_spect_desc

  ! ocean_ospect_desc of ocean_gpu.h: o_spect.x's run-directory inputs
  type, bind(C) :: ocean_ospect_desc
    integer(c_int32_t) :: ne
    real(c_double) :: el_ev
    real(c_double) :: eh_ev
    real(c_double) :: gam0_ev
    real(c_double) :: ebase_ev
    real(c_double) :: eps0
    real(c_double) :: nval
    real(c_double) :: ucvol
    integer(c_int32_t) :: is_val
  end type ocean_ospect_desc

  interface
    integer(c_int) function ocean_gpu_create(desc, h) bind(C, name='ocean_gpu_create')
      import :: c_int, c_ptr, ocean_bse_desc
      type(ocean_bse_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function ocean_gpu_destroy(h) bind(C, name='ocean_gpu_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function ocean_gpu_set_precision(h, precision) bind(C, name='ocean_gpu_set_precision')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: precision
    end function

    integer(c_int) function ocean_gpu_load_bloch(h, u_re, u_im) bind(C, name='ocean_gpu_load_bloch')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u_re(*), u_im(*)
    end function

    ! OBF form (lr_act_obf2): re/im_obf(num_obf, nx), re/im_obf2u(nb, num_obf, nk),
    ! xoff(3, nx) the wrapped x offsets of OCEAN_obf_make_phase
    integer(c_int) function ocean_gpu_load_obf(h, num_obf, obf_re, obf_im, obf2u_re, obf2u_im, xoff) &
        bind(C, name='ocean_gpu_load_obf')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: num_obf
      real(c_double), intent(in) :: obf_re(*), obf_im(*), obf2u_re(*), obf2u_im(*)
      integer(c_int32_t), intent(in) :: xoff(*)
    end function

    ! nbsemhsetup2 blocks and soprep formed on the device (see ocean_gpu.h)
    ! life_time: c_loc of the two core_broaden.ipt values, or c_null_ptr
    integer(c_int) function ocean_gpu_mult_setup(h, slater, scfac, tdlda, xi_ev, life_time, mhr, mhi, &
        somelr, someli, have_so) bind(C, name='ocean_gpu_mult_setup')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: slater(*)
      real(c_double), value :: scfac, xi_ev
      integer(c_int32_t), value :: tdlda
      type(c_ptr), value :: life_time
      real(c_double), intent(out) :: mhr(*), mhi(*), somelr(*), someli(*)
      integer(c_int32_t), intent(out) :: have_so
    end function

    integer(c_int) function ocean_gpu_load_w(h, w) bind(C, name='ocean_gpu_load_w')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: w(*)
    end function

    integer(c_int) function ocean_gpu_load_mult(h, mpcr, mpci, mpm, mhr, mhi, somelr, someli, have_so) &
        bind(C, name='ocean_gpu_load_mult')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: mpcr(*), mpci(*), mpm(*), mhr(*), mhi(*), somelr(*), someli(*)
      integer(c_int32_t), value :: have_so
    end function

    integer(c_int) function ocean_gpu_load_energies(h, e_re, e_im, allow) bind(C, name='ocean_gpu_load_energies')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: e_re(*), e_im(*), allow(*)
    end function

    integer(c_int) function ocean_gpu_xact(h, psi_re, psi_im, hpsi_re, hpsi_im, backwards, hflag) &
        bind(C, name='ocean_gpu_xact')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: psi_re(*), psi_im(*)
      real(c_double), intent(out) :: hpsi_re(*), hpsi_im(*)
      integer(c_int32_t), value :: backwards
      integer(c_int32_t), intent(in) :: hflag(6)
    end function

    integer(c_int) function ocean_gpu_haydock(h, psi_re, psi_im, niter, real_a, real_b, imag_a, niter_done) &
        bind(C, name='ocean_gpu_haydock')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: psi_re(*), psi_im(*)
      integer(c_int32_t), value :: niter
      real(c_double), intent(out) :: real_a(*), real_b(*), imag_a(*)
      integer(c_int32_t), intent(out) :: niter_done
    end function

    integer(c_int) function ocean_gpu_haydock_early(h, psi_re, psi_im, niter, spacing, threshold, kpref, &
        gam, el, eh, ne, real_a, real_b, imag_a, niter_done) bind(C, name='ocean_gpu_haydock_early')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: psi_re(*), psi_im(*)
      integer(c_int32_t), value :: niter, spacing
      real(c_double), value :: threshold, kpref, gam, el, eh
      integer(c_int32_t), value :: ne
      real(c_double), intent(out) :: real_a(*), real_b(*), imag_a(*)
      integer(c_int32_t), intent(out) :: niter_done
    end function

    integer(c_int) function ocean_gpu_gmres(h, hay_re, hay_im, ne, energies_ev, gres_ev, gprc_ev, ffff, &
        depth, kpref, out, x_re, x_im) bind(C, name='ocean_gpu_gmres')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: hay_re(*), hay_im(*)
      integer(c_int32_t), value :: ne
      real(c_double), intent(in) :: energies_ev(*)
      real(c_double), value :: gres_ev, gprc_ev, ffff
      integer(c_int32_t), value :: depth
      real(c_double), value :: kpref
      real(c_double), intent(out) :: out(*)
      type(c_ptr), value :: x_re, x_im
    end function

    integer(c_int) function ocean_gpu_spectrum(real_a, real_b, imag_a, niter, kpref, gam, el, eh, ne, &
        literal_cmplx, out) bind(C, name='ocean_gpu_spectrum')
      import :: c_int, c_double, c_int32_t
      real(c_double), intent(in) :: real_a(*), real_b(*), imag_a(*)
      integer(c_int32_t), value :: niter
      real(c_double), value :: kpref, gam, el, eh
      integer(c_int32_t), value :: ne, literal_cmplx
      real(c_double), intent(out) :: out(*)
    end function

    ! complex_haydock branch: OCEAN_haydock_nonHerm_do / haydock_abc_1
    ! (OCEAN2/OCEAN_haydock.f90:53-160, :1250-1433)
    integer(c_int) function ocean_gpu_haydock_abc(h, psi_re, psi_im, niter, real_a, imag_a, real_b, &
        imag_b, real_c, imag_c, niter_done) bind(C, name='ocean_gpu_haydock_abc')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: psi_re(*), psi_im(*)
      integer(c_int32_t), value :: niter
      real(c_double), intent(out) :: real_a(*), imag_a(*), real_b(*), imag_b(*), real_c(*), imag_c(*)
      integer(c_int32_t), intent(out) :: niter_done
    end function

    ! write_core's fraction with complex b and c (OCEAN2/OCEAN_haydock.f90:1836-1880)
    integer(c_int) function ocean_gpu_spectrum_abc(real_a, imag_a, real_b, imag_b, real_c, imag_c, &
        niter, kpref, gam, el, eh, ne, literal_cmplx, out) bind(C, name='ocean_gpu_spectrum_abc')
      import :: c_int, c_double, c_int32_t
      real(c_double), intent(in) :: real_a(*), imag_a(*), real_b(*), imag_b(*), real_c(*), imag_c(*)
      integer(c_int32_t), value :: niter
      real(c_double), value :: kpref, gam, el, eh
      integer(c_int32_t), value :: ne, literal_cmplx
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function ocean_gpu_write_core(real_a, real_b, imag_a, niter, kpref, sd, out, ncols) &
        bind(C, name='ocean_gpu_write_core')
      import :: c_int, c_double, c_int32_t, ocean_spect_desc
      real(c_double), intent(in) :: real_a(*), real_b(*), imag_a(*)
      integer(c_int32_t), value :: niter
      real(c_double), value :: kpref
      type(ocean_spect_desc), intent(in) :: sd
      real(c_double), intent(out) :: out(*)
      integer(c_int32_t), intent(out) :: ncols
    end function

    integer(c_int) function ocean_gpu_ospect(real_a, real_b, it, kpref, od, out, nrows, ncols) &
        bind(C, name='ocean_gpu_ospect')
      import :: c_int, c_double, c_int32_t, ocean_ospect_desc
      real(c_double), intent(in) :: real_a(*), real_b(*)
      integer(c_int32_t), value :: it
      real(c_double), value :: kpref
      type(ocean_ospect_desc), intent(in) :: od
      real(c_double), intent(out) :: out(*)
      integer(c_int32_t), intent(out) :: nrows, ncols
    end function

    integer(c_int) function ocean_gpu_tridiag_eigs(real_a, real_b, n, w) bind(C, name='ocean_gpu_tridiag_eigs')
      import :: c_int, c_double, c_int32_t
      real(c_double), intent(in) :: real_a(*), real_b(*)
      integer(c_int32_t), value :: n
      real(c_double), intent(out) :: w(*)
    end function

    ! ---- batches (many sites / edges / polarisations at once) ----
    integer(c_int) function ocean_gpu_batch_create(hs, n, b) bind(C, name='ocean_gpu_batch_create')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: b
    end function

    integer(c_int) function ocean_gpu_batch_destroy(b) bind(C, name='ocean_gpu_batch_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function

    integer(c_int) function ocean_gpu_batch_autotune(b, reps) bind(C, name='ocean_gpu_batch_autotune')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: reps
    end function

    integer(c_int) function ocean_gpu_batch_haydock(b, psi_re, psi_im, niter, real_a, real_b, imag_a, &
        niter_done) bind(C, name='ocean_gpu_batch_haydock')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      type(c_ptr), intent(in) :: psi_re(*), psi_im(*), real_a(*), real_b(*), imag_a(*)
      integer(c_int32_t), value :: niter
      integer(c_int32_t), intent(out) :: niter_done(*)
    end function

    ! ---- gradient of a GMRES spectral misfit ----
    integer(c_int) function ocean_gpu_spectral_grad(h, hay_re, hay_im, ne, energies_ev, gres_ev, gprc_ev, &
        ffff, depth, kpref, target, loss, sigma, g_ptab, g_scale, g_gamma_ev, g_scissor_ev) &
        bind(C, name='ocean_gpu_spectral_grad')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: hay_re(*), hay_im(*), energies_ev(*), target(*)
      integer(c_int32_t), value :: ne, depth
      real(c_double), value :: gres_ev, gprc_ev, ffff, kpref
      real(c_double), intent(out) :: loss, sigma(*), g_ptab(*), g_scale, g_gamma_ev, g_scissor_ev
    end function
  end interface
end module ocean_gpu_iface
