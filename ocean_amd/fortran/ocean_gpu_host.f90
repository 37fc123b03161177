! Fortran host mirror of OCEAN2's hot-path call surface over the HIP C-ABI.
!
!   OCEAN_xact_gpu            <- OCEAN_xact(sys, inter_scale, psi, new_psi, ierr, backwards, nhflag)
!                                OCEAN2/OCEAN_action.f90:286
!   OCEAN_haydock_do_gpu      <- OCEAN_haydock_do / OCEAN_haydock_Herm_do
!                                OCEAN2/OCEAN_haydock.f90:167, :190-320
!   OCEAN_haydock_nonHerm_do_gpu <- OCEAN_haydock_nonHerm_do (complex_haydock branch)
!                                OCEAN2/OCEAN_haydock.f90:53-160, :1250-1433
!   OCEAN_gmres_do_recycle_gpu <- OCEAN_gmres_do_recycle, OCEAN2/OCEAN_gmres.f90:508
!   write_core_gpu      <- write_core, OCEAN2/OCEAN_haydock.f90:1801-1894
!   write_lanczos_gpu   <- write_lanczos, OCEAN2/OCEAN_haydock.f90:2318-2358
!   write_lanczos_abc_gpu <- write_lanczos, complex branch, OCEAN2/OCEAN_haydock.f90:2337-2344
!
! ierr follows the reference: 0 on success, nonzero aborts up the stack.
module ocean_gpu_host
  use iso_c_binding
  use ocean_gpu_iface
  implicit none
  private

  type(c_ptr), save :: handle = c_null_ptr

  integer, parameter :: DP = kind(1.0d0)
  real(DP), parameter :: Hartree2eV = 27.211386245980_DP   ! OCEAN2/OCEAN_constants.f90:16

  public :: OCEAN_gpu_init, OCEAN_gpu_load, OCEAN_xact_gpu, OCEAN_haydock_do_gpu, &
            OCEAN_haydock_nonHerm_do_gpu, OCEAN_gmres_do_recycle_gpu, write_core_gpu, &
            write_lanczos_gpu, write_lanczos_abc_gpu, OCEAN_gpu_finalize, OCEAN_precision_gpu

contains

  subroutine OCEAN_gpu_init(desc, ierr)
    type(ocean_bse_desc), intent(in) :: desc
    integer, intent(inout) :: ierr
    if (c_associated(handle)) then
      ierr = ocean_gpu_destroy(handle)
      handle = c_null_ptr
    end if
    ierr = ocean_gpu_create(desc, handle)
  end subroutine OCEAN_gpu_init

  subroutine OCEAN_precision_gpu(precision, ierr)
    integer, intent(in) :: precision
    integer, intent(inout) :: ierr
    ierr = ocean_gpu_set_precision(handle, int(precision, c_int32_t))
  end subroutine OCEAN_precision_gpu

  ! The module state that lr_init / OCEAN_create_central / OCEAN_energies_load build.
  subroutine OCEAN_gpu_load(u_re, u_im, w, mpcr, mpci, mpm, mhr, mhi, somelr, someli, have_so, &
                            e_re, e_im, allow, ierr)
    real(DP), intent(in) :: u_re(*), u_im(*), w(*), mpcr(*), mpci(*), mpm(*), mhr(*), mhi(*)
    real(DP), intent(in) :: somelr(*), someli(*), e_re(*), e_im(*), allow(*)
    logical, intent(in) :: have_so
    integer, intent(inout) :: ierr
    integer(c_int32_t) :: so
    so = 0
    if (have_so) so = 1
    ierr = ocean_gpu_load_bloch(handle, u_re, u_im)
    if (ierr /= 0) return
    ierr = ocean_gpu_load_w(handle, w)
    if (ierr /= 0) return
    ierr = ocean_gpu_load_mult(handle, mpcr, mpci, mpm, mhr, mhi, somelr, someli, so)
    if (ierr /= 0) return
    ierr = ocean_gpu_load_energies(handle, e_re, e_im, allow)
  end subroutine OCEAN_gpu_load

  ! new_psi = H psi on the %r/%i full-storage planes (nb, nk, nalpha).
  subroutine OCEAN_xact_gpu(psi_r, psi_i, new_r, new_i, ierr, backwards, nhflag)
    real(DP), intent(in) :: psi_r(*), psi_i(*)
    real(DP), intent(out) :: new_r(*), new_i(*)
    integer, intent(inout) :: ierr
    logical, intent(in), optional :: backwards
    integer, intent(in), optional :: nhflag(6)
    integer(c_int32_t) :: hf(6), back
    hf = [1, 0, 0, 0, 1, 1]
    if (present(nhflag)) hf = int(nhflag, c_int32_t)
    back = 0
    if (present(backwards)) then
      if (backwards) back = 1
    end if
    ierr = ocean_gpu_xact(handle, psi_r, psi_i, new_r, new_i, back, hf)
  end subroutine OCEAN_xact_gpu

  ! Hermitian Haydock recursion; real_a/real_b/imag_a are (0:niter) as in the reference.
  ! With conv_spacing/conv_thresh present (haydockconv.ipt), the recursion stops where
  ! the reference's early exit does (OCEAN_haydock.f90:272-291): the relArea test of
  ! check_convergence over ne midpoint energies in [el, eh] (Hartree), broadening gam.
  subroutine OCEAN_haydock_do_gpu(psi_r, psi_i, niter, real_a, real_b, imag_a, niter_done, ierr, &
                                  conv_spacing, conv_thresh, kpref, gam, el, eh, ne)
    real(DP), intent(in) :: psi_r(*), psi_i(*)
    integer, intent(in) :: niter
    real(DP), intent(out) :: real_a(0:niter), real_b(0:niter), imag_a(0:niter)
    integer, intent(out) :: niter_done
    integer, intent(inout) :: ierr
    integer, intent(in), optional :: conv_spacing, ne
    real(DP), intent(in), optional :: conv_thresh, kpref, gam, el, eh
    integer(c_int32_t) :: done
    if (present(conv_spacing) .and. present(conv_thresh) .and. present(kpref) .and. present(gam) &
        .and. present(el) .and. present(eh) .and. present(ne)) then
      ierr = ocean_gpu_haydock_early(handle, psi_r, psi_i, int(niter, c_int32_t), int(conv_spacing, c_int32_t), &
                                     conv_thresh, kpref, gam, el, eh, int(ne, c_int32_t), real_a, real_b, &
                                     imag_a, done)
    else
      ierr = ocean_gpu_haydock(handle, psi_r, psi_i, int(niter, c_int32_t), real_a, real_b, imag_a, done)
    end if
    niter_done = done
  end subroutine OCEAN_haydock_do_gpu

  ! Non-Hermitian (bi-Lanczos) recursion, the reference's complex_haydock branch
  ! (OCEAN_haydock_nonHerm_do, OCEAN_haydock.f90:53-160, with haydock_abc_1 at
  ! :1250-1433); the six coefficient arrays are (0:niter) as in the reference.
  subroutine OCEAN_haydock_nonHerm_do_gpu(psi_r, psi_i, niter, real_a, imag_a, real_b, imag_b, &
                                          real_c, imag_c, niter_done, ierr)
    real(DP), intent(in) :: psi_r(*), psi_i(*)
    integer, intent(in) :: niter
    real(DP), intent(out), dimension(0:niter) :: real_a, imag_a, real_b, imag_b, real_c, imag_c
    integer, intent(out) :: niter_done
    integer, intent(inout) :: ierr
    integer(c_int32_t) :: done
    ierr = ocean_gpu_haydock_abc(handle, psi_r, psi_i, int(niter, c_int32_t), real_a, imag_a, real_b, &
                                 imag_b, real_c, imag_c, done)
    niter_done = done
  end subroutine OCEAN_haydock_nonHerm_do_gpu

  ! Recycled GMRES over energies (eV); writes the absspct rows of OCEAN_gmres.f90:864-869
  ! to unit fh when fh > 0.
  subroutine OCEAN_gmres_do_recycle_gpu(hay_r, hay_i, energies_ev, gres_ev, gprc_ev, ffff, depth, &
                                        kpref, rows, fh, ierr)
    real(DP), intent(in) :: hay_r(*), hay_i(*), energies_ev(:)
    real(DP), intent(in) :: gres_ev, gprc_ev, ffff, kpref
    integer, intent(in) :: depth, fh
    real(DP), intent(out) :: rows(6, size(energies_ev))
    integer, intent(inout) :: ierr
    integer :: i
    ierr = ocean_gpu_gmres(handle, hay_r, hay_i, int(size(energies_ev), c_int32_t), energies_ev, &
                           gres_ev, gprc_ev, ffff, int(depth, c_int32_t), kpref, rows, c_null_ptr, c_null_ptr)
    if (ierr /= 0) return
    if (fh > 0) then
      do i = 1, size(energies_ev)
        write(fh, '(1p,4(1e15.8,1x),1i6,1x,i10)') rows(1, i), rows(2, i), rows(3, i), rows(4, i), &
              nint(rows(5, i)), nint(rows(6, i))
      end do
    end if
  end subroutine OCEAN_gmres_do_recycle_gpu

  ! absspct writer (write_core, OCEAN_haydock.f90:1801-1894): nine columns, or the five
  ! of write_core_Voigt (:1896-1989) when gam_gauss > gam0/10.  Optional arguments
  ! default to sys%ebase = 0, oldXASbroaden off, gaussBroaden off, ucVol = 1.
  subroutine write_core_gpu(fh, niter, kpref, gam0, el, eh, ne, real_a, real_b, imag_a, ierr, &
                            ebase, old_broaden, nval, eps, ucvol, gam_gauss)
    integer, intent(in) :: fh, niter, ne
    real(DP), intent(in) :: kpref, gam0, el, eh
    real(DP), intent(in) :: real_a(0:niter), real_b(0:niter), imag_a(0:niter)
    integer, intent(inout) :: ierr
    real(DP), intent(in), optional :: ebase, nval, eps, ucvol, gam_gauss
    logical, intent(in), optional :: old_broaden
    type(ocean_spect_desc) :: sd
    real(DP), allocatable :: out(:, :)
    integer(c_int32_t) :: nc
    integer :: ie
    sd%gam0 = gam0
    sd%el = el
    sd%eh = eh
    sd%ne = int(ne, c_int32_t)
    sd%ebase_ev = 0.0_DP
    if (present(ebase)) sd%ebase_ev = ebase
    sd%gam_gauss = 0.0_DP
    if (present(gam_gauss)) sd%gam_gauss = gam_gauss
    sd%old_broaden = 0
    if (present(old_broaden)) sd%old_broaden = merge(1, 0, old_broaden)
    sd%nval = 0.0_DP
    if (present(nval)) sd%nval = nval
    sd%eps = 1.0_DP
    if (present(eps)) sd%eps = eps
    sd%ucvol = 1.0_DP
    if (present(ucvol)) sd%ucvol = ucvol
    sd%literal_cmplx = 1
    allocate(out(9, 0:ne))
    ierr = ocean_gpu_write_core(real_a, real_b, imag_a, int(niter, c_int32_t), kpref, sd, out, nc)
    if (ierr /= 0) return
    if (nc == 9) then
      write(fh, '(A,1i5,A,1e15.8,A,1e15.8,A,1e15.8)') '#   iter=', niter, '   gam=', gam0, '   kpref=', kpref, &
            'vol=', sd%ucvol
      write(fh, '(9(A15,1x))') '#   Energy', 'Spect', 'Spect(0)', 'SPKK', 'SPKK(0)', 'sigma (a.u.^2)', &
            'sigma( cm^2 )', 'sig.^2 (a.u.^5)', 'sig.^2 cm^4 sec'
    else
      write(fh, '(A,1i5,A,1e15.8,A,1e15.8)') '#   iter=', niter, '   gam=', gam0, '   kpref=', kpref
      write(fh, '(5(A15,1x))') '#   Energy', 'Spect', 'Spect(0)', 'SPKK', 'SPKK(0)'
    end if
    ! the C side packs rows of nc values contiguously
    call emit(out, nc)
    deallocate(out)
  contains
    subroutine emit(buf, ncol)
      real(DP), intent(in) :: buf(*)
      integer(c_int32_t), intent(in) :: ncol
      integer :: r
      do r = 0, ne
        if (ncol == 9) then
          write(fh, '(9(1e15.8,1x))') buf(9 * r + 1:9 * r + 9)
        else
          write(fh, '(5(1e15.8,1x))') buf(5 * r + 1:5 * r + 5)
        end if
      end do
    end subroutine emit
  end subroutine write_core_gpu

  ! abslanc writer (write_lanczos, OCEAN_haydock.f90:2318-2358); with_eigs appends the
  ! eigenvalue block redtrid passes in (:2273-2316), computed on the device.
  subroutine write_lanczos_gpu(fh, n, kpref, real_a, real_b, with_eigs, ierr)
    integer, intent(in) :: fh, n
    real(DP), intent(in) :: kpref, real_a(0:n), real_b(0:n)
    logical, intent(in), optional :: with_eigs
    integer, intent(inout), optional :: ierr
    real(DP), allocatable :: w(:)
    integer :: i, rc
    write(fh, '(1i8,1x,1ES24.17)') n, kpref
    do i = 0, n
      if (i == 0) then
        write(fh, '(2x,ES24.17)') real_a(i)
      else
        write(fh, '(2(2x,ES24.17))') real_a(i), real_b(i)
      end if
    end do
    if (present(with_eigs)) then
      if (with_eigs) then
        allocate(w(n + 1))
        rc = ocean_gpu_tridiag_eigs(real_a, real_b, int(n, c_int32_t), w)
        if (present(ierr)) ierr = rc
        if (rc == 0) write(fh, '(2x,2i5,1f20.10)') (i, n + 1, w(i), i = 1, n + 1)
        deallocate(w)
      end if
    end if
  end subroutine write_lanczos_gpu

  ! abslanc writer of the complex branch (write_lanczos, OCEAN_haydock.f90:2337-2344):
  ! a(0) and, for i = 1..n, a(i), b(i), c(i) as real and imaginary parts.
  subroutine write_lanczos_abc_gpu(fh, n, kpref, real_a, imag_a, real_b, imag_b, real_c, imag_c)
    integer, intent(in) :: fh, n
    real(DP), intent(in) :: kpref
    real(DP), intent(in), dimension(0:n) :: real_a, imag_a, real_b, imag_b, real_c, imag_c
    integer :: i
    write(fh, '(1i8,1x,1ES24.17)') n, kpref
    write(fh, '(2(2x,ES24.17))') real_a(0), imag_a(0)
    do i = 1, n
      write(fh, '(2x,6ES24.17)') real_a(i), imag_a(i), real_b(i), imag_b(i), real_c(i), imag_c(i)
    end do
  end subroutine write_lanczos_abc_gpu

  subroutine OCEAN_gpu_finalize(ierr)
    integer, intent(inout) :: ierr
    if (c_associated(handle)) ierr = ocean_gpu_destroy(handle)
    handle = c_null_ptr
  end subroutine OCEAN_gpu_finalize

end module ocean_gpu_host
