! Fortran host mirror of OCEAN2's hot-path call surface over the HIP C-ABI.
!
!   OCEAN_xact_gpu            <- OCEAN_xact(sys, inter_scale, psi, new_psi, ierr, backwards, nhflag)
!                                OCEAN2/OCEAN_action.f90:286
!   OCEAN_haydock_do_gpu      <- OCEAN_haydock_do / OCEAN_haydock_Herm_do
!                                OCEAN2/OCEAN_haydock.f90:167, :190-320
!   OCEAN_gmres_do_recycle_gpu <- OCEAN_gmres_do_recycle, OCEAN2/OCEAN_gmres.f90:508
!   write_core_gpu      <- write_core, OCEAN2/OCEAN_haydock.f90:1801-1894
!   write_lanczos_gpu   <- write_lanczos, OCEAN2/OCEAN_haydock.f90:2318-2358
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
            OCEAN_gmres_do_recycle_gpu, write_core_gpu, write_lanczos_gpu, &
            OCEAN_gpu_finalize, OCEAN_precision_gpu

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
  subroutine OCEAN_haydock_do_gpu(psi_r, psi_i, niter, real_a, real_b, imag_a, niter_done, ierr)
    real(DP), intent(in) :: psi_r(*), psi_i(*)
    integer, intent(in) :: niter
    real(DP), intent(out) :: real_a(0:niter), real_b(0:niter), imag_a(0:niter)
    integer, intent(out) :: niter_done
    integer, intent(inout) :: ierr
    integer(c_int32_t) :: done
    ierr = ocean_gpu_haydock(handle, psi_r, psi_i, int(niter, c_int32_t), real_a, real_b, imag_a, done)
    niter_done = done
  end subroutine OCEAN_haydock_do_gpu

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

  ! absspct writer in write_core's layout (Lorentzian, gaussian broadening off).
  subroutine write_core_gpu(fh, niter, kpref, gam0, el, eh, ne, real_a, real_b, imag_a, ierr)
    integer, intent(in) :: fh, niter, ne
    real(DP), intent(in) :: kpref, gam0, el, eh
    real(DP), intent(in) :: real_a(0:niter), real_b(0:niter), imag_a(0:niter)
    integer, intent(inout) :: ierr
    real(DP), allocatable :: out(:, :)
    integer :: ie
    allocate(out(5, 0:ne))
    ierr = ocean_gpu_spectrum(real_a, real_b, imag_a, int(niter, c_int32_t), kpref, gam0, el, eh, &
                              int(ne, c_int32_t), 1_c_int32_t, out)
    if (ierr /= 0) return
    write(fh, '(A,1i5,A,1e15.8,A,1e15.8)') '#   iter=', niter, '   gam=', gam0, '   kpref=', kpref
    write(fh, '(5(A15,1x))') '#   Energy', 'Spect', 'Spect(0)', 'SPKK', 'SPKK(0)'
    do ie = 0, ne
      write(fh, '(5(1e15.8,1x))') out(1, ie), out(2, ie), out(3, ie), out(4, ie), out(5, ie)
    end do
    deallocate(out)
  end subroutine write_core_gpu

  ! abslanc writer (write_lanczos without the eigenvalue block).
  subroutine write_lanczos_gpu(fh, n, kpref, real_a, real_b)
    integer, intent(in) :: fh, n
    real(DP), intent(in) :: kpref, real_a(0:n), real_b(0:n)
    integer :: i
    write(fh, '(1i8,1x,1ES24.17)') n, kpref
    do i = 0, n
      if (i == 0) then
        write(fh, '(2x,ES24.17)') real_a(i)
      else
        write(fh, '(2(2x,ES24.17))') real_a(i), real_b(i)
      end if
    end do
  end subroutine write_lanczos_gpu

  subroutine OCEAN_gpu_finalize(ierr)
    integer, intent(inout) :: ierr
    if (c_associated(handle)) ierr = ocean_gpu_destroy(handle)
    handle = c_null_ptr
  end subroutine OCEAN_gpu_finalize

end module ocean_gpu_host
