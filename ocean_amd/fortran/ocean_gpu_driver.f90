! Fortran host driver: reads one core-level problem (module state + starting
! vector), then runs OCEAN_xact_gpu, the Haydock recursion with its absspct /
! abslanc output, and optionally GMRES, all through the HIP C-ABI.
!
! usage: ocean_gpu_driver.x <input> <output-prefix> [<spacing> <threshold> <ne> | abc]
!   (the optional triple runs the recursion with the reference's early exit; abc runs
!   the non-Hermitian bi-Lanczos recursion of the complex_haydock branch instead and
!   writes its six-column abslanc)
! input (stream, little endian):
!   int32  nb k1 k2 k3 x1 x2 x3 nspn lc lmin lmax lvl lvh nproj(8) have_so niter ne nen depth precision
!   real8  celvol inter_scale kpref gam el eh gres gprc ffff
!   real8  u_re u_im (nb,nk,nx,nspn) W (nk,nx) e_re e_im allow (nb,nk,na)
!          mpcr mpci (nb,nk,npmax,2lmax+1,nl,nspn) mpm (npmax,npmax,nl) mhr mhi (sum mham^2)
!          somelr someli (na,na) psi_re psi_im (nb,nk,na) energies (nen)
! output <prefix>.bin (stream): int32 ierr niter_done; real8 hpsi_re hpsi_im a(0:niter) b(0:niter)
!   [abc: imag_a imag_b real_c imag_c (0:niter)] gmres(6,nen)
program ocean_gpu_driver
  use iso_c_binding
  use ocean_gpu_iface
  use ocean_gpu_host
  implicit none
  integer, parameter :: DP = kind(1.0d0)
  character(len=512) :: fin, prefix, arg
  integer :: conv_spacing, conv_ne
  real(DP) :: conv_thresh
  logical :: early, abc
  integer :: hdr(27), nb, nk, nx, nspn, lc, lmin, lmax, lvl, lvh, npmax, nl, na, niter, ne, nen, depth
  integer :: ierr, done, i, l, mhsum, have_so, prec
  real(DP) :: sc(9)
  real(DP), allocatable :: u_re(:), u_im(:), w(:), e_re(:), e_im(:), allow(:), mpcr(:), mpci(:), mpm(:)
  real(DP), allocatable :: mhr(:), mhi(:), sr(:), si(:), p_re(:), p_im(:), h_re(:), h_im(:)
  real(DP), allocatable :: ra(:), rb(:), ia(:), energies(:), rows(:, :), ib(:), rc(:), ic(:)
  type(ocean_bse_desc) :: d

  call get_command_argument(1, fin)
  call get_command_argument(2, prefix)
  early = command_argument_count() >= 5
  abc = .false.
  if (command_argument_count() == 3) then
    call get_command_argument(3, arg)
    abc = trim(arg) == 'abc'
  end if
  if (early) then
    call get_command_argument(3, arg)
    read(arg, *) conv_spacing
    call get_command_argument(4, arg)
    read(arg, *) conv_thresh
    call get_command_argument(5, arg)
    read(arg, *) conv_ne
  end if
  open(unit=11, file=trim(fin), form='unformatted', access='stream', status='old')
  read(11) hdr
  nb = hdr(1); nk = hdr(2) * hdr(3) * hdr(4); nx = hdr(5) * hdr(6) * hdr(7)
  nspn = hdr(8); lc = hdr(9); lmin = hdr(10); lmax = hdr(11); lvl = hdr(12); lvh = hdr(13)
  have_so = hdr(22); niter = hdr(23); ne = hdr(24); nen = hdr(25); depth = hdr(26); prec = hdr(27)
  na = 4 * (2 * lc + 1)
  nl = lmax - lmin + 1
  npmax = maxval(hdr(14:14 + nl - 1))
  mhsum = 0
  do l = lvl, lvh
    mhsum = mhsum + (na * (2 * l + 1) * hdr(14 + l - lmin)) ** 2
  end do
  read(11) sc

  d%num_bands = nb
  d%kmesh = hdr(2:4)
  d%xmesh = hdr(5:7)
  d%nspn = nspn
  d%lc = lc
  d%lmin = lmin
  d%lmax = lmax
  d%nproj = hdr(14:21)
  d%lvl = lvl
  d%lvh = lvh
  d%celvol = sc(1)
  d%inter_scale = sc(2)
  d%e0 = 1
  d%mult = 1
  d%long_range = 1
  d%precision = prec

  allocate(u_re(nb * nk * nx * nspn), u_im(nb * nk * nx * nspn), w(nk * nx))
  allocate(e_re(nb * nk * na), e_im(nb * nk * na), allow(nb * nk * na))
  allocate(mpcr(nb * nk * npmax * (2 * lmax + 1) * nl * nspn), mpci(nb * nk * npmax * (2 * lmax + 1) * nl * nspn))
  allocate(mpm(npmax * npmax * nl), mhr(mhsum), mhi(mhsum), sr(na * na), si(na * na))
  allocate(p_re(nb * nk * na), p_im(nb * nk * na), h_re(nb * nk * na), h_im(nb * nk * na))
  allocate(ra(0:niter), rb(0:niter), ia(0:niter), energies(max(nen, 1)), rows(6, max(nen, 1)))
  allocate(ib(0:niter), rc(0:niter), ic(0:niter))
  read(11) u_re, u_im, w, e_re, e_im, allow, mpcr, mpci, mpm, mhr, mhi, sr, si, p_re, p_im
  if (nen > 0) read(11) energies(1:nen)
  close(11)

  ierr = 0
  done = 0
  rows = 0.0_DP
  h_re = 0.0_DP
  h_im = 0.0_DP
  ra = 0.0_DP
  rb = 0.0_DP
  ia = 0.0_DP
  ib = 0.0_DP
  rc = 0.0_DP
  ic = 0.0_DP
  call OCEAN_gpu_init(d, ierr)
  if (ierr == 0) call OCEAN_gpu_load(u_re, u_im, w, mpcr, mpci, mpm, mhr, mhi, sr, si, have_so /= 0, &
                                     e_re, e_im, allow, ierr)
  if (ierr == 0) call OCEAN_xact_gpu(p_re, p_im, h_re, h_im, ierr)
  if (ierr == 0 .and. niter > 0 .and. abc) then
    call OCEAN_haydock_nonHerm_do_gpu(p_re, p_im, niter, ra, ia, rb, ib, rc, ic, done, ierr)
    if (ierr == 0) then
      open(unit=13, file=trim(prefix)//'.abslanc', form='formatted', status='unknown')
      call write_lanczos_abc_gpu(13, done, sc(3), ra(0:done), ia(0:done), rb(0:done), ib(0:done), &
                                 rc(0:done), ic(0:done))
      close(13)
    end if
  else if (ierr == 0 .and. niter > 0) then
    if (early) then
      call OCEAN_haydock_do_gpu(p_re, p_im, niter, ra, rb, ia, done, ierr, conv_spacing, conv_thresh, &
                                sc(3), sc(4), sc(5), sc(6), conv_ne)
    else
      call OCEAN_haydock_do_gpu(p_re, p_im, niter, ra, rb, ia, done, ierr)
    end if
    if (ierr == 0) then
      open(unit=12, file=trim(prefix)//'.absspct', form='formatted', status='unknown')
      ! the files describe the iterations actually run (haydump / redtrid take that count)
      call write_core_gpu(12, done, sc(3), sc(4), sc(5), sc(6), ne, ra(0:done), rb(0:done), ia(0:done), ierr, &
                          ucvol=sc(1))
      close(12)
      open(unit=13, file=trim(prefix)//'.abslanc', form='formatted', status='unknown')
      if (ierr == 0) call write_lanczos_gpu(13, done, sc(3), ra(0:done), rb(0:done), .true., ierr)
      close(13)
    end if
  end if
  if (ierr == 0 .and. nen > 0) then
    open(unit=14, file=trim(prefix)//'.gmres', form='formatted', status='unknown')
    call OCEAN_gmres_do_recycle_gpu(p_re, p_im, energies(1:nen), sc(7), sc(8), sc(9), depth, sc(3), &
                                    rows(:, 1:nen), 14, ierr)
    close(14)
  end if
  i = 0
  call OCEAN_gpu_finalize(i)

  open(unit=15, file=trim(prefix)//'.bin', form='unformatted', access='stream', status='replace')
  write(15) int(ierr, c_int32_t), int(done, c_int32_t)
  write(15) h_re, h_im, ra, rb
  if (abc) write(15) ia, ib, rc, ic
  if (nen > 0) write(15) rows(:, 1:nen)
  close(15)
  if (ierr /= 0) write(6, *) '!!', ierr
end program ocean_gpu_driver
