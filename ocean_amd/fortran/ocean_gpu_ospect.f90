! o_spect.x over the HIP C-ABI (OCEAN2/o_spect.f90:26-284): run from an OCEAN run
! directory after the Lanczos files exist.  Reads spect.in (ne el eh gam0 ebase, eV),
! epsilon, nval.h, avecsinbohr.ipt, runlist and, when present, do_avg.ipt with the
! photonN weights and niter_override; evaluates each run's continued fraction on the
! device (ocean_gpu_ospect) and writes the reference's files: absspct_/xesspct_/
! rxsspct_/opcons, and TotAvg when averaging.
program ocean_gpu_ospect_x
  use iso_c_binding
  use ocean_gpu_iface
  implicit none
  integer, parameter :: DP = kind(1.0d0)
  real(DP), parameter :: eV2Hartree = 0.036749322175666_DP   ! OCEAN2/OCEAN_constants.f90:15
  integer :: ne, nruns, run_iter, znl(3), indx, photon, rixs_energy, rixs_pol, n_recur, it, i, ii, ios
  character(len=2) :: elname, corelevel
  character(len=5) :: calc_type
  character(len=24) :: lancfile, photfile
  character(len=25) :: abs_filename
  character(len=256) :: line
  character(len=80) :: word
  real(DP) :: el, eh, gam0, ebase, eps0, nval, avec(3, 3), ucvol, kpref, avgweight
  real(DP), allocatable :: a(:), b(:), out(:), total(:, :)
  logical :: is_val, ex, do_avg
  type(ocean_ospect_desc) :: od
  integer(c_int32_t) :: nr, nc
  integer(c_int) :: rc

  open(unit=99, file='spect.in', form='formatted', status='old')
  read(99, *) ne, el, eh, gam0, ebase
  close(99)
  open(unit=99, file='epsilon', form='formatted', status='old')
  read(99, *) eps0
  close(99)
  open(unit=99, file='nval.h', form='formatted', status='old')
  read(99, *) nval
  close(99)
  open(unit=99, file='avecsinbohr.ipt', form='formatted', status='old')
  read(99, *) avec(:, :)
  close(99)
  ucvol = abs(avec(1, 1) * (avec(2, 2) * avec(3, 3) - avec(3, 2) * avec(2, 3)) &
            - avec(1, 2) * (avec(2, 1) * avec(3, 3) - avec(3, 1) * avec(2, 3)) &
            + avec(1, 3) * (avec(2, 1) * avec(3, 2) - avec(3, 1) * avec(2, 2)))

  inquire(file='do_avg.ipt', exist=ex)
  do_avg = .false.
  if (ex) then
    open(unit=99, file='do_avg.ipt', form='formatted', status='old')
    read(99, *) do_avg
    close(99)
  end if
  if (do_avg) then
    ! ne + 1 rows are accumulated; TotAvg holds the first ne (o_spect.f90:123-124, :275-277)
    allocate(total(5, ne + 1))
    total(:, :) = 0.0_DP
  end if

  od%ne = int(ne, c_int32_t)
  od%el_ev = el
  od%eh_ev = eh
  od%gam0_ev = gam0
  od%ebase_ev = ebase
  od%eps0 = eps0
  od%nval = nval
  od%ucvol = ucvol
  allocate(out(8 * (ne + 1)))

  open(unit=98, file='runlist', form='formatted', status='old')
  read(98, *) nruns
  do run_iter = 1, nruns
    read(98, '(A)') line
    read(line, *) znl(1), znl(2), znl(3), elname, corelevel, indx, photon, calc_type
    is_val = .false.
    select case (calc_type)
    case ('XES')
      write(abs_filename, '(A8,A2,A1,I4.4,A1,A2,A1,I2.2)') 'xesspct_', elname, '.', indx, '_', corelevel, '_', photon
      write(lancfile, '(A8,A2,A1,I4.4,A1,A2,A1,I2.2)') 'xeslanc_', elname, '.', indx, '_', corelevel, '_', photon
    case ('RXS')
      read(line, *) znl(1), znl(2), znl(3), elname, corelevel, indx, photon, calc_type, rixs_energy, rixs_pol
      is_val = .true.
      write(abs_filename, '(A8,A2,A1,A2,A1,I2.2,A1,I5.5,A1,I2.2)') 'rxsspct_', elname, '.', corelevel, '_', &
            photon, '.', rixs_energy, '.', rixs_pol
      write(lancfile, '(A7,A2,A1,A2,A1,I2.2,A1,I5.5,A1,I2.2)') 'rxlanc_', elname, '.', corelevel, '_', &
            photon, '.', rixs_energy, '.', rixs_pol
    case ('VAL')
      abs_filename = 'opcons'
      lancfile = 'oplanc'
      is_val = .true.
    case default
      write(abs_filename, '(A8,A2,A1,I4.4,A1,A2,A1,I2.2)') 'absspct_', elname, '.', indx, '_', corelevel, '_', photon
      write(lancfile, '(A8,A2,A1,I4.4,A1,A2,A1,I2.2)') 'abslanc_', elname, '.', indx, '_', corelevel, '_', photon
    end select

    if (do_avg) then
      ! photonN: a header line, two vector blocks closed by 'end', the energy line, the weight
      write(photfile, '(A,I0)') 'photon', photon
      open(unit=99, file=trim(photfile), form='formatted', status='old')
      read(99, *)
      do i = 1, 2
        do
          read(99, *, iostat=ios) word
          if (ios /= 0 .or. word == 'end') exit
        end do
      end do
      read(99, *)
      read(99, *) avgweight
      close(99)
    end if

    open(unit=99, file=trim(lancfile), form='formatted', status='old')
    read(99, *) n_recur, kpref
    allocate(a(0:n_recur), b(0:n_recur))
    b(0) = 0.0_DP
    read(99, *) a(0)
    do i = 1, n_recur
      read(99, *) a(i), b(i)
    end do
    close(99)

    it = n_recur
    inquire(file='niter_override', exist=ex)
    if (ex) then
      open(unit=99, file='niter_override', form='formatted', status='old')
      read(99, *) it
      close(99)
      it = min(max(1, it), n_recur)
    end if

    od%is_val = merge(1, 0, is_val)
    rc = ocean_gpu_ospect(a, b, int(it, c_int32_t), kpref, od, out, nr, nc)
    if (rc /= 0) then
      write(6, *) 'ocean_gpu_ospect failed: ', rc
      stop 1
    end if

    open(unit=99, file=trim(abs_filename), form='formatted', status='unknown')
    if (is_val) then
      write(99, "(a)") "#   omega (eV)      epsilon_1       epsilon_2       n"// &
        "               kappa           mu (cm^(-1))    R"// &
        "               epsinv"
      do ii = 0, nr - 1
        write(99, '(8(1E24.16,1X))') out(8 * ii + 1:8 * ii + 8)
      end do
    else
      write(99, '(A,1i5,A,1e15.8,A,1e15.8)') '#   iter=', it, '   gam=', gam0 * eV2Hartree, &
            '   kpref=', kpref
      write(99, '(5(A15,1x))') '#   Energy', 'Spect', 'Spect(0)', 'SPKK', 'SPKK(0)'
      do ii = 0, nr - 1
        write(99, '(5(1e15.8,1x))') out(5 * ii + 1:5 * ii + 5)
        if (do_avg) then
          total(1, ii + 1) = out(5 * ii + 1)
          total(2, ii + 1) = total(2, ii + 1) + out(5 * ii + 5) * avgweight
          total(3, ii + 1) = total(3, ii + 1) + out(5 * ii + 3) * avgweight
          total(4, ii + 1) = total(4, ii + 1) + out(5 * ii + 4) * avgweight
          total(5, ii + 1) = total(5, ii + 1) + out(5 * ii + 2) * avgweight
        end if
      end do
    end if
    close(99)
    deallocate(a, b)
  end do
  close(98)

  if (do_avg) then
    open(unit=99, file='TotAvg', form='formatted')
    do ii = 1, ne
      write(99, '(5(E24.16,1X))') total(:, ii)
    end do
    close(99)
  end if
end program ocean_gpu_ospect_x
