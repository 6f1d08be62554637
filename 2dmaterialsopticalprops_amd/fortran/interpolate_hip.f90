! interpolate_hip: drop-in for interpolate.x (Chi_RPA/interpolate.f90) on one
! MI355X.
!
! Same &config namelist (interpolate.f90:33-36; pol, Nlf, No_interp, No, omin,
! omax, eta_intra, c0, write_GiGj), same config file rule (the only argument,
! else config.interp.in; interpolate.f90:229-242), same inputs Corrfun_{pol}
! and Qeff_{pol} (the files Optabs calc = 1 writes) and the same outputs:
!   Pi_RPA_{pol}_dense        oi, Re, Im of Pi_tot_interp(1,1)
!   Pi_RPA_{pol}_dense_inter  the same values (interpolate.f90:178, 182)
!   Pi_RPA_{pol}_dense_intra  oi, Re, Im of Qeff(1,1) oi/(oi + i eta_intra)
!   Pi_RPA_{pol}_GiGj_dense   oi, Pi_tot_interp(io,iG,jG) for every point and
!                             pair, when write_GiGj (interpolate.f90:198-209)
! with oi = counter * domega_tot in eV.  The Kramers-Kronig transform, the
! diamagnetic term, the interpolation and the intraband term run on the
! device in one call, rpa_interpolate.
program interpolate_hip
  use iso_c_binding
  use rpa_hip
  use rpa_files, only: dp, hartree
  implicit none

  character(len=100) :: pol = 'xx'
  logical :: write_GiGj = .false.
  integer :: No = 0, Nlf = 0, No_interp = 1
  real(dp) :: omin = 0.0_dp, omax = 0.0_dp, eta_intra = 0.0_dp, c0 = 0.0_dp
  namelist /config/ pol, Nlf, No_interp, No, omin, omax, eta_intra, c0, write_GiGj

  character(len=256) :: config_file
  character(len=100) :: dato1, dato2, dato3, dato4, dato5, dato6
  integer :: u, ios, io, iG, jG, No_tot, ncols
  real(dp) :: domega, domega_tot, oi
  complex(dp), allocatable, target :: S0(:, :, :), Qeff(:, :), Pi_all(:, :, :), Pi11(:), Pi_intra(:)
  type(rpa_interp_params_t) :: p
  integer(c_int) :: rc

  if (command_argument_count() > 1) then
    write(*, '(a)') 'ERROR: Please provide a single argument corresponding to the config_file path.'
    error stop 1
  else if (command_argument_count() == 0) then
    config_file = 'config.interp.in'
  else
    call get_command_argument(1, config_file)
  end if
  write(*, '(a,a)') ' Config file: ', trim(config_file)
  open(newunit=u, file=trim(config_file), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a)') 'ERROR: cannot open '//trim(config_file)
    error stop 1
  end if
  read(u, nml=config, iostat=ios)
  close(u)

  ! unit conversions and grids, interpolate.f90:54-61
  No_tot = (No_interp - 1) * (No - 2)
  eta_intra = eta_intra / hartree
  omin = omin / hartree
  omax = omax / hartree + omin
  domega = (omax - omin) / (No - 1)
  domega_tot = (omax - omin) / No_tot

  dato1 = 'Corrfun_'//adjustl(trim(pol))
  dato2 = 'Qeff_'//adjustl(trim(pol))
  dato6 = 'Pi_RPA_'//adjustl(trim(pol))//'_GiGj_dense'
  dato3 = 'Pi_RPA_'//adjustl(trim(pol))//'_dense'
  dato4 = 'Pi_RPA_'//adjustl(trim(pol))//'_dense_inter'
  dato5 = 'Pi_RPA_'//adjustl(trim(pol))//'_dense_intra'
  write(*, '(a,i0,a,i0,a,i0)') ' No: ', No, ' No_interp: ', No_interp, ' No_tot: ', No_tot

  ! Corrfun_{pol}: per frequency a header line, then Nlf*Nlf lines (jG
  ! fastest); Qeff_{pol}: Nlf*Nlf lines (interpolate.f90:76-101)
  allocate(S0(-No:No, Nlf, Nlf), Qeff(Nlf, Nlf))
  open(newunit=u, file=trim(dato1), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a)') 'ERROR: cannot open '//trim(dato1)
    error stop 1
  end if
  do io = -No, No
    read(u, *)
    do iG = 1, Nlf
      do jG = 1, Nlf
        read(u, '(2F15.10)') S0(io, iG, jG)
      end do
    end do
  end do
  close(u)
  open(newunit=u, file=trim(dato2), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a)') 'ERROR: cannot open '//trim(dato2)
    error stop 1
  end if
  do iG = 1, Nlf
    do jG = 1, Nlf
      read(u, '(2F15.10)') Qeff(iG, jG)
    end do
  end do
  close(u)
  write(*, '(a)') ' COMPLETED: Read S0 and Qeff matrices.'

  ! the dense response: every pair when write_GiGj, else (1,1) only
  ncols = merge(Nlf * Nlf, 1, write_GiGj)
  allocate(Pi_intra(max(No_tot, 1)))
  if (write_GiGj) then
    allocate(Pi_all(max(No_tot, 1), Nlf, Nlf))
  else
    allocate(Pi11(max(No_tot, 1)))
  end if
  p%no = No
  p%nlf = Nlf
  p%no_interp = No_interp
  p%all = merge(1, 0, write_GiGj)
  p%device = 0
  p%domega = domega
  p%domega_tot = domega_tot
  p%eta_intra = eta_intra
  if (write_GiGj) then
    rc = rpa_interpolate(p, c_loc(S0), c_loc(Qeff), c_loc(Pi_all), c_loc(Pi_intra))
  else
    rc = rpa_interpolate(p, c_loc(S0), c_loc(Qeff), c_loc(Pi11), c_loc(Pi_intra))
  end if
  if (rc /= RPA_OK) then
    write(*, '(a)') 'ERROR: rpa_interpolate: '//rpa_last_error(c_null_ptr)
    error stop 1
  end if
  write(*, '(a,i0,a,i0)') ' COMPLETED: current-current response and interpolation, No_tot: ', No_tot, &
                          ', pairs: ', ncols

  call write_dense(dato3, .false.)
  call write_dense(dato4, .false.)
  call write_dense(dato5, .true.)
  if (write_GiGj) then
    open(newunit=u, file=trim(dato6))
    do io = 1, No_tot
      oi = io * domega_tot
      do iG = 1, Nlf
        do jG = 1, Nlf
          write(u, *) oi * hartree, Pi_all(io, iG, jG)
        end do
      end do
    end do
    close(u)
  end if
  write(*, '(a,a)') ' COMPLETED: Writting interpoalted current-current response to file. Polarization: ', &
                    trim(pol)

contains

  ! one (1,1) column per line: oi in eV, Re, Im (interpolate.f90:183-185)
  subroutine write_dense(name, intra)
    character(len=*), intent(in) :: name
    logical, intent(in) :: intra
    integer :: w, c
    complex(dp) :: v
    open(newunit=w, file=trim(name))
    do c = 1, No_tot
      if (intra) then
        v = Pi_intra(c)
      else if (write_GiGj) then
        v = Pi_all(c, 1, 1)
      else
        v = Pi11(c)
      end if
      write(w, *) c * domega_tot * hartree, dble(v), aimag(v)
    end do
    close(w)
  end subroutine write_dense

end program interpolate_hip
