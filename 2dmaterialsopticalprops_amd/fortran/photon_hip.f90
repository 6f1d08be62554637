! photon_hip: drop-in for Photon.x (BSE/Photonew.f90) on one MI355X.
!
! Same config file (the only argument, else config.in; Photonew.f90:372-386)
! with the same namelists &directories, &system, &config, &parameters and
! &parallel (Photonew.f90:26-90), the same inputs (scf.out for the reciprocal
! axes, charge-density.dat for the G list, the unscreened current-current
! tensors rpa_xx_file / rpa_yy_file / rpa_zz_file as loadPi0 reads them,
! 665-696) and the same output files, appended per frequency as the
! reference's writePi / writeSigma / writeSpectra do (1066-1171):
!   pi{xx,yy,zz}0_q#{iq}_theta#{it}, pi{xx,yy,zz}_q#.., sigma0_{xx,yy,zz}_q#..,
!   sigmaSc_{xx,yy,zz}_q#.., pi{xx,yy,zz}_sumGiGj_q#..,
!   sigmaSc_sumGiGj_{xx,yy,zz}_q#.., absorption / transmission / reflection{id}.
! Every (angle, q, frequency) point of the reference's loops (205-330) runs on
! the device in one call, rpa_photon.
program photon_hip
  use iso_c_binding
  use rpa_hip
  use rpa_files, only: dp, hartree, load_kc, load_g
  implicit none

  real(dp), parameter :: pi = 4.0_dp * atan(1.0_dp)
  real(dp), parameter :: gamma = 1.0_dp / 137.0_dp

  integer :: No = 0, NG = 0, NGd = 0, NkI = 0, Nlfd = 0
  character(len=200) :: rundir = './', savedir = './', scf_file = 'scf.out', &
                        rpa_xx_file = 'Pi_RPA_xx_GiGj_dense', rpa_yy_file = 'Pi_RPA_yy_GiGj_dense', &
                        rpa_zz_file = 'Pi_RPA_zz_GiGj_dense'
  character(len=3) :: lf = 'z'
  integer :: qmin = 1, qmax = 1, theta_min = 1, theta_max = 1
  real(dp) :: dq = 0.0_dp, omin = 0.0_dp, omax = 0.0_dp, eta = 0.0_dp, Ecut = 0.0_dp
  real(dp) :: a0 = 1.0_dp, c0 = 1.0_dp, Vcell = 0.0_dp
  integer :: Nthreads = 1, NthreadsMKL = 1
  namelist /directories/ rundir, savedir, scf_file, rpa_xx_file, rpa_yy_file, rpa_zz_file
  namelist /system/ lf, qmin, qmax, theta_min, theta_max, dq, omin, omax, eta, Ecut
  namelist /config/ No, NG, NGd, NkI, Nlfd
  namelist /parameters/ a0, c0, Vcell
  namelist /parallel/ Nthreads, NthreadsMKL

  character(len=256) :: config_file
  integer :: u, ios, ngm, nlf, Ntheta, Nq, itheta, iq, io, nline, l, iG
  real(dp) :: gcar, domega, dtheta, kc(3, 3), o
  integer, allocatable :: mill(:, :), parg(:), lf_idx(:), qmode(:)
  real(dp), allocatable :: g(:, :), glf(:, :), glfz(:), qline(:), thline(:)
  complex(dp), allocatable, target :: Pixx0(:, :, :), Piyy0(:, :, :), Pizz0(:, :, :)
  complex(dp), allocatable, target :: pi11(:, :, :), psum(:, :, :)
  real(dp), allocatable, target :: spec(:, :, :)
  integer, allocatable :: line_iq(:), line_it(:)
  type(rpa_photon_params_t) :: p
  integer(c_int) :: rc

  write(*, '(a)') ' PROGRAM PHOTON RUN STARTED'
  if (command_argument_count() > 1) then
    write(*, '(a)') 'ERROR: Please provide a single argument corresponding to the config_file path.'
    error stop 1
  else if (command_argument_count() == 0) then
    config_file = 'config.in'
  else
    call get_command_argument(1, config_file)
  end if
  open(newunit=u, file=trim(config_file), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a)') 'ERROR: cannot open '//trim(config_file)
    error stop 1
  end if
  read(u, nml=directories, iostat=ios); rewind(u)
  read(u, nml=system, iostat=ios); rewind(u)
  read(u, nml=config, iostat=ios); rewind(u)
  read(u, nml=parameters, iostat=ios); rewind(u)
  read(u, nml=parallel, iostat=ios)
  close(u)

  ! constants and grids, Photonew.f90:98-106, 155-161
  gcar = 2.0_dp * pi / a0
  eta = eta / hartree
  omin = omin / hartree
  omax = omax / hartree + omin
  Ntheta = theta_max - theta_min
  Nq = qmax - qmin
  domega = (omax - omin) / (No - 1)
  if (Ntheta /= 0) then
    dtheta = pi / (2.0_dp * dble(Ntheta))
  else
    dtheta = 1.0_dp
  end if
  if (Ntheta /= 0 .and. Nq /= 0) then  ! Photonew.f90:240-242
    write(*, '(a)') 'FATAL ERROR: Both Ntheta>1 and Nq>1. Choose one or the other.'
    error stop 1
  end if

  ! reciprocal axes, G list, local fields with parity (163-186)
  call load_kc(trim(rundir)//'/'//trim(scf_file), kc)
  call load_g(trim(savedir), ngm, mill)
  allocate(g(3, ngm), glf(3, max(Nlfd, 1)), parg(max(Nlfd, 1)), lf_idx(max(Nlfd, 1)))
  call rpa_check(rpa_gcart(ngm, mill, kc, g), c_null_ptr, 'G vectors')
  call rpa_check(rpa_gen_glf(trim(lf)//c_null_char, Ecut, gcar, ngm, g, mill, Nlfd, nlf, glf, parg, &
                             lf_idx), c_null_ptr, 'genGlfandParity')
  write(*, '(a,i0,a,i0)') ' Nlf: ', nlf, ' Nlfd: ', Nlfd
  allocate(glfz(nlf))
  do iG = 1, nlf
    glfz(iG) = gcar * glf(3, iG)
  end do

  call load_pi0(trim(rpa_xx_file), Pixx0)
  call load_pi0(trim(rpa_yy_file), Piyy0)
  call load_pi0(trim(rpa_zz_file), Pizz0)
  write(*, '(a)') ' status: Current-current response tensor Pi0(omega,G,G'') loaded.'

  ! the spectrum lines in the reference's loop order (angle outer, q inner)
  ! and the q of each: (iq-1) dq, or gamma o sin(theta) for angles past the
  ! first when there is one q (Photonew.f90:205-249)
  nline = (theta_max - theta_min + 1) * (qmax - qmin + 1)
  allocate(qline(nline), thline(nline), qmode(nline), line_iq(nline), line_it(nline))
  l = 0
  do itheta = theta_min, theta_max
    do iq = qmin, qmax
      l = l + 1
      line_iq(l) = iq
      line_it(l) = itheta
      thline(l) = (itheta - 1) * dtheta
      qline(l) = (iq - 1) * dq
      qmode(l) = merge(1, 0, itheta /= 1 .and. Nq == 0)
    end do
  end do

  allocate(pi11(3, No - 1, nline), psum(3, No - 1, nline), spec(2, No - 1, nline))
  p%no = No
  p%nlf = nlf
  p%nline = nline
  p%device = 0
  p%omin = omin
  p%domega = domega
  p%eta = eta
  p%c0 = c0
  rc = rpa_photon(p, glfz, parg, c_loc(Pixx0), c_loc(Piyy0), c_loc(Pizz0), c_null_ptr, c_null_ptr, &
                  qline, thline, qmode, c_loc(pi11), c_loc(psum), c_loc(spec), c_null_ptr)
  if (rc /= RPA_OK) then
    write(*, '(a)') 'ERROR: rpa_photon: '//rpa_last_error(c_null_ptr)
    error stop 1
  end if

  do l = 1, nline
    do io = 1, No - 1
      o = omin + (io - 1) * domega
      call write_pi(o, Pixx0(io, 1, 1), 'pixx0', l)
      call write_pi(o, Piyy0(io, 1, 1), 'piyy0', l)
      call write_pi(o, Pizz0(io, 1, 1), 'pizz0', l)
      call write_pi(o, pi11(1, io, l), 'pixx', l)
      call write_pi(o, pi11(2, io, l), 'piyy', l)
      call write_pi(o, pi11(3, io, l), 'pizz', l)
      call write_sigma(o, Pixx0(io, 1, 1), 'sigma0_xx', l)
      call write_sigma(o, Piyy0(io, 1, 1), 'sigma0_yy', l)
      call write_sigma(o, Pizz0(io, 1, 1), 'sigma0_zz', l)
      call write_sigma(o, pi11(1, io, l), 'sigmaSc_xx', l)
      call write_sigma(o, pi11(2, io, l), 'sigmaSc_yy', l)
      call write_sigma(o, pi11(3, io, l), 'sigmaSc_zz', l)
      call write_pi(o, psum(1, io, l), 'pixx_sumGiGj', l)
      call write_pi(o, psum(2, io, l), 'piyy_sumGiGj', l)
      call write_pi(o, psum(3, io, l), 'pizz_sumGiGj', l)
      call write_sigma(o, psum(1, io, l), 'sigmaSc_sumGiGj_xx', l)
      call write_sigma(o, psum(2, io, l), 'sigmaSc_sumGiGj_yy', l)
      call write_sigma(o, psum(3, io, l), 'sigmaSc_sumGiGj_zz', l)
      call write_spectra(l, o, spec(1, io, l), spec(2, io, l))
    end do
  end do
  write(*, '(a)') ' PROGRAM PHOTON RUN ENDED'

contains

  ! int2str of Modules/utility.f90:24-29
  function int2str(k) result(s)
    integer, intent(in) :: k
    character(len=20) :: s
    write(s, *) k
    s = adjustl(s)
  end function int2str

  function line_name(stem, l) result(name)
    character(len=*), intent(in) :: stem
    integer, intent(in) :: l
    character(len=:), allocatable :: name
    name = trim(stem//'_q#'//int2str(line_iq(l)))//trim('_theta#'//int2str(line_it(l)))
  end function line_name

  subroutine append(name, u)
    character(len=*), intent(in) :: name
    integer, intent(out) :: u
    open(newunit=u, file=trim(adjustl(name)), position='append', action='write')
  end subroutine append

  ! writePi (1121-1144): o in eV, Re, Im
  subroutine write_pi(o, v, stem, l)
    real(dp), intent(in) :: o
    complex(dp), intent(in) :: v
    character(len=*), intent(in) :: stem
    integer, intent(in) :: l
    integer :: w
    call append(line_name(stem, l), w)
    write(w, *) o * hartree, real(v), aimag(v)
    close(w)
  end subroutine write_pi

  ! writeSigma (1146-1171): o in eV, -i 4 c0 Pi / o
  subroutine write_sigma(o, v, stem, l)
    real(dp), intent(in) :: o
    complex(dp), intent(in) :: v
    character(len=*), intent(in) :: stem
    integer, intent(in) :: l
    integer :: w
    call append(line_name(stem, l), w)
    write(w, *) o * hartree, real(-dcmplx(0.0_dp, 1.0_dp) * 4.0_dp * c0 * v / o), &
                aimag(-dcmplx(0.0_dp, 1.0_dp) * 4.0_dp * c0 * v / o)
    close(w)
  end subroutine write_sigma

  ! writeSpectra (1066-1119): absorption, transmission = 1 - A - R, reflection
  subroutine write_spectra(l, o, a_p, r_p)
    integer, intent(in) :: l
    real(dp), intent(in) :: o, a_p, r_p
    character(len=20) :: id
    integer :: w
    if (Ntheta /= 0 .and. Nq == 0) then
      id = '_theta=#'//int2str(line_it(l))
    else if (Ntheta == 0 .and. Nq /= 0) then
      id = '_q#'//int2str(line_iq(l))
    else
      id = ''
    end if
    call append('absorption'//trim(id), w)
    write(w, *) o * hartree, a_p
    close(w)
    call append('transmission'//trim(id), w)
    write(w, *) o * hartree, 1 - a_p - r_p
    close(w)
    call append('reflection'//trim(id), w)
    write(w, *) o * hartree, r_p
    close(w)
  end subroutine write_spectra

  ! loadPi0 (665-696): per frequency and (iG, jG) one list-directed record
  ! 'o, Pi'
  subroutine load_pi0(path, P0)
    character(len=*), intent(in) :: path
    complex(dp), allocatable, target, intent(out) :: P0(:, :, :)
    integer :: w, i, j, f
    real(dp) :: od
    allocate(P0(No, nlf, nlf))
    open(newunit=w, file=adjustl(trim(path)), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, '(a)') 'ERROR: cannot open '//trim(path)
      error stop 1
    end if
    do f = 1, No
      do i = 1, nlf
        do j = 1, nlf
          read(w, *) od, P0(f, i, j)
        end do
      end do
    end do
    close(w)
  end subroutine load_pi0

end program photon_hip
