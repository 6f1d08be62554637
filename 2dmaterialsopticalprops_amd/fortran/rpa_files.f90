! Input readers and output writers of the Optabs / Sloss drivers.
!
! The formats are the reference's:
!   scf_file   PointR.f90:41-91 (Nsym, symmetry matrices), loadKC
!              (Sloss.f90:1279-1307, Optabs.f90:652-680)
!   band_file  loadkIandE (Sloss.f90:1492-1523, Optabs.f90:1320-1352)
!   savedir    QE >= 6 charge-density.dat and wfc{ik}.dat records
!              (Modules/io_qe.f90:11-166)
! Output files are written by the drivers with the reference's formats.
module rpa_files
  use iso_fortran_env, only: dp => real64
  implicit none
  private
  public :: dp, hartree, point_r, load_kc, load_ki_and_e, load_g, wfc_size, load_cs_full, &
            check_fbz_integration, numbered

  real(dp), parameter :: hartree = 2.0_dp * 13.6056923_dp  ! Sloss.f90:71

contains

  ! Number of symmetry operations and their cartesian matrices R(48,3,3):
  ! Nsym is read list-directed three lines after the line that starts with
  ! '     atomic'; each matrix is the three '(19X,3F11.3)' rows that follow
  ! three lines after a line starting with ' cryst.' (PointR.f90:41-91).
  subroutine point_r(path, nsym, R)
    character(len=*), intent(in) :: path
    integer, intent(out) :: nsym
    real(dp), intent(out) :: R(48, 3, 3)
    character(len=200) :: line
    integer :: u, ios, i, is
    nsym = 0
    R = 0.0_dp
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('cannot open scf file '//trim(path))
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) call die('no "     atomic" line in '//trim(path))
      if (line(1:11) == '     atomic') then
        do i = 1, 3
          read(u, *)
        end do
        read(u, *, iostat=ios) nsym
        if (ios /= 0) call die('cannot read Nsym from '//trim(path))
        exit
      end if
    end do
    if (nsym < 1 .or. nsym > 48) call die('Nsym out of range')
    is = 0
    do while (is < nsym)
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) call die('symmetry matrices missing in '//trim(path))
      if (line(1:7) == ' cryst.') then
        is = is + 1
        do i = 1, 3
          read(u, *)
        end do
        do i = 1, 3
          read(u, '(19X,3F11.3)', iostat=ios) R(is, i, 1), R(is, i, 2), R(is, i, 3)
          if (ios /= 0) call die('bad symmetry matrix row in '//trim(path))
        end do
      end if
    end do
    close(u)
  end subroutine point_r

  ! Reciprocal axes KC(3,3), cartesian = KC * crystal, from the three
  ! '(23X,3F10.3)' rows after '     reciprocal axes: (cart. coord.'.
  subroutine load_kc(path, kc)
    character(len=*), intent(in) :: path
    real(dp), intent(out) :: kc(3, 3)
    character(len=35) :: tag
    character(len=200) :: line
    integer :: u, ios, j
    tag = '     reciprocal axes: (cart. coord.'
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('cannot open scf file '//trim(path))
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) call die('no reciprocal axes in '//trim(path))
      if (line(1:35) == tag) exit
    end do
    do j = 1, 3
      read(u, '(23X,3F10.3)', iostat=ios) kc(1, j), kc(2, j), kc(3, j)
      if (ios /= 0) call die('bad reciprocal axes row in '//trim(path))
    end do
    close(u)
  end subroutine load_kc

  ! IBZ k-points kI(3,NkI) (2pi/a0) and E(NkI,Nband) in Hartree from a
  ! bands.x file: one header line, then per k '(10X,3F10.3)' and Nband
  ! energies '(10F8.4)' in eV; bands above nfill get + shift (Hartree):
  ! Sloss 1 eV (Sloss.f90:1519-1521), Optabs dGW (Optabs.f90:1348-1350).
  subroutine load_ki_and_e(path, nki, nband, nfill, shift, ki, E)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nki, nband, nfill
    real(dp), intent(in) :: shift
    real(dp), intent(out) :: ki(3, nki), E(nki, nband)
    integer :: u, ios, ik, i
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('cannot open band file '//trim(path))
    read(u, *)
    do ik = 1, nki
      read(u, '(10X,3F10.3)', iostat=ios) ki(1, ik), ki(2, ik), ki(3, ik)
      if (ios /= 0) call die('bad k line in band file')
      read(u, '(10F8.4)', iostat=ios) (E(ik, i), i = 1, nband)
      if (ios /= 0) call die('bad energy line in band file')
    end do
    close(u)
    E = E / hartree
    E(:, nfill + 1:nband) = E(:, nfill + 1:nband) + shift
  end subroutine load_ki_and_e

  ! Miller indices mill(3,NG) of the density G list (charge-density.dat:
  ! gamma_only, ngm, nspin / b1 b2 b3 / mill), io_qe.f90:11-71.
  subroutine load_g(savedir, ngm, mill)
    character(len=*), intent(in) :: savedir
    integer, intent(out) :: ngm
    integer, allocatable, intent(out) :: mill(:, :)
    integer :: u, ios, nspin
    logical(4) :: gamma_only
    open(newunit=u, file=trim(savedir)//'/charge-density.dat', form='unformatted', &
         action='read', status='old', iostat=ios)
    if (ios /= 0) call die('cannot open '//trim(savedir)//'/charge-density.dat')
    read(u, iostat=ios) gamma_only, ngm, nspin
    if (ios /= 0) call die('cannot read NG from charge-density.dat')
    read(u)
    allocate(mill(3, ngm))
    read(u, iostat=ios) mill
    if (ios /= 0) call die('cannot read Miller indices from charge-density.dat')
    close(u)
    if (any(mill(:, 1) /= 0)) write(*, '(a)') 'WARNING: G vectors input is wrong. G(1:3,1) is not (0,0,0)!'
  end subroutine load_g

  ! igwx and nbnd of wfc{ik}.dat (second record).
  subroutine wfc_size(savedir, ik, igwx, nbnd)
    character(len=*), intent(in) :: savedir
    integer, intent(in) :: ik
    integer, intent(out) :: igwx, nbnd
    integer :: u, ios, ngw, npol
    open(newunit=u, file=wfc_path(savedir, ik), form='unformatted', action='read', &
         status='old', iostat=ios)
    if (ios /= 0) call die('cannot open '//wfc_path(savedir, ik))
    read(u)
    read(u) ngw, igwx, npol, nbnd
    close(u)
    if (npol /= 1) call die('spinor wavefunctions (npol > 1) are not supported')
  end subroutine wfc_size

  ! Miller list mill(3,igwx) and all bands evc(igwx,nbnd) of wfc{ik}.dat
  ! (io_qe.loadCs_full, io_qe.f90:124-166, which skips the Miller record).
  subroutine load_cs_full(savedir, ik, nband, mill, evc, igwx)
    character(len=*), intent(in) :: savedir
    integer, intent(in) :: ik, nband
    integer, intent(out) :: mill(:, :)
    complex(dp), intent(out) :: evc(:, :)
    integer, intent(out) :: igwx
    integer :: u, ios, ngw, npol, nbnd, i
    open(newunit=u, file=wfc_path(savedir, ik), form='unformatted', action='read', &
         status='old', iostat=ios)
    if (ios /= 0) call die('cannot open '//wfc_path(savedir, ik))
    read(u)
    read(u) ngw, igwx, npol, nbnd
    read(u)
    if (nband > nbnd) call die('wfc file holds fewer bands than Nband')
    read(u, iostat=ios) mill(1:3, 1:igwx)
    if (ios /= 0) call die('cannot read Miller indices of '//wfc_path(savedir, ik))
    do i = 1, nband
      read(u, iostat=ios) evc(1:igwx, i)
      if (ios /= 0) call die('cannot read coefficients of '//wfc_path(savedir, ik))
    end do
    close(u)
  end subroutine load_cs_full

  ! Electron count of the FBZ sum, Nel = 2 * #(E(K1(k), n) < Efermi) / Ntot
  ! over the FBZ k and their IBZ points K1 (checkFBZintegration,
  ! Sloss.f90:761-822, with findKinIBZ's K1).
  subroutine check_fbz_integration(nki, nband, ntot, K1, efermi, E, nel)
    integer, intent(in) :: nki, nband, ntot, K1(ntot)
    real(dp), intent(in) :: efermi, E(nki, nband)
    real(dp), intent(out) :: nel
    integer :: ik, n
    nel = 0.0_dp
    do ik = 1, ntot
      do n = 1, nband
        if (E(K1(ik), n) < efermi) nel = nel + 1.0_dp
      end do
    end do
    nel = 2.0_dp * nel / ntot
  end subroutine check_fbz_integration

  ! 'W_Qi' -> 'W_Q7' etc.: the digits of iq replace the 'i'
  ! (writeWT_Qi, Sloss.f90:1583-1591).
  function numbered(stem, iq) result(name)
    character(len=*), intent(in) :: stem
    integer, intent(in) :: iq
    character(len=:), allocatable :: name
    character(len=12) :: digits
    integer :: p
    write(digits, '(i0)') iq
    p = index(stem, 'i')
    name = stem(1:p - 1)//trim(digits)//stem(p + 1:)
  end function numbered

  function wfc_path(savedir, ik) result(path)
    character(len=*), intent(in) :: savedir
    integer, intent(in) :: ik
    character(len=:), allocatable :: path
    character(len=12) :: digits
    write(digits, '(i0)') ik
    path = trim(savedir)//'/wfc'//trim(digits)//'.dat'
  end function wfc_path

  subroutine die(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a)') 'ERROR: '//msg
    error stop 1
  end subroutine die

end module rpa_files
