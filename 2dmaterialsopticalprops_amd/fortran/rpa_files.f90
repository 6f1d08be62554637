! Input readers and output writers of the Optabs / Sloss drivers.
!
! The formats are the reference's:
!   scf_file   PointR.f90:41-91 (Nsym, symmetry matrices), loadKC
!              (Sloss.f90:1279-1307, Optabs.f90:652-680)
!   band_file  loadkIandE (Sloss.f90:1492-1523, Optabs.f90:1320-1352)
!   savedir    QE >= 6 charge-density.dat and wfc{ik}.dat records
!              (Modules/io_qe.f90:11-166)
!   xml_file   QE's data-file-schema.xml (Modules/io_xml.f90:24-235): the
!              lattices, symmetry operations, k-points, bands and scalars
! Output files are written by the drivers with the reference's formats.
module rpa_files
  use iso_fortran_env, only: dp => real64
  implicit none
  private
  public :: dp, hartree, point_r, load_kc, load_ki_and_e, load_g, wfc_size, load_cs_full, &
            check_fbz_integration, numbered, qe_xml, load_xml_qe, xml_kc, xml_point_ops

  ! What loadXML_qe returns (io_xml.f90:24-41): a(i,:) and b(i,:) are the
  ! direct (bohr) and reciprocal (2pi/alat) lattice vectors i, R(:,:,op) the
  ! operations as the file lists them (crystal axes, the file's three lines
  ! the three columns), ki(3,NkI) cartesian in 2pi/alat, eig and occ
  ! (Nband, NkI) in Hartree / fractions.
  type :: qe_xml
    integer :: nband = 0, nki = 0, nrot = 0, nsym = 0, nmp(3) = 0
    real(dp) :: ecutwfc = 0.0_dp, ecutrho = 0.0_dp, alat = 0.0_dp, temp = 0.0_dp, efermi = 0.0_dp, &
                nelec = 0.0_dp
    real(dp) :: a(3, 3) = 0.0_dp, b(3, 3) = 0.0_dp
    logical :: noinv = .false., force_symmorphic = .false.
    real(dp), allocatable :: R(:, :, :), ki(:, :), eig(:, :), occ(:, :)
    integer, allocatable :: npw(:)
  end type qe_xml

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

  ! data-file-schema.xml of a pw.x run (io_xml.f90:24-235).  Two passes: the
  ! counts first (<nrot>, <nsym>, <nks>, <nbnd>, the last occurrence of a
  ! scalar tag counting, as in the reference), then the arrays: every
  ! <rotation> (nine values), every <k_point> (three), <npw>, <eigenvalues>
  ! and <occupations> (nbnd values each, on as many lines as the file uses).
  ! The lattices come from the first <cell> and <reciprocal_lattice> blocks,
  ! alat = a(1,1) as the reference sets it, T from the degauss attribute of
  ! <smearing>.  Stops as the reference does when the run used time reversal
  ! (noinv false) or fractional translations (force_symmorphic false); warns
  ! when Nsym /= Nrot.
  subroutine load_xml_qe(path, x)
    character(len=*), intent(in) :: path
    type(qe_xml), intent(out) :: x
    character(len=1024) :: line
    integer :: u, ios, pass, irot, ikp, inpw, ieig, iocc, icell, irec
    real(dp) :: v(9)
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('cannot open '//trim(path))
    do pass = 1, 2
      rewind(u)
      irot = 0; ikp = 0; inpw = 0; ieig = 0; iocc = 0; icell = 0; irec = 0
      do
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) exit
        if (pass == 1) then
          call scalar_i(line, 'nrot', x%nrot)
          call scalar_i(line, 'nsym', x%nsym)
          call scalar_i(line, 'nks', x%nki)
          call scalar_i(line, 'nbnd', x%nband)
          call scalar_r(line, 'nelec', x%nelec)
          call scalar_r(line, 'fermi_energy', x%efermi)
          call scalar_r(line, 'ecutwfc', x%ecutwfc)
          call scalar_r(line, 'ecutrho', x%ecutrho)
          if (index(line, '<noinv>') > 0) x%noinv = index(line, 'true') > 0
          if (index(line, '<force_symmorphic>') > 0) x%force_symmorphic = index(line, 'true') > 0
          if (index(line, '<smearing') > 0) call attr_r(line, 'degauss', x%temp)
          if (index(line, '<monkhorst_pack') > 0) then
            call attr_i(line, 'nk1', x%nmp(1))
            call attr_i(line, 'nk2', x%nmp(2))
            call attr_i(line, 'nk3', x%nmp(3))
          end if
          if (index(line, '<cell>') > 0 .and. icell == 0) then
            icell = 1
            call read_vectors(u, line, x%a)
          end if
          if (index(line, '<reciprocal_lattice>') > 0 .and. irec == 0) then
            irec = 1
            call read_vectors(u, line, x%b)
          end if
        else
          if (index(line, '<rotation') > 0) then
            irot = irot + 1
            call collect(u, line, 'rotation', 9, v)
            if (irot <= x%nrot) x%R(:, :, irot) = reshape(v(1:9), [3, 3])
          else if (index(line, '<k_point') > 0 .and. index(line, '<k_points') == 0) then
            ikp = ikp + 1
            call collect(u, line, 'k_point', 3, v)
            if (ikp <= x%nki) x%ki(:, ikp) = v(1:3)
          else if (index(line, '<npw>') > 0) then
            inpw = inpw + 1
            if (inpw <= x%nki) call scalar_i(line, 'npw', x%npw(inpw))
          else if (index(line, '<eigenvalues') > 0) then
            ieig = ieig + 1
            if (ieig <= x%nki) call collect(u, line, 'eigenvalues', x%nband, x%eig(:, ieig))
          else if (index(line, '<occupations') > 0 .and. index(line, '<occupations>') == 0) then
            iocc = iocc + 1
            if (iocc <= x%nki) call collect(u, line, 'occupations', x%nband, x%occ(:, iocc))
          end if
        end if
      end do
      if (pass == 1) then
        x%alat = x%a(1, 1)
        allocate(x%R(3, 3, max(x%nrot, 1)), x%ki(3, max(x%nki, 1)), x%npw(max(x%nki, 1)), &
                 x%eig(max(x%nband, 1), max(x%nki, 1)), x%occ(max(x%nband, 1), max(x%nki, 1)))
        x%R = 0.0_dp; x%ki = 0.0_dp; x%npw = 0; x%eig = 0.0_dp; x%occ = 0.0_dp
      else
        if (irot < x%nrot) call die('fewer <rotation> elements than nrot in '//trim(path))
        if (ikp < x%nki) call die('fewer <k_point> elements than nks in '//trim(path))
      end if
    end do
    close(u)
    if (x%nsym /= x%nrot) write(*, '(a)') 'WARNING: Quantum Espresso calculation must not use '// &
      'non-rotational symmetries. Use force_symmorphic=.true. and noinv=.true. in QE input.'
    if (.not. x%noinv) call die('Quantum espresso uses time reversal (k=-k) symmetry. '// &
                                'Use noinv=.true. in QE input.')
    if (.not. x%force_symmorphic) call die('Quantum espresso uses fractionary translation symmetry. '// &
                                           'Use force_symmorphic=.true. in QE input.')
  end subroutine load_xml_qe

  ! Reciprocal axes KC(3,3) of the drivers (cartesian = KC * crystal, 2pi/alat):
  ! column j is b_j.
  subroutine xml_kc(x, kc)
    type(qe_xml), intent(in) :: x
    real(dp), intent(out) :: kc(3, 3)
    kc = transpose(x%b)
  end subroutine xml_kc

  ! The first Nsym operations in the cartesian form of point_r, R(op,i,j):
  ! a crystal operation s acts on reciprocal crystal components (pw.x's
  ! k-point rotation, k' = s k), so R = KC s KC^-1.
  subroutine xml_point_ops(x, nsym, R)
    type(qe_xml), intent(in) :: x
    integer, intent(out) :: nsym
    real(dp), intent(out) :: R(48, 3, 3)
    real(dp) :: kc(3, 3), kci(3, 3), det
    integer :: op
    call xml_kc(x, kc)
    det = kc(1, 1) * (kc(2, 2) * kc(3, 3) - kc(2, 3) * kc(3, 2)) - kc(1, 2) * (kc(2, 1) * kc(3, 3) - &
          kc(2, 3) * kc(3, 1)) + kc(1, 3) * (kc(2, 1) * kc(3, 2) - kc(2, 2) * kc(3, 1))
    if (abs(det) < 1.0e-12_dp) call die('singular reciprocal lattice in the xml file')
    kci(1, 1) = kc(2, 2) * kc(3, 3) - kc(2, 3) * kc(3, 2)
    kci(1, 2) = kc(1, 3) * kc(3, 2) - kc(1, 2) * kc(3, 3)
    kci(1, 3) = kc(1, 2) * kc(2, 3) - kc(1, 3) * kc(2, 2)
    kci(2, 1) = kc(2, 3) * kc(3, 1) - kc(2, 1) * kc(3, 3)
    kci(2, 2) = kc(1, 1) * kc(3, 3) - kc(1, 3) * kc(3, 1)
    kci(2, 3) = kc(1, 3) * kc(2, 1) - kc(1, 1) * kc(2, 3)
    kci(3, 1) = kc(2, 1) * kc(3, 2) - kc(2, 2) * kc(3, 1)
    kci(3, 2) = kc(1, 2) * kc(3, 1) - kc(1, 1) * kc(3, 2)
    kci(3, 3) = kc(1, 1) * kc(2, 2) - kc(1, 2) * kc(2, 1)
    kci = kci / det
    nsym = x%nsym
    if (nsym < 1 .or. nsym > 48 .or. nsym > x%nrot) call die('Nsym out of range in the xml file')
    R = 0.0_dp
    do op = 1, nsym
      R(op, :, :) = matmul(kc, matmul(x%R(:, :, op), kci))
    end do
    where (abs(R) < 1.0e-12_dp) R = 0.0_dp
  end subroutine xml_point_ops

  ! the text between '<tag ...>' and '</tag>' of a one-line element
  function element_text(line, tag, found) result(t)
    character(len=*), intent(in) :: line, tag
    logical, intent(out) :: found
    character(len=len(line)) :: t
    integer :: p, q, r
    t = ''
    found = .false.
    p = index(line, '<'//tag//'>')
    if (p == 0) p = index(line, '<'//tag//' ')
    if (p == 0) return
    q = index(line(p:), '>') + p
    r = index(line, '</'//tag)
    if (r == 0) r = len_trim(line) + 1
    t = line(q:r - 1)
    found = .true.
  end function element_text

  subroutine scalar_i(line, tag, v)
    character(len=*), intent(in) :: line, tag
    integer, intent(inout) :: v
    character(len=len(line)) :: t
    logical :: found
    integer :: ios, w
    t = element_text(line, tag, found)
    if (.not. found) return
    read(t, *, iostat=ios) w
    if (ios == 0) v = w
  end subroutine scalar_i

  subroutine scalar_r(line, tag, v)
    character(len=*), intent(in) :: line, tag
    real(dp), intent(inout) :: v
    character(len=len(line)) :: t
    logical :: found
    integer :: ios
    real(dp) :: w
    t = element_text(line, tag, found)
    if (.not. found) return
    read(t, *, iostat=ios) w
    if (ios == 0) v = w
  end subroutine scalar_r

  ! value of attribute name="..." on the line
  subroutine attr_text(line, name, t, found)
    character(len=*), intent(in) :: line, name
    character(len=*), intent(out) :: t
    logical, intent(out) :: found
    integer :: p, q
    t = ''
    found = .false.
    p = index(line, ' '//name//'="')
    if (p == 0) return
    p = p + len(name) + 3
    q = index(line(p:), '"') + p - 2
    t = line(p:q)
    found = .true.
  end subroutine attr_text

  subroutine attr_r(line, name, v)
    character(len=*), intent(in) :: line, name
    real(dp), intent(inout) :: v
    character(len=128) :: t
    logical :: found
    integer :: ios
    real(dp) :: w
    call attr_text(line, name, t, found)
    if (.not. found) return
    read(t, *, iostat=ios) w
    if (ios == 0) v = w
  end subroutine attr_r

  subroutine attr_i(line, name, v)
    character(len=*), intent(in) :: line, name
    integer, intent(inout) :: v
    character(len=128) :: t
    logical :: found
    integer :: ios, w
    call attr_text(line, name, t, found)
    if (.not. found) return
    read(t, *, iostat=ios) w
    if (ios == 0) v = w
  end subroutine attr_i

  ! the numbers of an element that starts on `line`: those after its '>' on
  ! that line, then whole lines until n values are read or '</tag' is met
  subroutine collect(u, line, tag, n, v)
    integer, intent(in) :: u, n
    character(len=*), intent(in) :: line, tag
    real(dp), intent(out) :: v(n)
    character(len=1024) :: cur
    integer :: got, p, ios
    logical :: last
    v = 0.0_dp
    got = 0
    p = index(line, '<'//tag)
    p = index(line(p:), '>') + p
    cur = line(p:)
    do
      last = index(cur, '</'//tag) > 0
      if (last) cur = cur(1:index(cur, '</'//tag) - 1)
      call numbers(cur, n, v, got)
      if (got >= n .or. last) exit
      read(u, '(a)', iostat=ios) cur
      if (ios /= 0) exit
    end do
    if (got < n) call die('element <'//tag//'> holds fewer values than expected')
  end subroutine collect

  ! appends the blank-separated numbers of s to v(got+1:), at most n in all
  subroutine numbers(s, n, v, got)
    character(len=*), intent(in) :: s
    integer, intent(in) :: n
    real(dp), intent(inout) :: v(n)
    integer, intent(inout) :: got
    integer :: i, j, ln, ios
    ln = len_trim(s)
    i = 1
    do while (i <= ln .and. got < n)
      do while (i <= ln)
        if (s(i:i) /= ' ' .and. s(i:i) /= char(9)) exit
        i = i + 1
      end do
      if (i > ln) exit
      j = i
      do while (j <= ln)
        if (s(j:j) == ' ' .or. s(j:j) == char(9)) exit
        j = j + 1
      end do
      got = got + 1
      read(s(i:j - 1), *, iostat=ios) v(got)
      if (ios /= 0) call die('not a number in the xml file: '//s(i:j - 1))
      i = j
    end do
  end subroutine numbers

  ! three vectors <x1>..</x1> .. <x3>..</x3> on the lines after a block tag
  subroutine read_vectors(u, first, m)
    integer, intent(in) :: u
    character(len=*), intent(in) :: first
    real(dp), intent(out) :: m(3, 3)
    character(len=1024) :: line
    integer :: i, ios, got
    real(dp) :: v(3)
    do i = 1, 3
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) call die('lattice block cut short after '//trim(first))
      line = line(index(line, '>') + 1:)
      if (index(line, '</') > 0) line = line(1:index(line, '</') - 1)
      got = 0
      call numbers(line, 3, v, got)
      if (got /= 3) call die('bad lattice vector line after '//trim(first))
      m(i, :) = v
    end do
  end subroutine read_vectors

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
