! sloss_hip: drop-in for Sloss.x (W/Sloss.f90) on one MI355X.
!
! Same config.in namelists (Sloss.f90:39-172; the file name may also be given
! as the first argument), same inputs (scf_file, band_file, QE >= 6 savedir)
! and the same output files per q: Info, W_Qi{iq}, Kramers-Kron_Qi{iq},
! fort.20008 / 10008 / 22308 / 12308.  Added: Loss_Qi{iq} with the energy-loss
! spectrum -Im eps^-1_00(q, omega) and the run time of the device chain.
!
! Extra namelist entries (absent from the reference, defaults in brackets):
!   &directories xml_file = ''  QE's data-file-schema.xml (path under rundir):
!           when given, the operations and reciprocal axes come from it
!           (Modules/io_xml.f90) instead of scf_file.
!   &system vertex_map = 'direct' | 'igk'  ['direct'] coefficient-to-G map:
!           'direct' is the reference drivers' first-NGd convention
!           (Sloss.f90:907-946), 'igk' the Modules/vertices.f90 _v2 one.
!   &system precision = 'fp64' | 'mixed'  ['fp64'] arithmetic of the device
!           chain: 'fp64' throughout, or complex FP32 vertices and BF16 hi/lo
!           split-product S0 with FP64 weights, reduction, KK and Dyson/W.
!
! Each q runs the device chain with rpa_run (its kernels on one stream; measured
! faster than a replay of the captured chain, rpa_run_graph, which gives the
! same results bit for bit).
! A zero pivot in eps = I - chi0 V prints the reference's 'singular matrix
! in gjel' (Chi_RPA/gjel.f:59-61) and the run goes on, as the reference's.
program sloss_hip
  use iso_c_binding
  use rpa_hip
  use rpa_files
  implicit none

  character(len=100) :: rundir = './', savedir = './', scf_file = 'scf.out', band_file = 'bands.dat', &
                        xml_file = ''
  character(len=3) :: lf = 'z'
  character(len=8) :: vertex_map = 'direct', precision = 'fp64'
  integer :: jump = 1, qmin = 1, qmax = 1
  real(dp) :: omin = 0.0_dp, omax = 1.0_dp
  integer :: NkI = 0, Nband = 0, Nocc = 0, NelQE = 0, NGd = 0, NG = 0, no = 0, nq = 1, Nlfd = 50
  real(dp) :: Efermi = 0.0_dp, a0 = 1.0_dp, c0 = 1.0_dp, eps = 1.0e-4_dp, T = 0.0_dp, &
              eta = 0.05_dp, Ecut = 0.0_dp, Vcell = 1.0_dp
  integer :: Nthreads = 1
  namelist /directories/ rundir, savedir, scf_file, band_file, xml_file
  namelist /system/ lf, jump, omin, omax, qmin, qmax, vertex_map, precision
  namelist /config/ NkI, Nband, Nocc, NelQE, NGd, NG, no, nq, Nlfd
  namelist /parameters/ Efermi, a0, c0, eps, T, eta, Ecut, Vcell
  namelist /parallel/ Nthreads

  real(dp), parameter :: pi = 4.0_dp * atan(1.0_dp)
  character(len=256) :: cfgfile
  integer :: u, ios, nsym, ngm, ik, iq, io, igwx, nbnd, npwx, ntot, nlf, i, rc
  real(dp) :: R(48, 3, 3), RI(48, 3, 3), q0(3), kc(3, 3), gcar, domega, nel, q(3), oi, fact
  real(dp) :: kks, skk, wind, windkk, absq, err
  real(dp), allocatable :: ki(:, :), E(:, :), ktot(:, :), g(:, :), glf(:, :), V(:, :), loss(:), &
                           w1(:), w2(:)
  integer, allocatable :: mill(:, :), millk(:, :), parg(:), lf_idx(:), K1(:), K2(:), iG0(:), R1(:), R2(:)
  complex(dp), allocatable, target :: evc(:, :), WT(:, :, :)
  real(dp), allocatable, target :: Vt(:, :), losst(:)
  type(c_ptr) :: ctx
  type(qe_xml) :: qx
  type(rpa_dims_t) :: dims
  type(rpa_params_t) :: p

  cfgfile = 'config.in'
  if (command_argument_count() >= 1) call get_command_argument(1, cfgfile)
  open(newunit=u, file=trim(cfgfile), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*, '(a)') 'ERROR: cannot open '//trim(cfgfile)
    error stop 1
  end if
  read(u, nml=directories, iostat=ios); rewind(u)
  read(u, nml=system, iostat=ios); rewind(u)
  read(u, nml=config, iostat=ios); rewind(u)
  read(u, nml=parameters, iostat=ios); rewind(u)
  read(u, nml=parallel, iostat=ios)
  close(u)

  ! unit conversions, Sloss.f90:190-212
  T = T / hartree
  Efermi = Efermi / hartree
  eta = eta / hartree
  gcar = 2.0_dp * pi / a0
  domega = (omax - omin) / (no - 1)

  if (len_trim(xml_file) > 0) then
    ! operations and reciprocal axes from data-file-schema.xml (io_xml.f90)
    call load_xml_qe(trim(rundir)//trim(xml_file), qx)
    call xml_point_ops(qx, nsym, R)
  else
    call point_r(trim(rundir)//trim(scf_file), nsym, R)
  end if
  RI = 0.0_dp
  call rpa_check(rpa_point_inverse(nsym, R, RI), c_null_ptr, 'PointR')
  allocate(ki(3, NkI), E(NkI, Nband), ktot(3, 48 * NkI))
  call load_ki_and_e(trim(rundir)//trim(band_file), NkI, Nband, Nocc, 1.0_dp / hartree, ki, E)
  call rpa_check(rpa_gen_fbz(NkI, nsym, eps, ki, R, ntot, ktot), c_null_ptr, 'genFBZ')
  if (len_trim(xml_file) > 0) then
    call xml_kc(qx, kc)
  else
    call load_kc(trim(rundir)//trim(scf_file), kc)
  end if
  call load_g(trim(savedir), ngm, mill)
  if (ngm /= NG) write(*, '(a,i0,a,i0)') 'note: NG = ', NG, ' in config, charge-density.dat has ', ngm
  allocate(g(3, ngm), glf(3, Nlfd), parg(Nlfd), lf_idx(Nlfd))
  call rpa_check(rpa_gcart(ngm, mill, kc, g), c_null_ptr, 'G vectors')
  call rpa_check(rpa_gen_glf(trim(lf)//c_null_char, Ecut, gcar, ngm, g, mill, Nlfd, nlf, glf, parg, &
                             lf_idx), c_null_ptr, 'genGlfandParity')
  ! findKinIBZ for every FBZ k (K1 does not depend on q), then the FBZ
  ! electron count over those IBZ points
  allocate(R1(ntot), K1(ntot), R2(ntot), K2(ntot), iG0(ntot))
  q0 = 0.0_dp
  call rpa_check(rpa_find_kq_sym(ntot, ktot, NkI, ki, ngm, g, q0, nsym, RI, eps, R1, K1, R2, K2, iG0), &
                 c_null_ptr, 'findKinIBZ')
  call check_fbz_integration(NkI, Nband, ntot, K1, Efermi, E, nel)
  write(*, '(a,i0,a,i0)') ' Nlf: ', nlf, ' Nlfd: ', Nlfd

  npwx = 0
  do ik = 1, NkI
    call wfc_size(trim(savedir), ik, igwx, nbnd)
    npwx = max(npwx, igwx)
  end do
  dims = rpa_dims_t(NkI, Nband, npwx, Nlfd, no, RPA_SLOSS, rpa_precision_of(precision), 0)
  call rpa_check(rpa_ctx_create(dims, ctx), c_null_ptr, 'rpa_ctx_create')
  call rpa_check(rpa_set_gvectors(ctx, ngm, mill, kc), ctx, 'rpa_set_gvectors')
  if (nsym > 1) call rpa_check(rpa_set_symmetry(ctx, nsym, R), ctx, 'rpa_set_symmetry')
  allocate(millk(3, npwx), evc(npwx, Nband))
  do ik = 1, NkI
    call load_cs_full(trim(savedir), ik, Nband, millk, evc, igwx)
    call rpa_check(rpa_upload_basis(ctx, ik, igwx, millk, evc(1:igwx, 1:Nband)), ctx, 'rpa_upload_basis')
  end do
  call rpa_check(rpa_set_bands(ctx, E), ctx, 'rpa_set_bands')
  call rpa_check(rpa_set_local_fields(ctx, nlf, lf_idx, parg), ctx, 'rpa_set_local_fields')
  p%nval = Nocc
  p%vertex_map = merge(RPA_MAP_IGK, RPA_MAP_DIRECT, trim(vertex_map) == 'igk')
  p%ngd = NGd
  p%pol = RPA_POL_X
  p%eps = eps
  p%gcar = gcar
  p%domega = domega
  p%eta = eta
  p%lor_cut = 1.0e-3_dp
  p%df_cut = 0.0_dp
  p%efermi = Efermi
  p%temp = T
  p%vcell = Vcell
  p%c0 = c0
  p%eta_intra = 0.0_dp
  call rpa_check(rpa_set_params(ctx, p), ctx, 'rpa_set_params')

  allocate(WT(no - 1, nlf, nlf), Vt(nlf, nlf), losst(no - 1), &
           w1(no - 1), w2(no - 1))
  do iq = qmin, qmax
    call rpa_check(rpa_find_min_q(ntot, ktot, iq, q), ctx, 'findMinQ')
    call write_info(q)
    if (nsym > 1) then
      call rpa_check(rpa_find_kq_sym(ntot, ktot, NkI, ki, ngm, g, q, nsym, RI, eps, R1, K1, R2, K2, iG0), &
                     ctx, 'findKQinBZ')
      call rpa_check(rpa_set_q_sym(ctx, ntot, K1, K2, iG0, R1, R2, ktot, q), ctx, 'rpa_set_q_sym')
    else
      call rpa_check(rpa_find_kq(ntot, ktot, NkI, ki, ngm, g, q, eps, K1, K2, iG0), ctx, 'findKQinBZ')
      call rpa_check(rpa_set_q(ctx, ntot, K1, K2, iG0, ktot, q), ctx, 'rpa_set_q')
    end if
    rc = rpa_run(ctx)
    if (rc == RPA_E_SINGULAR) then
      print *, 'singular matrix in gjel'
    else
      call rpa_check(rc, ctx, 'rpa_run')
    end if
    write(*, '(a,i0,a,f10.3,a)') ' q ', iq, ': device chain ', rpa_last_run_ms(ctx), ' ms'
    call rpa_check(rpa_get_result(ctx, RPA_RES_W, c_loc(WT)), ctx, 'W')
    call rpa_check(rpa_get_result(ctx, RPA_RES_V, c_loc(Vt)), ctx, 'V')
    call rpa_check(rpa_get_result(ctx, RPA_RES_LOSS, c_loc(losst)), ctx, 'loss')

    ! per-frequency W traces, Sloss.f90:472-479
    call open_trace(20008, iq == qmin)
    call open_trace(10008, iq == qmin)
    call open_trace(22308, iq == qmin)
    call open_trace(12308, iq == qmin)
    do io = 1, no - 1
      oi = (io - 1) * domega
      write(20008, *) oi * hartree, aimag(WT(io, 1, 1))
      write(10008, *) oi * hartree, real(WT(io, 1, 1))
      if (nlf >= 2) write(22308, *) oi * hartree, aimag(WT(io, 2, 2))
      if (nlf >= 3) write(12308, *) oi * hartree, real(WT(io, 2, 3))
    end do
    close(20008); close(10008); close(22308); close(12308)

    ! W_Qi: io = 1 only, '(10F15.5)' (writeWT_Qi, Sloss.f90:1572-1601)
    open(newunit=u, file=numbered('W_Qi', iq))
    write(u, *) 'omega=', 0.0_dp, 'Hartree'
    write(u, '(10F15.5)') ((WT(1, i, io), io = 1, nlf), i = 1, nlf)
    close(u)

    ! Kramers-Kronig check of W (Sloss.f90:489-534, 1603-1647)
    call rpa_check(rpa_w_kk(ctx, w1, w2), ctx, 'rpa_w_kk')
    kks = 0.0_dp
    skk = 0.0_dp
    do io = 1, no - 1
      wind = real(WT(io, 1, 1) - Vt(1, 1))
      windkk = w1(io) - w2(io)
      fact = domega
      if (io == 1 .or. io == no - 1) fact = 0.5_dp * domega
      kks = kks + fact * (windkk - wind) * (windkk - wind)
      skk = skk + fact * wind * wind
    end do
    absq = sqrt(sum((gcar * q)**2))
    open(newunit=u, file=numbered('Kramers-Kron_Qi', iq))
    write(u, '(a25,3f10.4,a5)') 'Wave vector (qx,qy,qz)=(', gcar * q(1) * gcar, gcar * q(2) * gcar, &
      gcar * q(3) * gcar, ') a.u.'
    write(u, '(a25,f8.4,a5)') '|(qx,qy,qz)|=', absq * gcar, 'a.u.'
    write(u, *) 'int(WindKK-Wind)^2 =  ', kks
    write(u, *) 'int(Wind)^2 =  ', skk
    write(u, *) '****************************************'
    write(u, *) 'Kramers-Kronig relation relative error'
    write(u, '(5X,f7.2,a2)') 100.0_dp * abs(kks / skk), '%'
    write(u, *) 'Usporedba Gamma i WT'
    write(u, *) 'real[Gamma(o=0,1,1)]=', w1(1)
    write(u, *) 'real[WT(o=0,1,1)]/2=', real(WT(1, 1, 1) - Vt(1, 1)) / 2.0_dp
    close(u)

    ! energy-loss spectrum
    open(newunit=u, file=numbered('Loss_Qi', iq))
    do io = 1, no - 1
      write(u, '(2es24.15)') (io - 1) * domega * hartree, losst(io)
    end do
    close(u)
  end do

  call rpa_check(rpa_ctx_destroy(ctx), c_null_ptr, 'rpa_ctx_destroy')

contains

  ! fort.<unit>: the reference writes these units without opening them, so a
  ! run starts them afresh and every q of the run appends
  subroutine open_trace(unit, first)
    integer, intent(in) :: unit
    logical, intent(in) :: first
    character(len=16) :: name
    write(name, '(a,i0)') 'fort.', unit
    if (first) then
      open(unit, file=trim(name), status='replace', action='write')
    else
      open(unit, file=trim(name), position='append', action='write')
    end if
  end subroutine open_trace

  ! Info (writeInfo, Sloss.f90:1525-1570); q in 2pi/a0 units
  subroutine write_info(qv)
    real(dp), intent(in) :: qv(3)
    integer :: w
    absq = sqrt(sum(qv**2))
    open(newunit=w, file='Info')
    write(w, *) '***************General***********************'
    write(w, *) ''
    write(w, *) 'Number of point symmetry operation is', nsym
    write(w, '(a25,3f10.4,a5)') 'Wave vector (qx,qy,qz)=(', qv(1) * gcar, qv(2) * gcar, qv(3) * gcar, ') a.u.'
    write(w, '(a25,f8.4,a5)') '|(qx,qy,qz)|=', absq * gcar, 'a.u.'
    write(w, *) 'Local field effcts in '//lf//'-dir'
    write(w, *) 'Number of local field vectors is', nlf
    write(w, *) 'Number of different K vectors in 1.B.Z. is', ntot
    write(w, *) 'Number of K vectors in I.B.Z. is', NkI
    write(w, *) 'Number of bands is               ', Nband
    write(w, '(a25,f8.4,a5)') 'Eta damping is ', eta * hartree * 1000.0_dp, 'meV'
    write(w, '(a25,f8.4,a5)') 'Temperature is  ', T * hartree * 1000.0_dp, 'meV'
    write(w, *) ''
    write(w, *) '************* Checking 1BZ integration*******'
    write(w, *) ''
    write(w, '(a40,f7.4)') 'Number of electrons(1BZ integration)=', nel
    write(w, *) 'Number of electrons(unit cell)=', NelQE
    err = abs((NelQE - nel) / NelQE)
    write(w, '(a25,f8.4,a5)') 'Relative error=', err * 100.0_dp, '%'
    if (err > 0.05_dp) then
      write(w, *) 'WARRNING!!-1BZ INTEGRATION IS BAD!.'
      close(w)
      error stop 1
    end if
    close(w)
  end subroutine write_info

end program sloss_hip
