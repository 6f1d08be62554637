! optabs_hip: drop-in for Optabs.x (Chi_RPA/Optabs.f90) on one MI355X.
!
! Same config.in namelists (Optabs.f90:14-103, 142-150; the file name may
! also be given as the first argument), same inputs and the same outputs:
!   calc = 1  S0 -> Corrfun_{pol}, Qeff_{pol}            (Optabs.f90:480-510)
!   calc = 2  Corrfun_{pol}, Qeff_{pol} -> KK -> Pi_RPA_{pol}, fort.401/402
!   calc = 3  both, the second part reading the files the first one wrote
!             (Optabs.f90:513-614), plus Info_{pol}.out.
!
! Extra namelist entries (absent from the reference, defaults in brackets):
!   &directories xml_file = ''  QE's data-file-schema.xml (path under rundir):
!           when given, the operations and reciprocal axes come from it
!           (Modules/io_xml.f90) instead of scf_file.
!   &system vertex_map = 'direct' | 'igk' ['direct']
!   &system precision = 'fp64' | 'mixed'  ['fp64'] arithmetic of the device
!           chain: 'fp64' throughout, or complex FP32 vertices and BF16 hi/lo
!           split-product S0 with FP64 weights, reduction, KK and Pi.
! Each q of calc = 1 / 3 runs the device chain with rpa_run (measured faster
! than a replay of the captured chain, rpa_run_graph; same results).
program optabs_hip
  use iso_c_binding
  use rpa_hip
  use rpa_files
  implicit none

  character(len=100) :: rundir = './', savedir = './', scf_file = 'scf.out', band_file = 'bands.dat', &
                        xml_file = ''
  character(len=3) :: lf = 'z', pol = 'xx'
  character(len=8) :: vertex_map = 'direct', precision = 'fp64'
  integer :: calc = 3, jump = 1, qmin = 1, qmax = 1
  real(dp) :: omin = 0.0_dp, omax = 1.0_dp, eta_intra = 0.01_dp, eta_inter = 0.05_dp, &
              Lor_cut = 1.0e-3_dp, df_cut = 1.0e-3_dp
  integer :: NG = 0, NGd = 0, NkI = 0, Nband = 0, Nval = 0, NelQE = 0, no = 0, Nlfd = 50
  real(dp) :: Efermi = 0.0_dp, dGW = 0.0_dp, eps = 1.0e-4_dp, T = 0.0_dp, Ecut = 0.0_dp, &
              a0 = 1.0_dp, c0 = 1.0_dp, Vcell = 1.0_dp
  integer :: Nthreads = 1
  namelist /directories/ rundir, savedir, scf_file, band_file, xml_file
  namelist /system/ lf, pol, calc, jump, omin, omax, qmin, qmax, eta_intra, eta_inter, Lor_cut, &
                    df_cut, vertex_map, precision
  namelist /config/ NG, NGd, NkI, Nband, Nval, NelQE, no, Nlfd
  namelist /parameters/ Efermi, dGW, eps, T, Ecut, a0, c0, Vcell
  namelist /parallel/ Nthreads

  real(dp), parameter :: pi = 4.0_dp * atan(1.0_dp)
  character(len=256) :: cfgfile
  character(len=100) :: dato1, dato2, dato3
  integer :: u, ios, nsym, ngm, ik, iq, io, iG, jG, igwx, nbnd, npwx, ntot, nlf
  real(dp) :: R(48, 3, 3), RI(48, 3, 3), q0(3), kc(3, 3), gcar, domega, nel, q(3), o, oi
  real(dp), allocatable :: ki(:, :), E(:, :), ktot(:, :), g(:, :), glf(:, :)
  integer, allocatable :: mill(:, :), millk(:, :), parg(:), lf_idx(:), K1(:), K2(:), iG0(:), R1(:), R2(:)
  complex(dp), allocatable :: evc(:, :)
  complex(dp), allocatable, target :: S0(:, :, :), Qeff(:, :), pi11(:)
  real(dp), allocatable, target :: sig(:, :)
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

  dato1 = 'Corrfun_'//adjustl(trim(pol))
  dato2 = 'Qeff_'//adjustl(trim(pol))
  dato3 = 'Pi_RPA_'//adjustl(trim(pol))

  ! unit conversions, Optabs.f90:174-204
  T = T / hartree
  Efermi = Efermi / hartree
  gcar = 2.0_dp * pi / a0
  dGW = dGW / hartree
  eta_intra = eta_intra / hartree
  eta_inter = eta_inter / hartree
  omin = omin / hartree
  omax = omax / hartree + omin
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
  call load_ki_and_e(trim(rundir)//trim(band_file), NkI, Nband, Nval, dGW, ki, E)
  call rpa_check(rpa_gen_fbz(NkI, nsym, eps, ki, R, ntot, ktot), c_null_ptr, 'genFBZ')
  if (len_trim(xml_file) > 0) then
    call xml_kc(qx, kc)
  else
    call load_kc(trim(rundir)//trim(scf_file), kc)
  end if
  call load_g(trim(savedir), ngm, mill)
  allocate(g(3, ngm), glf(3, Nlfd), parg(Nlfd), lf_idx(Nlfd))
  call rpa_check(rpa_gcart(ngm, mill, kc, g), c_null_ptr, 'G vectors')
  call rpa_check(rpa_gen_glf(trim(lf)//c_null_char, Ecut, gcar, ngm, g, mill, Nlfd, nlf, glf, parg, &
                             lf_idx), c_null_ptr, 'genGlf')
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
  dims = rpa_dims_t(NkI, Nband, npwx, Nlfd, no, RPA_OPTABS, rpa_precision_of(precision), 0)
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
  p%nval = Nval
  p%vertex_map = merge(RPA_MAP_IGK, RPA_MAP_DIRECT, trim(vertex_map) == 'igk')
  p%ngd = NGd
  select case (trim(pol))
  case ('xx'); p%pol = RPA_POL_X
  case ('yy'); p%pol = RPA_POL_Y
  case ('zz'); p%pol = RPA_POL_Z
  case ('yz'); p%pol = RPA_POL_YZ
  case default
    write(*, '(a)') 'ERROR: pol must be xx, yy, zz or yz'
    error stop 1
  end select
  p%eps = eps
  p%gcar = gcar
  p%domega = domega
  p%eta = eta_inter
  p%lor_cut = Lor_cut
  p%df_cut = df_cut
  p%efermi = Efermi
  p%temp = T
  p%vcell = Vcell
  p%c0 = c0
  p%eta_intra = eta_intra
  call rpa_check(rpa_set_params(ctx, p), ctx, 'rpa_set_params')

  allocate(S0(-no:no, nlf, nlf), Qeff(nlf, nlf), pi11(no - 1), &
           sig(no - 1, 2))
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

    if (calc /= 2) then
      call rpa_check(rpa_run(ctx), ctx, 'rpa_run')
      write(*, '(a,i0,a,f10.3,a)') ' q ', iq, ': device chain ', rpa_last_run_ms(ctx), ' ms'
      call rpa_check(rpa_get_result(ctx, RPA_RES_S0, c_loc(S0)), ctx, 'S0')
      call rpa_check(rpa_get_result(ctx, RPA_RES_QEFF, c_loc(Qeff)), ctx, 'Qeff')
      ! Corrfun_{pol} and Qeff_{pol}, Optabs.f90:480-508
      open(newunit=u, file=trim(dato1))
      do io = -no, no
        o = io * domega
        write(u, *) 'omega=', o, 'Hartree'
        do iG = 1, nlf
          do jG = 1, nlf
            write(u, '(2F15.10)') real(S0(io, iG, jG)), aimag(S0(io, iG, jG))
          end do
        end do
      end do
      close(u)
      open(newunit=u, file=trim(dato2))
      do iG = 1, nlf
        do jG = 1, nlf
          write(u, '(2F15.10)') real(Qeff(iG, jG)), aimag(Qeff(iG, jG))
        end do
      end do
      close(u)
      write(*, '(a)') ' PROGRAM EXECUTION ENDED FOR CALC = 1'
    end if

    if (calc /= 1) then
      ! second part: read the text files back, KK on the device, Pi
      ! (Optabs.f90:513-614)
      open(newunit=u, file=trim(dato1), status='old', action='read', iostat=ios)
      if (ios /= 0) then
        write(*, '(a)') 'ERROR: cannot open '//trim(dato1)
        error stop 1
      end if
      do io = -no, no
        read(u, *)
        do iG = 1, nlf
          do jG = 1, nlf
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
      do iG = 1, nlf
        do jG = 1, nlf
          read(u, '(2F15.10)') Qeff(iG, jG)
        end do
      end do
      close(u)
      call rpa_check(rpa_load_s0(ctx, c_loc(S0), c_loc(Qeff)), ctx, 'rpa_load_s0')
      call rpa_check(rpa_kk(ctx, c_null_ptr), ctx, 'rpa_kk')
      call rpa_check(rpa_epilogue(ctx, c_null_ptr, c_loc(pi11), c_loc(sig)), ctx, 'rpa_epilogue')
      open(newunit=u, file=trim(dato3))
      call open_trace(401, iq == qmin)
      call open_trace(402, iq == qmin)
      do io = 1, no - 1
        oi = (io - 1) * domega
        write(u, *) oi * hartree, pi11(io)
        if (io > 1) then
          write(401, *) oi * hartree, sig(io, 1)
          write(402, *) oi * hartree, sig(io, 2)
        end if
      end do
      close(u); close(401); close(402)
      write(*, '(a)') ' PROGRAM EXECUTION ENDED FOR CALC = 2'
    end if
  end do

  call rpa_check(rpa_ctx_destroy(ctx), c_null_ptr, 'rpa_ctx_destroy')

contains

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

  ! Info_{pol}.out (writeInfo, Optabs.f90:682-731); q in 2pi/a0 units
  subroutine write_info(qv)
    real(dp), intent(in) :: qv(3)
    integer :: w
    real(dp) :: absq, err
    absq = sqrt(sum(qv**2))
    open(newunit=w, file='Info_'//adjustl(trim(pol))//'.out')
    write(w, *) '***************General***********************'
    write(w, *) ' Currently we calculate         ---->', adjustl(trim(dato1))
    write(w, *) ' Currently we calculate         ---->', adjustl(trim(dato2))
    write(w, *) ''
    write(w, *) 'Number of point symmetry operation is', nsym
    write(w, '(A25,3F10.4,A5)') 'Wave vector (qx,qy,qz)=(', qv(1) * gcar, qv(2) * gcar, qv(3) * gcar, ') a.u.'
    write(w, '(A25,F7.3,A5)') '|(qx,qy,qz)|=', absq * gcar, 'a.u.'
    write(w, *) 'Local field effcts in '//trim(lf)//'-dir'
    write(w, *) 'Polarization in '//trim(pol)//'-dir'
    write(w, *) 'Number of local field vectors is', nlf
    write(w, *) 'Number of different K vectors in 1.B.Z. is', ntot
    write(w, *) 'Number of K vectors in I.B.Z. is', NkI
    write(w, *) 'Number of bands is               ', Nband
    write(w, '(A25,F7.3,A5)') 'eta_intra is  ', eta_intra * hartree * 1000.0_dp, 'meV'
    write(w, '(A25,F7.3,A5)') 'eta_inter is  ', eta_inter * hartree * 1000.0_dp, 'meV'
    write(w, '(A25,F7.3,A5)') 'Temperature is      ', T * hartree * 1000.0_dp, 'meV'
    write(w, *) ''
    write(w, *) '-Im(Chi(io,G1,G2))/pi is in file---->', adjustl(trim(dato1))
    write(w, *) ' Qeff complex matrix is in file ---->', adjustl(trim(dato2))
    write(w, *) ' Pi_munu is in file             ---->', adjustl(trim(dato3))
    write(w, *) ''
    write(w, *) '************* Checking 1BZ integration*******'
    write(w, *) ''
    write(w, '(A40,F8.4)') 'Number of electrons(1BZ integration)=', nel
    write(w, *) 'Number of electrons(unit cell)=', NelQE
    err = abs((NelQE - nel) / NelQE)
    write(w, '(A25,F7.3,A5)') 'Relative error=', err * 100.0_dp, '%'
    if (err > 0.05_dp) then
      write(w, *) 'WARRNING!!-1BZ INTEGRATION IS BAD!.'
      close(w)
      error stop 1
    end if
    close(w)
  end subroutine write_info

end program optabs_hip
