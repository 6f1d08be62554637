! ISO_C_BINDING interface to librpa_hip.so (include/rpa_hip.h).
!
! The reference programs call their hot path as contained subroutines
! (Chi_RPA/Optabs.f90, W/Sloss.f90); the drivers of this package call the
! C ABI below instead.  Every function returns 0 on success; rpa_check turns
! a nonzero code into the reference's fail-stop (notifications.f90:20-37:
! message, then error stop).
module rpa_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: RPA_OK = 0, RPA_E_SINGULAR = 6
  integer(c_int), parameter :: RPA_SLOSS = 0, RPA_OPTABS = 1
  integer(c_int), parameter :: RPA_FP64 = 0, RPA_MIXED = 1
  integer(c_int), parameter :: RPA_MAP_DIRECT = 0, RPA_MAP_IGK = 1
  integer(c_int), parameter :: RPA_POL_X = 1, RPA_POL_Y = 2, RPA_POL_Z = 3, RPA_POL_YZ = 4
  integer(c_int), parameter :: RPA_RES_S0 = 0, RPA_RES_QEFF = 1, RPA_RES_CHI0 = 2, &
                               RPA_RES_V = 3, RPA_RES_EPSINV = 4, RPA_RES_CHI = 5, &
                               RPA_RES_W = 6, RPA_RES_LOSS = 7, RPA_RES_PI11 = 8, &
                               RPA_RES_SIG = 9

  type, bind(C) :: rpa_dims_t
    integer(c_int) :: nki, nband, npwx, nlfd, no, program, precision, device
  end type rpa_dims_t

  type, bind(C) :: rpa_params_t
    integer(c_int) :: nval, vertex_map, ngd, pol
    real(c_double) :: eps, gcar, domega, eta, lor_cut, df_cut, efermi, temp, vcell, c0, eta_intra
  end type rpa_params_t

  type, bind(C) :: rpa_fock_params_t
    integer(c_int) :: nk, nlf, nq, ntot, device, pad
    real(c_double) :: gcar, qmin, c0, vcell
  end type rpa_fock_params_t

  type, bind(C) :: rpa_ladder_params_t
    integer(c_int) :: nk, no, ntot, device
    real(c_double) :: omin, domega, eta, vcell, valley
  end type rpa_ladder_params_t

  type, bind(C) :: rpa_photon_params_t
    integer(c_int) :: no, nlf, nline, device
    real(c_double) :: omin, domega, eta, c0
  end type rpa_photon_params_t

  type, bind(C) :: rpa_interp_params_t
    integer(c_int) :: no, nlf, no_interp, all, device
    real(c_double) :: domega, domega_tot, eta_intra
  end type rpa_interp_params_t

  interface
    integer(c_int) function rpa_ctx_create(dims, ctx) bind(C, name="rpa_ctx_create")
      import :: c_int, c_ptr, rpa_dims_t
      type(rpa_dims_t), intent(in) :: dims
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function rpa_ctx_destroy(ctx) bind(C, name="rpa_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function rpa_last_error_c(ctx) bind(C, name="rpa_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    real(c_double) function rpa_last_run_ms(ctx) bind(C, name="rpa_last_run_ms")
      import :: c_double, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rpa_set_gvectors(ctx, ngm, mill, kc) bind(C, name="rpa_set_gvectors")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngm
      integer(c_int), intent(in) :: mill(3, *)
      real(c_double), intent(in) :: kc(3, 3)
    end function
    integer(c_int) function rpa_upload_basis(ctx, ik, ngk, mill, evc) bind(C, name="rpa_upload_basis")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ik, ngk
      integer(c_int), intent(in) :: mill(3, *)
      complex(c_double_complex), intent(in) :: evc(*)
    end function
    integer(c_int) function rpa_set_bands(ctx, E) bind(C, name="rpa_set_bands")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: E(*)
    end function
    integer(c_int) function rpa_set_local_fields(ctx, nlf, lf_idx, parg) bind(C, name="rpa_set_local_fields")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nlf
      integer(c_int), intent(in) :: lf_idx(*), parg(*)
    end function
    integer(c_int) function rpa_set_params(ctx, p) bind(C, name="rpa_set_params")
      import :: c_int, c_ptr, rpa_params_t
      type(c_ptr), value :: ctx
      type(rpa_params_t), intent(in) :: p
    end function
    integer(c_int) function rpa_set_q(ctx, ntot, K1, K2, iG0, ktot, q) bind(C, name="rpa_set_q")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ntot
      integer(c_int), intent(in) :: K1(*), K2(*), iG0(*)
      real(c_double), intent(in) :: ktot(3, *), q(3)
    end function
    integer(c_int) function rpa_set_q_sym(ctx, ntot, K1, K2, iG0, R1, R2, ktot, q) bind(C, name="rpa_set_q_sym")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ntot
      integer(c_int), intent(in) :: K1(*), K2(*), iG0(*), R1(*), R2(*)
      real(c_double), intent(in) :: ktot(3, *), q(3)
    end function
    integer(c_int) function rpa_set_symmetry(ctx, nsym, R) bind(C, name="rpa_set_symmetry")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsym
      real(c_double), intent(in) :: R(48, 3, 3)
    end function
    integer(c_int) function rpa_run(ctx) bind(C, name="rpa_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rpa_run_graph(ctx) bind(C, name="rpa_run_graph")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rpa_get_result(ctx, which, out) bind(C, name="rpa_get_result")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      type(c_ptr), value :: out
    end function
    integer(c_int) function rpa_load_s0(ctx, S0, qeff) bind(C, name="rpa_load_s0")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, S0, qeff
    end function
    integer(c_int) function rpa_kk(ctx, chi0_out) bind(C, name="rpa_kk")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, chi0_out
    end function
    integer(c_int) function rpa_epilogue(ctx, loss, pi11, sig) bind(C, name="rpa_epilogue")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, loss, pi11, sig
    end function
    integer(c_int) function rpa_w_kk(ctx, w1, w2) bind(C, name="rpa_w_kk")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: w1(*), w2(*)
    end function
    integer(c_int) function rpa_gen_fbz(nki, nsym, eps, ki, R, ntot, ktot) bind(C, name="rpa_gen_fbz")
      import :: c_int, c_double
      integer(c_int), value :: nki, nsym
      real(c_double), value :: eps
      real(c_double), intent(in) :: ki(3, *), R(48, 3, 3)
      integer(c_int), intent(out) :: ntot
      real(c_double), intent(out) :: ktot(3, *)
    end function
    integer(c_int) function rpa_find_min_q(ntot, ktot, iq, q) bind(C, name="rpa_find_min_q")
      import :: c_int, c_double
      integer(c_int), value :: ntot, iq
      real(c_double), intent(in) :: ktot(3, *)
      real(c_double), intent(out) :: q(3)
    end function
    integer(c_int) function rpa_find_kq(ntot, ktot, nki, ki, ng, g, q, eps, K1, K2, iG0) &
        bind(C, name="rpa_find_kq")
      import :: c_int, c_double
      integer(c_int), value :: ntot, nki, ng
      real(c_double), intent(in) :: ktot(3, *), ki(3, *), g(3, *), q(3)
      real(c_double), value :: eps
      integer(c_int), intent(out) :: K1(*), K2(*), iG0(*)
    end function
    integer(c_int) function rpa_find_kq_sym(ntot, ktot, nki, ki, ng, g, q, nsym, RI, eps, R1, K1, R2, K2, iG0) &
        bind(C, name="rpa_find_kq_sym")
      import :: c_int, c_double
      integer(c_int), value :: ntot, nki, ng, nsym
      real(c_double), intent(in) :: ktot(3, *), ki(3, *), g(3, *), q(3), RI(48, 3, 3)
      real(c_double), value :: eps
      integer(c_int), intent(out) :: R1(*), K1(*), R2(*), K2(*), iG0(*)
    end function
    integer(c_int) function rpa_point_inverse(nsym, R, RI) bind(C, name="rpa_point_inverse")
      import :: c_int, c_double
      integer(c_int), value :: nsym
      real(c_double), intent(in) :: R(48, 3, 3)
      real(c_double), intent(out) :: RI(48, 3, 3)
    end function
    integer(c_int) function rpa_gen_glf(lf, ecut, gcar, ng, g, mill, nlfd, nlf, glf, parg, lf_idx) &
        bind(C, name="rpa_gen_glf")
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: lf(*)
      real(c_double), value :: ecut, gcar
      integer(c_int), value :: ng, nlfd
      real(c_double), intent(in) :: g(3, *)
      integer(c_int), intent(in) :: mill(3, *)
      integer(c_int), intent(out) :: nlf
      real(c_double), intent(out) :: glf(3, *)
      integer(c_int), intent(out) :: parg(*), lf_idx(*)
    end function
    integer(c_int) function rpa_gcart(n, mill, kc, g) bind(C, name="rpa_gcart")
      import :: c_int, c_double
      integer(c_int), value :: n
      integer(c_int), intent(in) :: mill(3, *)
      real(c_double), intent(in) :: kc(3, 3)
      real(c_double), intent(out) :: g(3, *)
    end function
    integer(c_long) function c_strlen(s) bind(C, name="strlen")
      import :: c_long, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function rpa_interpolate(p, S0, qeff, pi_dense, pi_intra) bind(C, name="rpa_interpolate")
      import :: c_int, c_ptr, rpa_interp_params_t
      type(rpa_interp_params_t), intent(in) :: p
      type(c_ptr), value :: S0, qeff, pi_dense, pi_intra
    end function
    integer(c_int) function rpa_gamma_w(ctx, gammap, gammam) bind(C, name="rpa_gamma_w")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, gammap, gammam
    end function
    integer(c_int) function rpa_gamma_propagator(no, nser, domega, D, gammap, gammam, kks, skk, device) &
        bind(C, name="rpa_gamma_propagator")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: no, nser, device
      real(c_double), value :: domega
      type(c_ptr), value :: D, gammap, gammam
      real(c_double), intent(out) :: kks(*), skk(*)
    end function
    integer(c_int) function rpa_fock_kernel(p, kbz, wt, mnm, fockk) bind(C, name="rpa_fock_kernel")
      import :: c_int, c_double, c_ptr, rpa_fock_params_t
      type(rpa_fock_params_t), intent(in) :: p
      real(c_double), intent(in) :: kbz(3, *)
      type(c_ptr), value :: wt, mnm, fockk
    end function
    integer(c_int) function rpa_ladder(p, fockk, delta, current, chi_ladd, chi_para) bind(C, name="rpa_ladder")
      import :: c_int, c_double, c_ptr, rpa_ladder_params_t
      type(rpa_ladder_params_t), intent(in) :: p
      real(c_double), intent(in) :: delta(*)
      type(c_ptr), value :: fockk, current, chi_ladd, chi_para
    end function
    integer(c_int) function rpa_photon(p, glfz, parg, pixx0, piyy0, pizz0, piyz0, pizy0, q_line, theta_line, &
                                       qmode_line, pi11, psum, spectra, pi_full) bind(C, name="rpa_photon")
      import :: c_int, c_double, c_ptr, rpa_photon_params_t
      type(rpa_photon_params_t), intent(in) :: p
      real(c_double), intent(in) :: glfz(*), q_line(*), theta_line(*)
      integer(c_int), intent(in) :: parg(*), qmode_line(*)
      type(c_ptr), value :: pixx0, piyy0, pizz0, piyz0, pizy0, pi11, psum, spectra, pi_full
    end function
  end interface

contains

  ! Message of the last failed call on ctx.
  function rpa_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = rpa_last_error_c(ctx)
    if (.not. c_associated(p)) then
      msg = 'unknown error'
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function rpa_last_error

  ! Fail-stop on a nonzero return code, as notifications.error does.
  ! The &system entry precision = 'fp64' | 'mixed' as enum rpa_precision;
  ! anything else stops with a message.
  integer(c_int) function rpa_precision_of(name) result(pr)
    character(len=*), intent(in) :: name
    select case (trim(adjustl(name)))
    case ('fp64', 'FP64')
      pr = RPA_FP64
    case ('mixed', 'MIXED')
      pr = RPA_MIXED
    case default
      write(*, '(a)') 'ERROR: precision must be ''fp64'' or ''mixed'', got '''//trim(name)//''''
      error stop 1
    end select
  end function rpa_precision_of

  subroutine rpa_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc == RPA_OK) return
    if (c_associated(ctx)) then
      write(*, '(a)') 'ERROR: '//what//': '//rpa_last_error(ctx)
    else
      write(*, '(a,i0)') 'ERROR: '//what//' failed with code ', rc
    end if
    error stop 1
  end subroutine rpa_check

end module rpa_hip
