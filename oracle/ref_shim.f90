! bind(C) entry points over the compile-clean reference modules, so the
! oracle tests can call the reference's own routines (KK stencils, gjel,
! qe_gksort, local_field).  Compiled together with the reference sources
! where they lie; the result goes to oracle/_ref/ only.  Test infrastructure.
module ref_shim
  use iso_c_binding
  use iso_fortran_env, only: dp => real64
  use response_function_KK, only: genReChi0, genImChi0
  use qe_gksort, only: init_igk
  use local_field, only: genGlf
  implicit none
  private
  public :: ref_kk_cmplx, ref_kk_real, ref_gjel, ref_init_igk, ref_genglf

  interface
    subroutine gjel(a, n, np, b, m, mp)
      integer :: n, np, m, mp
      double complex :: a(np, np), b(np, mp)
    end subroutine gjel
  end interface

contains

  ! Modules/response_function_KK.f90:37-71 (complex S0(-No:No,Nlf,Nlf))
  subroutine ref_kk_cmplx(io, no, nlf, ig, jg, oi, domega, s0, re, im) bind(C, name="ref_kk_cmplx")
    integer(c_int), value :: io, no, nlf, ig, jg
    real(c_double), value :: oi, domega
    complex(c_double_complex), intent(in) :: s0(-no:no, nlf, nlf)
    real(c_double), intent(out) :: re, im
    real(dp) :: r, i
    call genReChi0(io, no, nlf, ig, jg, oi, domega, s0, r)
    call genImChi0(io, no, nlf, ig, jg, oi, domega, s0, i)
    re = r
    im = i
  end subroutine ref_kk_cmplx

  ! Modules/response_function_KK.f90:25-59 (real S0(-No:No,Nlf,Nlf))
  subroutine ref_kk_real(io, no, nlf, ig, jg, oi, domega, s0, re, im) bind(C, name="ref_kk_real")
    integer(c_int), value :: io, no, nlf, ig, jg
    real(c_double), value :: oi, domega
    real(c_double), intent(in) :: s0(-no:no, nlf, nlf)
    real(c_double), intent(out) :: re, im
    real(dp) :: r, i
    call genReChi0(io, no, nlf, ig, jg, oi, domega, s0, r)
    call genImChi0(io, no, nlf, ig, jg, oi, domega, s0, i)
    re = r
    im = i
  end subroutine ref_kk_real

  ! Chi_RPA/gjel.f:1-91
  subroutine ref_gjel(a, n, np, b, m, mp) bind(C, name="ref_gjel")
    integer(c_int), value :: n, np, m, mp
    complex(c_double_complex), intent(inout) :: a(np, np), b(np, mp)
    call gjel(a, n, np, b, m, mp)
  end subroutine ref_gjel

  ! Modules/qe_gksort.f90:11-42
  subroutine ref_init_igk(nks, xk, npwx, ngm, g, gcutw, igk) bind(C, name="ref_init_igk")
    integer(c_int), value :: nks, npwx, ngm
    real(c_double), intent(in) :: xk(3, nks), g(3, ngm)
    real(c_double), value :: gcutw
    integer(c_int), intent(out) :: igk(npwx, nks)
    integer, allocatable :: igk_k(:, :)
    call init_igk(nks, xk, npwx, ngm, g, gcutw, igk_k)
    igk = igk_k
  end subroutine ref_init_igk

  ! Modules/local_field.f90:11-94; parity input is the Miller z index as the
  ! reference's loadG stores it (Modules/io_qe.f90:64).
  subroutine ref_genglf(lfcode, ecut, gcar, ng, g, millz, nlfmax, nlf, glf, parg) bind(C, name="ref_genglf")
    integer(c_int), value :: lfcode, ng, nlfmax
    real(c_double), value :: ecut, gcar
    real(c_double), intent(in) :: g(3, ng)
    integer(c_int), intent(in) :: millz(ng)
    integer(c_int), intent(out) :: nlf
    real(c_double), intent(out) :: glf(3, nlfmax)
    integer(c_int), intent(out) :: parg(nlfmax)
    real(dp), allocatable :: glf_a(:, :)
    integer, allocatable :: p(:)
    integer :: n
    character(len=3) :: lf
    if (lfcode == 1) then
      lf = 'z'
    else
      lf = 'xyz'
    end if
    allocate(p(ng))
    p = millz
    call genGlf(lf, ecut, gcar, g, n, glf_a, p)
    nlf = n
    n = min(n, nlfmax)
    glf(:, 1:n) = glf_a(:, 1:n)
    parg(1:n) = p(1:n)
  end subroutine ref_genglf

end module ref_shim
