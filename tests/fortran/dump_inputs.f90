! Reads a run directory with the drivers' readers (module rpa_files) and
! writes what it read to dump.bin (stream access) for comparison with the
! arrays the synthetic inputs were written from.
! usage: dump_inputs.x <rundir> <savedir> <NkI> <Nband> <Nfill> <shift_Ha>
program dump_inputs
  use rpa_files
  implicit none
  character(len=256) :: a(6)
  integer :: i, nki, nband, nfill, nsym, ngm, ik, igwx, nbnd, u, npwx
  real(dp) :: shift, R(48, 3, 3), kc(3, 3), nel
  real(dp), allocatable :: ki(:, :), E(:, :)
  integer, allocatable :: mill(:, :), millk(:, :), k1(:)
  complex(dp), allocatable :: evc(:, :)
  do i = 1, 6
    call get_command_argument(i, a(i))
  end do
  read(a(3), *) nki
  read(a(4), *) nband
  read(a(5), *) nfill
  read(a(6), *) shift
  call point_r(trim(a(1))//'scf.out', nsym, R)
  call load_kc(trim(a(1))//'scf.out', kc)
  allocate(ki(3, nki), E(nki, nband))
  call load_ki_and_e(trim(a(1))//'synthetic.band', nki, nband, nfill, shift, ki, E)
  allocate(k1(nki))
  k1 = [(ik, ik = 1, nki)]
  call check_fbz_integration(nki, nband, nki, k1, 0.0_dp, E, nel)
  call load_g(trim(a(2)), ngm, mill)
  open(newunit=u, file='dump.bin', access='stream', form='unformatted', status='replace')
  write(u) nsym, R(1:nsym, :, :), kc, ki, E, nel, ngm, mill
  npwx = 0
  do ik = 1, nki
    call wfc_size(trim(a(2)), ik, igwx, nbnd)
    npwx = max(npwx, igwx)
  end do
  allocate(millk(3, npwx), evc(npwx, nband))
  do ik = 1, nki
    call load_cs_full(trim(a(2)), ik, nband, millk, evc, igwx)
    write(u) igwx, millk(:, 1:igwx), evc(1:igwx, :)
  end do
  close(u)
end program dump_inputs
