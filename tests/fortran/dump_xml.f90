! Reads a data-file-schema.xml with the drivers' reader (module rpa_files)
! and writes what it read to dump_xml.bin (stream access).
! usage: dump_xml.x <path to data-file-schema.xml>
program dump_xml
  use rpa_files
  implicit none
  character(len=256) :: path
  type(qe_xml) :: x
  integer :: u, nsym
  real(dp) :: R(48, 3, 3), kc(3, 3)
  call get_command_argument(1, path)
  call load_xml_qe(trim(path), x)
  call xml_point_ops(x, nsym, R)
  call xml_kc(x, kc)
  open(newunit=u, file='dump_xml.bin', access='stream', form='unformatted', status='replace')
  write(u) x%nband, x%nki, x%nrot, x%nsym, x%nmp, x%ecutwfc, x%ecutrho, x%alat, x%temp, x%efermi, x%nelec
  write(u) x%a, x%b, x%R(:, :, 1:x%nrot), x%ki(:, 1:x%nki), x%npw(1:x%nki)
  write(u) x%eig(1:x%nband, 1:x%nki), x%occ(1:x%nband, 1:x%nki)
  write(u) nsym, R(1:nsym, :, :), kc
  close(u)
end program dump_xml
