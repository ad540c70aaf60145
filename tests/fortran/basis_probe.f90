! Fortran host calling libimsrg_hip.so through the generated bind(C) module:
! builds the eMax-2 basis from an orbit table read on stdin (norb, then
! n l 2j 2tz hole per line) and prints the sizes; host-only entry points.
program basis_probe
  use, intrinsic :: iso_c_binding
  use imsrg_hip_c
  implicit none
  integer(c_int) :: norb, i, rc
  integer(c_int), allocatable :: n(:), l(:), jj(:), tz(:), hole(:)
  type(c_ptr) :: b
  read(*, *) norb
  allocate(n(norb), l(norb), jj(norb), tz(norb), hole(norb))
  do i = 1, norb
    read(*, *) n(i), l(i), jj(i), tz(i), hole(i)
  end do
  rc = imsrg_basis_create(norb, n, l, jj, tz, hole, b)
  if (rc /= 0) stop 'imsrg_basis_create failed'
  write(*, '(A,I0)') 'OP ', imsrg_basis_size(b, 5_c_int)
  write(*, '(A,I0)') 'NBLOCKS ', imsrg_basis_size(b, 4_c_int)
  write(*, '(A,I0)') 'WCC ', imsrg_basis_size(b, 8_c_int)
  rc = imsrg_set_path(b, 7_c_int)
  write(*, '(A,I0)') 'BADPATH ', rc
  rc = imsrg_basis_destroy(b)
end program basis_probe
