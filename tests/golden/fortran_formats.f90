! Writes the formatted records of the reference's output headers
! (pic_out.f90: fields_out '(4i14)', '(4e14.5)'; part_pdata_out
! '(a13,e11.4)', '(a12,i8)', '(4i20)'; list-directed strings) for a fixed
! set of values, so that the Python writers can be checked against what a
! Fortran compiler prints.  Built and run by make_formats.py.
program fortran_formats
  implicit none
  integer, parameter :: sp = kind(1.0), i8 = selected_int_kind(18)
  real(sp) :: r(20)
  integer :: n(20)
  integer(i8) :: big
  integer :: q
  r = [1.0_sp, -2.5e-3_sp, 0.0_sp, 123456.78_sp, 3.0e-30_sp, 0.999995_sp, 9.99995e5_sp, &
       -1.0e10_sp, 7.2e-8_sp, 1.0_sp/3.0_sp, 0.05_sp, -0.2_sp, 25.6_sp, 1.0e-3_sp, &
       2.81794033_sp, 0.510998928_sp, 11.1487_sp, -6.4_sp, 12345.0_sp, 123445.0_sp]
  n = [1, 1, 1, 256, 128, 128, 128, 128, 1, 0, 0, 1, 1, 1, 3, 8, 2, 2, 2, 1]
  big = 33554432_i8
  open (10, file='fortran_formats.txt', form='formatted')
  write (10, *) ' Integer parameters'
  write (10, '(4i14)') n
  write (10, *) ' Real parameters'
  write (10, '(4e14.5)') r
  do q = 1, 20
    write (10, '(a13,e11.4)') ' time =      ', r(q)
  end do
  do q = 1, 20
    write (10, '(a12,i8)') ' npe =      ', n(q)
  end do
  write (10, *) ' Number of particles in the output box'
  write (10, '(4i20)') big
  close (10)
end program
