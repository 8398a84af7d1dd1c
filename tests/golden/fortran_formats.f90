! Prints sample values with the edit descriptors of RUMDEED's output files
! (emitted.dt, field.dt, absorbed*.dt, ramo_current.dt, lifetime.dt,
! volt.dt, integration.dt), one
! record per line, so that tests/test_output_formats.py can compare the
! Python writer against a Fortran compiler's own formatting.
program fortran_formats
  implicit none
  double precision :: x(12)
  integer :: i
  x = [1.25d-15, 0.0d0, -3.5d-3, 7.123456789d5, 9.99999951d-1, 1.0d0, &
       & -2.0d-300, 6.02214076d23, 1.0d100, 4.9999999d-5, 123.456d0, -0.0d0]
  do i = 1, size(x)
    write (*, "(E14.6, *(tr8, i6))") x(i), i, 10*i, -i
    write (*, "(i8, tr2, E16.8, tr2, E16.8, tr2, E16.8)") i, x(i), -x(i), x(i)*3.0d0
    write (*, "(ES12.4, *(tr2, i8))") x(i), i, 1000*i, 0
    write (*, "(ES12.4, tr2, i6, tr2, i6)") x(i), i, 7*i
    ! volt.dt: four items for three descriptors (format reversion)
    write (*, "(ES12.4, tr2, i8, tr2, ES18.8)") x(i), i, 2.0d0*x(i) + 1.0d0, 0.0d0
    ! integration.dt
    write (*, "(i3, tr2, i8, tr2, i8, tr2, i4, tr2, ES12.4, tr2, ES12.4, tr2, ES12.4)") &
      & i, 1024*i, 1280*i, mod(i, 2), x(i), -x(i), 0.0d0
  end do
end program fortran_formats
