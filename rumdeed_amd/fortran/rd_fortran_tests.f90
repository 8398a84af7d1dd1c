! Fortran driver of the drop-in layer: the reference's own test scenarios
! (src/mod_tests.F90) run through rumdeed_dropin, the way the reference's
! EMISSION_MODE = 0 binary runs them.  Exits with `error stop 1` on failure.
program rd_fortran_tests
  use, intrinsic :: iso_c_binding
  use rumdeed_hip
  use rumdeed_dropin
  implicit none
  double precision, parameter :: pi = 3.141592653589793238462643383279502884197169399375105820974944592307816406286d0
  double precision, parameter :: mu_0 = 1.25663706212d-6, c = 299792458.0d0
  double precision, parameter :: epsilon_0 = 1.0d0/(mu_0 * c**2)
  double precision, parameter :: q_0 = 1.602176634d-19, m_0 = 9.1093837015d-31
  double precision, parameter :: m_u = 1.66053906660d-27
  double precision, parameter :: m_N2p = 28.0134d0*m_u - m_0
  double precision, parameter :: length_scale = 1.0d-9
  integer :: nfail = 0

  call test_three_body()
  call test_beeman()
  call test_main_loop()
  if (nfail > 0) then
    print '(a, i0)', 'FAILED: ', nfail
    error stop 1
  end if
  print '(a)', 'ALL PASSED'

contains

  subroutine check(ok, name)
    logical, intent(in) :: ok
    character(len=*), intent(in) :: name
    if (ok) then
      print '(a, a)', 'PASS ', name
    else
      print '(a, a)', 'FAIL ', name
      nfail = nfail + 1
    end if
  end subroutine check

  logical function close_vec(a, b, rtol)
    double precision, intent(in) :: a(3), b(3), rtol
    close_vec = norm2(a - b) <= rtol * norm2(b)
  end function close_vec

  ! Test_Acceleration_Without_Image_Charge (src/mod_tests.F90:405-519)
  subroutine test_three_body()
    type(rd_geometry) :: g
    double precision :: R(3, 3), q(3), m(3), acc(3, 3), a_exp(3, 3), E(3), probe(3), f(3)
    double precision :: k
    integer :: i, j, sp(3)
    g = rd_geometry(RD_GEOM_PLANAR, 0, 0, 0, 2.0d0, 100.0d0*length_scale, 0d0, 0d0, 0d0, 0d0)
    call RD_Init(16, g, 0.25d-15)
    R(:, 1) = (/ 3.0d0, -10.0d0, 2.0d0 /) * length_scale
    R(:, 2) = (/ -9.0d0, 26.0d0, 80.0d0 /) * length_scale
    R(:, 3) = (/ 6.0d0, -24.0d0, 56.53d0 /) * length_scale
    q = (/ -q_0, -q_0, q_0 /)
    m = (/ m_0, m_0, m_N2p /)
    sp = (/ 1, 1, 2 /)
    acc = 0.0d0
    call Calculate_Acceleration_Particles(3, R, q, m, sp, acc)
    k = 1.0d0/(4.0d0*pi*epsilon_0)
    E = (/ 0.0d0, 0.0d0, -2.0d0/(100.0d0*length_scale) /)
    do i = 1, 3
      a_exp(:, i) = q(i)*E/m(i)
      do j = 1, 3
        if (i /= j) a_exp(:, i) = a_exp(:, i) + k*q(i)*q(j)/m(i)*(R(:, i) - R(:, j))/norm2(R(:, i) - R(:, j))**3
      end do
      call check(close_vec(acc(:, i), a_exp(:, i), 1.0d-9), 'three-body acceleration')
    end do
    probe = (/ -4.55, -2.34, 96.44 /) * length_scale
    f = Calc_Field_at(probe)
    call check(close_vec(f, (/ -314559.29097098d0, 1423979.07058996d0, -20246038.87978313d0 /), &
               0.02d0), 'field at probe vs the reference test value')
    call RD_Finalize()
  end subroutine test_three_body

  ! Test_Beeman_Kinematics (src/mod_tests.F90:1403-1452)
  subroutine test_beeman()
    type(rd_geometry) :: g
    type(rd_step_record) :: rec, rec2
    double precision :: pos(3, 1), prev(3, 1), vel(3, 1), a0(3, 1), a1(3, 1), a2(3, 1)
    integer(c_int) :: ints(5, 1)
    double precision :: d, V, dt, a_z
    integer :: i
    d = 1000.0d0*length_scale
    V = 2.0d3
    dt = 0.25d-15
    g = rd_geometry(RD_GEOM_PLANAR, 0, 0, 0, V, d, 0d0, 0d0, 0d0, 0d0)
    call RD_Init(16, g, dt)
    call Add_Particle((/ 0.0d0, 0.0d0, 500.0d0*length_scale /), (/ 1.0d3, 0.0d0, 0.0d0 /), 1, 1, 1, -1)
    do i = 1, 3
      call Update_Position(i, rec)
      call Remove_Particles(i, rec2)
    end do
    call rd_check(rd_get_state(rd_ctx, pos, prev, vel, a0, a1, a2, ints), 'rd_get_state')
    a_z = q_0*V/(m_0*d)
    call check(abs((pos(3, 1) - 500.0d0*length_scale) - 4.5d0*a_z*dt**2) < 1.0d-9*4.5d0*a_z*dt**2, &
               'Beeman z displacement')
    call check(abs(vel(3, 1) - 3.0d0*a_z*dt) < 1.0d-9*3.0d0*a_z*dt, 'Beeman z velocity')
    call check(abs(rec%ramo(1) - q_0*3.0d0*a_z*dt/d) < 1.0d-9*q_0*3.0d0*a_z*dt/d, 'Ramo current')
    call RD_Finalize()
  end subroutine test_beeman

  ! The main loop (src/main.F90:175-219) for a planar field-emission diode.
  subroutine test_main_loop()
    type(rd_geometry) :: g
    type(rd_emitter), target :: em
    type(rd_step_record) :: r1, r2, r3
    double precision, target :: wf(1)
    integer :: i, emitted, absorbed
    g = rd_geometry(RD_GEOM_PLANAR, 1, 1, 0, 1600.0d0, 200.0d0*length_scale, 0d0, 0d0, 0d0, 0d0)
    call RD_Init(4096, g, 0.5d-15)
    call rd_check(rd_set_seed(rd_ctx, 5_c_int64_t, 0_c_int64_t), 'rd_set_seed')
    wf(1) = 4.7d0
    em = rd_emitter(RD_EMIT_PLANAR_FE, 0, -50.0d0*length_scale, -50.0d0*length_scale, &
                    100.0d0*length_scale, 100.0d0*length_scale, 300.0d0, 4.7d0, 1, 1, &
                    c_loc(wf), 0, 0)
    call rd_check(rd_set_emitter(rd_ctx, em), 'rd_set_emitter')
    emitted = 0
    absorbed = 0
    do i = 1, 45
      call Do_Emission(i, r1)
      call Update_Position(i, r2)
      call Remove_Particles(i, r3)
      emitted = emitted + r1%emitted
      absorbed = absorbed + r3%absorbed
    end do
    print '(a, i0, a, i0, a, i0)', 'main loop: emitted ', emitted, ' absorbed ', absorbed, &
          ' in flight ', RD_Count()
    call check(emitted > 100 .and. absorbed > 0, 'main loop emits and absorbs')
    call check(RD_Count() == emitted - absorbed, 'charge bookkeeping')
    call RD_Finalize()
  end subroutine test_main_loop
end program rd_fortran_tests
