! Fortran driver of the drop-in modules (mod_global, mod_verlet, mod_pair and
! the emission modules of this directory).
!
!   rd_fortran_tests
!       the reference's own test scenarios (src/mod_tests.F90) through the
!       drop-in routines; exits with `error stop 1` on failure
!   rd_fortran_tests mainloop <fe|gtf|tip> <steps> <fp64|mixed>
!       the reference's main loop (src/main.F90:175-219) unchanged in shape:
!         call ptr_Do_Emission(i); call Update_Position(i); call Remove_Particles(i)
!       after the Init_* routine of the emission model, writing each step's
!       record (rd_step_record) to records.bin and the final particles
!       (n, then positions, velocities, ids) to state.bin in the current
!       directory, for comparison with the oracle.
program rd_fortran_tests
  use, intrinsic :: iso_c_binding
  use mod_global
  use rumdeed_hip
  use rd_device
  use mod_verlet
  use mod_pair
  use mod_field_emission_v2
  use mod_field_thermo_emission
  use mod_emission_tip
  implicit none
  double precision, parameter :: t_mu0 = 1.25663706212d-6, t_c = 299792458.0d0
  double precision, parameter :: t_eps0 = 1.0d0/(t_mu0 * t_c**2)
  double precision, parameter :: t_mu = 1.66053906660d-27
  double precision, parameter :: t_mion = 28.0134d0*t_mu - m_0
  integer :: nfail = 0
  character(len=32) :: arg1, arg2, arg3, arg4

  call get_command_argument(1, arg1)
  if (trim(arg1) == 'mainloop') then
    call get_command_argument(2, arg2)
    call get_command_argument(3, arg3)
    call get_command_argument(4, arg4)
    call main_loop(trim(arg2), arg3, trim(arg4))
    stop
  end if
  call test_three_body()
  call test_beeman()
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

  ! The particle arrays, as src/main.F90:422-456 allocates them (with a
  ! capacity of `cap` instead of MAX_PARTICLES).
  subroutine allocate_arrays(cap)
    integer, intent(in) :: cap
    if (allocated(particles_cur_pos)) then
      deallocate(particles_cur_pos, particles_prev_pos, particles_cur_vel, particles_cur_accel)
      deallocate(particles_prev_accel, particles_prev2_accel, particles_charge, particles_mass)
      deallocate(particles_species, particles_step, particles_emitter, particles_section)
      deallocate(particles_id)
    end if
    allocate(particles_cur_pos(1:3, 1:cap), particles_prev_pos(1:3, 1:cap))
    allocate(particles_cur_vel(1:3, 1:cap), particles_cur_accel(1:3, 1:cap))
    allocate(particles_prev_accel(1:3, 1:cap), particles_prev2_accel(1:3, 1:cap))
    allocate(particles_charge(1:cap), particles_mass(1:cap), particles_species(1:cap))
    allocate(particles_step(1:cap), particles_emitter(1:cap), particles_section(1:cap))
    allocate(particles_id(1:cap))
    particles_cur_accel = 0.0d0
    if (.not. allocated(emitters_pos)) then
      allocate(emitters_pos(1:3, 1:MAX_EMITTERS), emitters_dim(1:3, 1:MAX_EMITTERS))
      allocate(emitters_delay(1:MAX_EMITTERS))
    end if
    emitters_delay = 0
    nrEmit = 1
  end subroutine allocate_arrays

  subroutine planar_input(V, gap, ic, nic, dt)
    double precision, intent(in) :: V, gap, dt
    logical, intent(in) :: ic
    integer, intent(in) :: nic
    V_s = V
    d = gap
    box_dim = (/ 0.0d0, 0.0d0, gap /)
    image_charge = ic
    N_ic_max = nic
    time_step = dt
    time_step2 = dt*dt
  end subroutine planar_input

  ! Test_Acceleration_Without_Image_Charge (src/mod_tests.F90:405-519)
  subroutine test_three_body()
    double precision :: R(3, 3), q(3), m(3), a_exp(3, 3), E(3), probe(3), f(3), k
    integer :: i, j, sp(3)
    call allocate_arrays(16)
    call planar_input(2.0d0, 100.0d0*length_scale, .false., 0, 0.25d-15)
    call RD_Open(RD_GEOM_PLANAR)
    R(:, 1) = (/ 3.0d0, -10.0d0, 2.0d0 /) * length_scale
    R(:, 2) = (/ -9.0d0, 26.0d0, 80.0d0 /) * length_scale
    R(:, 3) = (/ 6.0d0, -24.0d0, 56.53d0 /) * length_scale
    q = (/ -q_0, -q_0, q_0 /)
    m = (/ m_0, m_0, t_mion /)
    sp = (/ species_elec, species_elec, species_ion /)
    do i = 1, 3
      call Add_Particle(R(:, i), (/ 0.0d0, 0.0d0, 0.0d0 /), sp(i), 1, 1, -1)
    end do
    call check(nrPart == 3 .and. nrElec == 2 .and. nrIon == 1, 'Add_Particle counters')
    particles_cur_accel = 0.0d0
    call Calculate_Acceleration_Particles()
    k = 1.0d0/(4.0d0*pi*t_eps0)
    E = (/ 0.0d0, 0.0d0, -2.0d0/(100.0d0*length_scale) /)
    do i = 1, 3
      a_exp(:, i) = q(i)*E/m(i)
      do j = 1, 3
        if (i /= j) a_exp(:, i) = a_exp(:, i) + k*q(i)*q(j)/m(i)*(R(:, i) - R(:, j))/norm2(R(:, i) - R(:, j))**3
      end do
      call check(close_vec(particles_cur_accel(:, i), a_exp(:, i), 1.0d-9), 'three-body acceleration')
    end do
    probe = (/ -4.55, -2.34, 96.44 /) * length_scale
    f = Calc_Field_at(probe)
    call check(close_vec(f, (/ -314559.29097098d0, 1423979.07058996d0, -20246038.87978313d0 /), &
               0.02d0), 'field at probe vs the reference test value')
    f = Calc_Field_at(probe, R(:, 1), .false.)
    call check(close_vec(f, Calc_Field_at(probe), 0.0d0), 'optional arguments of Calc_Field_at')
    call RD_Close()
  end subroutine test_three_body

  ! Test_Beeman_Kinematics (src/mod_tests.F90:1403-1452)
  subroutine test_beeman()
    double precision :: a_z, V, dt
    integer :: i
    V = 2.0d3
    dt = 0.25d-15
    call allocate_arrays(16)
    call planar_input(V, 1000.0d0*length_scale, .false., 0, dt)
    call RD_Open(RD_GEOM_PLANAR)
    call Add_Particle((/ 0.0d0, 0.0d0, 500.0d0*length_scale /), (/ 1.0d3, 0.0d0, 0.0d0 /), &
                      species_elec, 1, 1, -1)
    do i = 1, 3
      call Update_Position(i)
      call Remove_Particles(i)
    end do
    call RD_Sync_To_Host()
    a_z = q_0*V/(m_0*d)
    call check(abs((particles_cur_pos(3, 1) - 500.0d0*length_scale) - 4.5d0*a_z*dt**2) &
               < 1.0d-9*4.5d0*a_z*dt**2, 'Beeman z displacement')
    call check(abs(particles_cur_vel(3, 1) - 3.0d0*a_z*dt) < 1.0d-9*3.0d0*a_z*dt, 'Beeman z velocity')
    call check(abs(particles_cur_pos(1, 1) - 3.0d3*dt) < 1.0d-9*3.0d3*dt, 'Beeman x displacement')
    call check(abs(ramo_current(species_elec) - q_0*3.0d0*a_z*dt/d) < 1.0d-9*q_0*3.0d0*a_z*dt/d, &
               'Ramo current')
    call check(abs(cur_time - 3*dt/time_scale) < 1.0d-15, 'cur_time')
    call RD_Close()
  end subroutine test_beeman

  subroutine write_work(table, nx, ny)
    integer, intent(in) :: nx, ny
    double precision, intent(in) :: table(nx, ny)
    integer :: u, i
    open(newunit=u, file='work', status='REPLACE', action='WRITE')
    write(u, '(i0)') 1
    write(u, '(i0, 1x, i0)') ny, nx
    do i = 1, ny
      write(u, '(*(es24.16, 1x))') table(:, i)
    end do
    close(u)
  end subroutine write_work

  ! The reference's main loop for one emission model; see the header.
  subroutine main_loop(model, nsteps_arg, prec)
    character(len=*), intent(in) :: model, nsteps_arg, prec
    type(rd_step_record) :: rec
    integer :: i, u, nsteps, emitted, absorbed
    double precision :: wf(2, 1)
    integer(c_int), allocatable :: ints(:, :)
    read(nsteps_arg, *) nsteps
    call allocate_arrays(65536)
    call RD_Use_Seed(5_c_int64_t)
    if (prec == 'mixed') call RD_Set_Precision(RD_PREC_MIXED)
    select case (model)
    case ('fe')
      call planar_input(1600.0d0, 200.0d0*length_scale, .true., 1, 0.5d-15)
      emitters_pos(:, 1) = (/ -50.0d0, -50.0d0, 0.0d0 /) * length_scale
      emitters_dim(:, 1) = (/ 100.0d0, 100.0d0, 0.0d0 /) * length_scale
      wf(1, 1) = 4.7d0
      call write_work(wf(1:1, :), 1, 1)
      call Init_Field_Emission_v2()
    case ('gtf')
      call planar_input(6000.0d0, 1000.0d0*length_scale, .true., 0, 0.25d-15)
      emitters_pos(:, 1) = (/ -200.0d0, -200.0d0, 0.0d0 /) * length_scale
      emitters_dim(:, 1) = (/ 400.0d0, 400.0d0, 0.0d0 /) * length_scale
      T_temp = 1500.0d0
      wf(:, 1) = (/ 4.0d0, 4.5d0 /)
      call write_work(wf, 2, 1)
      call Init_Field_Thermo_Emission()
    case ('tip')
      V_s = 1000.0d0
      image_charge = .true.
      N_ic_max = 0
      time_step = 0.25d-15
      time_step2 = time_step**2
      emitters_dim(:, 1) = (/ 1000.0d0, 250.0d0, 500.0d0 /) * length_scale
      d = emitters_dim(1, 1) + emitters_dim(3, 1)
      box_dim = (/ 0.0d0, 0.0d0, 1000.0d0*length_scale /)
      call Init_Emission_Tip()
    case default
      print '(a, a)', 'unknown model ', model
      error stop 2
    end select
    open(newunit=u, file='records.bin', status='REPLACE', access='STREAM', form='UNFORMATTED')
    emitted = 0
    absorbed = 0
    do i = 1, nsteps
      call ptr_Do_Emission(i)
      call Update_Position(i)
      call Remove_Particles(i)
      call RD_Last_Record(rec)
      write(u) rec
      emitted = emitted + rec%emitted
      absorbed = absorbed + rec%absorbed
    end do
    close(u)
    call RD_Sync_To_Host()
    open(newunit=u, file='state.bin', status='REPLACE', access='STREAM', form='UNFORMATTED')
    write(u) int(nrPart, c_int)
    write(u) particles_cur_pos(:, 1:nrPart), particles_cur_vel(:, 1:nrPart)
    allocate(ints(1, max(nrPart, 1)))
    ints(1, 1:nrPart) = int(particles_id(1:nrPart), c_int)
    write(u) ints(1, 1:nrPart)
    close(u)
    print '(a, a, a, i0, a, i0, a, i0)', 'main loop ', model, ': emitted ', emitted, &
          ' absorbed ', absorbed, ' in flight ', nrPart
    if (nrPart /= emitted - absorbed) then
      print '(a)', 'FAIL charge bookkeeping'
      error stop 1
    end if
  end subroutine main_loop
end program rd_fortran_tests
