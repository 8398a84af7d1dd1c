! ORACLE -- test infrastructure only.
!
! A thin bind(C) harness over the RUMDEED reference modules (compiled from the
! untouched sources under the reference tree by oracle/Makefile into
! oracle/_ref/).  It lets the tests and the golden-vector script call the
! reference's own routines:
!   Calculate_Acceleration_Particles_Planar / _Generic / _ACC  (src/mod_verlet.F90:625-884, :1217-1429)
!   Calc_Field_at                                              (src/mod_verlet.F90:1466-1529)
!   Update_Position (Beeman + boundary + Ramo)                 (src/mod_verlet.F90:115-509)
!   Add_Particle / Mark_Particles_Remove / Remove_Particles    (src/mod_pair.F90:29-562)
!   Get_Kevin_Jgtf_v2                                          (src/mod_kevin_rjgtf_v2.f90:56-74)
!   w_theta_xy after Read_work_function                        (src/mod_work_function.F90:53-194)
!   xyz_corr / xi,eta,phi_coor / surface_normal / Field_normal /
!   Tip_Area / field_E_Hyperboloid                             (src/mod_hyperboloid_tip.f90:25-163)
!   Get_MB_Velocity through Init_Emission_Velocity             (src/mod_velocity.f90:20-68)
! Only the glue lives here: array allocation (as main.F90:422-456 does it),
! scratch output units, geometry set-up, and the Ramo unit vector of the
! planar and tip geometries (E_zunit_planar / E_zunit_tip restated, because
! their modules also pull in the Cuba integrator, which is not available).
module rd_ref_harness
  use, intrinsic :: iso_c_binding
  use mod_global
  use mod_pair
  use mod_verlet
  use mod_hyperboloid_tip
  use mod_kevin_rjgtf_v2
  use mod_work_function
  use mod_velocity
  !$ use omp_lib
  implicit none

  integer :: ref_cap = 0
  integer :: ref_geom = 1

contains

  function hz_planar(pos) result(e)
    double precision, dimension(1:3), intent(in) :: pos
    double precision, dimension(1:3) :: e
    e(1:2) = 0.0d0
    e(3) = -1.0d0/d
  end function hz_planar

  function hz_tip(pos) result(e)
    double precision, dimension(1:3), intent(in) :: pos
    double precision, dimension(1:3) :: e
    e = field_E_Hyperboloid(pos)
    e = e * pre_fac_E_tip_unit_voltage / pre_fac_E_tip
  end function hz_tip

  subroutine open_scratch()
    open(newunit=ud_density_emit_elec, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_density_emit_ion, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_density_emit_atom, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_density_emit, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_density_absorb_top, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_density_absorb_bot, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_density_absorb_recom, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_ramo_sec, status='scratch', form='unformatted', access='stream')
    open(newunit=ud_absorb, status='scratch', form='formatted')
    open(newunit=ud_absorb_top, status='scratch', form='formatted')
    open(newunit=ud_absorb_bot, status='scratch', form='formatted')
    open(newunit=ud_absorb_recom, status='scratch', form='formatted')
    open(newunit=ud_ramo, status='scratch', form='formatted')
    open(newunit=ud_volt, status='scratch', form='formatted')
  end subroutine open_scratch

  subroutine ref_init(cap) bind(C, name='ref_init')
    integer(c_int), value :: cap
    if (ref_cap > 0) return
    ref_cap = cap
    allocate(particles_cur_pos(1:3, 1:cap), particles_prev_pos(1:3, 1:cap))
    allocate(particles_last_col_pos(1:3, 1:cap), particles_cur_vel(1:3, 1:cap))
    allocate(particles_cur_accel(1:3, 1:cap), particles_prev_accel(1:3, 1:cap))
    allocate(particles_prev2_accel(1:3, 1:cap))
    allocate(particles_nearest_dist(1:cap), particles_nearest_id(1:cap))
    allocate(particles_charge(1:cap), particles_species(1:cap), particles_mass(1:cap))
    allocate(particles_step(1:cap), particles_emitter(1:cap), particles_section(1:cap))
    allocate(particles_life(1:cap), particles_id(1:cap))
    allocate(particles_mask(1:cap), particles_elec_mask(1:cap), particles_ion_mask(1:cap))
    allocate(particles_atom_mask(1:cap))
    allocate(particles_elec_pointer(1:cap), particles_ion_pointer(1:cap))
    allocate(particles_atom_pointer(1:cap), particles_inverse_pointer(1:cap))
    allocate(particles_cur_energy(1:cap), particles_ion_cross_sec(1:cap))
    allocate(particles_ion_cross_rad(1:cap), particles_recom_cross_rad(1:cap))
    allocate(particles_tot_cross_sec(1:cap))
    allocate(life_time(1:MAX_LIFE_TIME, 1:nrSpecies), ramo_current(1:nrSpecies))
    allocate(ramo_current_emit(1:MAX_SECTIONS, 1:MAX_EMITTERS))
    allocate(nrElec_remove_top_emit(1:MAX_EMITTERS))
    allocate(emitters_pos(1:3, 1:MAX_EMITTERS), emitters_dim(1:3, 1:MAX_EMITTERS))
    allocate(emitters_type(1:MAX_EMITTERS), emitters_delay(1:MAX_EMITTERS))
    emitters_pos = 0.0d0
    emitters_dim = 0.0d0
    emitters_type = 2
    emitters_delay = 0
    nrEmit = 1
    planes_N = 0
    call open_scratch()
    call ref_clear()
  end subroutine ref_init

  subroutine ref_set_threads(n) bind(C, name='ref_set_threads')
    integer(c_int), value :: n
    !$ call omp_set_num_threads(n)
  end subroutine ref_set_threads

  ! Reset the particle store, as main.F90:458-510 / mod_tests.F90:324-400 do.
  subroutine ref_clear() bind(C, name='ref_clear')
    particles_cur_pos = 0.0d0
    particles_prev_pos = 0.0d0
    particles_last_col_pos = 0.0d0
    particles_cur_vel = 0.0d0
    particles_cur_accel = 0.0d0
    particles_prev_accel = 0.0d0
    particles_prev2_accel = 0.0d0
    particles_charge = 0.0d0
    particles_species = 0
    particles_mass = 0.0d0
    particles_step = 0
    particles_emitter = 0
    particles_section = 0
    particles_life = 0
    particles_mask = .true.
    particles_elec_mask = .true.
    particles_ion_mask = .true.
    particles_atom_mask = .true.
    particles_id = 0
    particles_cur_energy = 0.0d0
    particles_ion_cross_sec = 0.0d0
    particles_ion_cross_rad = 0.0d0
    particles_recom_cross_rad = 0.0d0
    particles_tot_cross_sec = 0.0d0
    ramo_current = 0.0d0
    ramo_current_emit = 0.0d0
    life_time = 0
    nrPart = 0; nrElec = 0; nrIon = 0; nrAtom = 0; nrElecIon = 0
    nrPart_remove = 0; nrElec_remove = 0; nrIon_remove = 0; nrAtom_remove = 0
    nrPart_remove_top = 0; nrPart_remove_bot = 0; nrPart_remove_recom = 0; nrPart_remove_ion = 0
    nrElec_remove_top = 0; nrElec_remove_bot = 0; nrElec_remove_recom = 0; nrElec_remove_ion = 0
    nrIon_remove_top = 0; nrIon_remove_bot = 0; nrIon_remove_recom = 0; nrAtom_remove_ion = 0
    nrElec_remove_top_emit = 0
    nrPart_dropped = 0
    nrID = 0
    two_time_step = .false.
  end subroutine ref_clear

  ! Constants as the reference builds them (mod_global).
  subroutine ref_consts(out) bind(C, name='ref_consts')
    real(c_double), intent(out) :: out(12)
    out(1) = pi
    out(2) = epsilon_0
    out(3) = div_fac_c
    out(4) = h_bar
    out(5) = q_0
    out(6) = m_0
    out(7) = m_N2p
    out(8) = length_scale**2
    out(9) = q_02
    out(10) = k_b
    out(11) = m_u
    out(12) = mu_0
  end subroutine ref_consts

  subroutine ref_setup_planar(d_m, V, dt, ic, nic) bind(C, name='ref_setup_planar')
    real(c_double), value :: d_m, V, dt
    integer(c_int), value :: ic, nic
    box_dim = (/ 100.0d0*length_scale, 100.0d0*length_scale, d_m /)
    d = d_m
    V_s = V
    V_d = V
    E_z = -1.0d0*V_d/d
    time_step = dt
    time_step2 = dt**2
    image_charge = (ic /= 0)
    N_ic_max = nic
    ptr_field_E => field_E_planar
    ptr_Image_Charge_effect => Force_Image_charges_v2
    ptr_Check_Boundary => Check_Boundary_Planar
    ptr_E_zunit => hz_planar
    ref_geom = 1
  end subroutine ref_setup_planar

  ! Geometry formulas of Init_Emission_Tip (src/mod_emission_tip.f90:105-125).
  subroutine ref_setup_tip(V, dtip, rbase, htip, dt, ic) bind(C, name='ref_setup_tip')
    real(c_double), value :: V, dtip, rbase, htip, dt
    integer(c_int), value :: ic
    R_base = rbase
    h_tip = htip
    d = dtip + htip
    V_s = V
    V_d = V
    E_z = 0.0d0
    time_step = dt
    time_step2 = dt**2
    box_dim = (/ 0.0d0, 0.0d0, d /)
    max_xi = h_tip/dtip + 1.0d0
    a_foci = sqrt(dtip**2*R_base**2 / (h_tip**2 + 2*dtip*h_tip) + dtip**2)
    eta_1 = -1.0d0 * dtip / a_foci
    theta_tip = acos(dtip/a_foci)
    r_tip = a_foci * sin(theta_tip) * tan(theta_tip)
    shift_z = abs(a_foci * eta_1 * max_xi)
    pre_fac_E_tip = log( (1.0d0 + eta_1)/(1.0d0 - eta_1) * (1.0d0 - eta_2)/(1.0d0 + eta_2) )
    pre_fac_E_tip_unit_voltage = 2.0d0 * 1.0d0 / (a_foci * pre_fac_E_tip)
    pre_fac_E_tip = 2.0d0 * V_s / (a_foci * pre_fac_E_tip)
    image_charge = (ic /= 0)
    ptr_field_E => field_E_Hyperboloid
    ptr_Image_Charge_effect => Sphere_IC_field
    ptr_Check_Boundary => Check_Boundary_Planar
    ptr_E_zunit => hz_tip
    ref_geom = 2
  end subroutine ref_setup_tip

  subroutine ref_tip_params(out) bind(C, name='ref_tip_params')
    real(c_double), intent(out) :: out(8)
    out(1) = a_foci
    out(2) = shift_z
    out(3) = pre_fac_E_tip
    out(4) = pre_fac_E_tip_unit_voltage
    out(5) = r_tip
    out(6) = eta_1
    out(7) = h_tip
    out(8) = theta_tip
  end subroutine ref_tip_params

  ! two_time_step with atom_time_interval k (0: off), as Read_Input_Variables
  ! sets them (src/main.F90:391-395).
  subroutine ref_set_two_time_step(k) bind(C, name='ref_set_two_time_step')
    integer(c_int), value :: k
    atom_time_interval = k
    two_time_step = (k /= 0)
  end subroutine ref_set_two_time_step

  subroutine ref_add_particle(pos, vel, spc, step, emit, sec) bind(C, name='ref_add_particle')
    real(c_double), intent(in) :: pos(3), vel(3)
    integer(c_int), value :: spc, step, emit, sec
    call Add_Particle(pos, vel, spc, step, emit, -1, sec)
  end subroutine ref_add_particle

  ! Write a snapshot straight into the particle arrays (positions, charge,
  ! mass, species), for the force and field kernels.
  subroutine ref_put_snapshot(n, pos, q, m, spc) bind(C, name='ref_put_snapshot')
    integer(c_int), value :: n
    real(c_double), intent(in) :: pos(3, n), q(n), m(n)
    integer(c_int), intent(in) :: spc(n)
    call ref_clear()
    particles_cur_pos(:, 1:n) = pos
    particles_charge(1:n) = q
    particles_mass(1:n) = m
    particles_species(1:n) = spc
    nrPart = n
    nrElec = count(spc == species_elec)
    nrIon = count(spc == species_ion)
    nrElecIon = nrElec + nrIon
  end subroutine ref_put_snapshot

  subroutine ref_accel(which, n, out) bind(C, name='ref_accel')
    integer(c_int), value :: which, n
    real(c_double), intent(out) :: out(3, n)
    particles_cur_accel(:, 1:nrPart) = 0.0d0
    select case (which)
    case (0)
      call Calculate_Acceleration_Particles_Planar()
    case (1)
      call Calculate_Acceleration_Particles_Generic()
    case default
      call Calculate_Acceleration_Particles_ACC(ref_geom)
    end select
    out = particles_cur_accel(:, 1:n)
  end subroutine ref_accel

  subroutine ref_field_at(m, pts, out) bind(C, name='ref_field_at')
    integer(c_int), value :: m
    real(c_double), intent(in) :: pts(3, m)
    real(c_double), intent(out) :: out(3, m)
    integer :: k
    do k = 1, m
      out(:, k) = Calc_Field_at(pts(:, k))
    end do
  end subroutine ref_field_at

  subroutine ref_update_position(step) bind(C, name='ref_update_position')
    integer(c_int), value :: step
    call Update_Position(step)
  end subroutine ref_update_position

  subroutine ref_remove_particles(step) bind(C, name='ref_remove_particles')
    integer(c_int), value :: step
    call Remove_Particles(step)
  end subroutine ref_remove_particles

  subroutine ref_mark_remove(i, m) bind(C, name='ref_mark_remove')
    integer(c_int), value :: i, m
    call Mark_Particles_Remove(i, m)
  end subroutine ref_mark_remove

  ! State read-back: n, positions, velocities, the three acceleration
  ! histories, previous positions, charge, mass and the integer columns.
  subroutine ref_get_state(n, pos, prev, vel, a0, a1, a2, q, m, ints) bind(C, name='ref_get_state')
    integer(c_int), value :: n
    real(c_double), intent(out) :: pos(3, n), prev(3, n), vel(3, n), a0(3, n), a1(3, n), a2(3, n)
    real(c_double), intent(out) :: q(n), m(n)
    integer(c_int), intent(out) :: ints(6, n)
    pos = particles_cur_pos(:, 1:n)
    prev = particles_prev_pos(:, 1:n)
    vel = particles_cur_vel(:, 1:n)
    a0 = particles_cur_accel(:, 1:n)
    a1 = particles_prev_accel(:, 1:n)
    a2 = particles_prev2_accel(:, 1:n)
    q = particles_charge(1:n)
    m = particles_mass(1:n)
    ints(1, :) = particles_species(1:n)
    ints(2, :) = particles_step(1:n)
    ints(3, :) = particles_emitter(1:n)
    ints(4, :) = particles_section(1:n)
    ints(5, :) = particles_id(1:n)
    ints(6, :) = merge(1, 0, particles_mask(1:n))
  end subroutine ref_get_state

  subroutine ref_counts(out) bind(C, name='ref_counts')
    integer(c_int), intent(out) :: out(16)
    out(1) = nrPart
    out(2) = nrElec
    out(3) = nrIon
    out(4) = nrPart_remove
    out(5) = nrElec_remove
    out(6) = nrIon_remove
    out(7) = nrPart_remove_top
    out(8) = nrPart_remove_bot
    out(9) = nrElec_remove_top
    out(10) = nrElec_remove_bot
    out(11) = nrIon_remove_top
    out(12) = nrIon_remove_bot
    out(13) = nrID
    out(14) = particles_charge_rev
    out(15) = nrElec_remove_top_emit(1)
    out(16) = nrPart_dropped
  end subroutine ref_counts

  subroutine ref_ramo(out) bind(C, name='ref_ramo')
    real(c_double), intent(out) :: out(4)
    out(1:3) = ramo_current(1:3)
    out(4) = sum(ramo_current)
  end subroutine ref_ramo

  subroutine ref_life_time(out) bind(C, name='ref_life_time')
    integer(c_int), intent(out) :: out(MAX_LIFE_TIME, nrSpecies)
    out = life_time
  end subroutine ref_life_time

  function ref_jgtf(F, T, w) bind(C, name='ref_jgtf') result(j)
    real(c_double), value :: F, T, w
    real(c_double) :: j
    j = Get_Kevin_Jgtf_v2(F, T, w)
  end function ref_jgtf

  ! Read the work-function file "work" from the current directory, with the
  ! emitter at (x0, y0) of size (lx, ly), in metres.
  subroutine ref_read_work(x0, y0, lx, ly) bind(C, name='ref_read_work')
    real(c_double), value :: x0, y0, lx, ly
    emitters_pos(1:3, 1) = (/ x0, y0, 0.0d0 /)
    emitters_dim(1:3, 1) = (/ lx, ly, 0.0d0 /)
    call Read_work_function()
  end subroutine ref_read_work

  subroutine ref_work_at(m, pts, w, sec) bind(C, name='ref_work_at')
    integer(c_int), value :: m
    real(c_double), intent(in) :: pts(3, m)
    real(c_double), intent(out) :: w(m)
    integer(c_int), intent(out) :: sec(m)
    integer :: k, s
    do k = 1, m
      w(k) = w_theta_xy(pts(:, k), 1, s)
      sec(k) = s
    end do
  end subroutine ref_work_at

  ! ---- hyperboloid tip surface (after ref_setup_tip) ----------------------
  subroutine ref_tip_xyz(xi, eta, phi, out) bind(C, name='ref_tip_xyz')
    real(c_double), value :: xi, eta, phi
    real(c_double), intent(out) :: out(3)
    out = xyz_corr(xi, eta, phi)
  end subroutine ref_tip_xyz

  subroutine ref_tip_coords(pos, out) bind(C, name='ref_tip_coords')
    real(c_double), intent(in) :: pos(3)
    real(c_double), intent(out) :: out(3)
    out(1) = xi_coor(pos(1), pos(2), pos(3))
    out(2) = eta_coor(pos(1), pos(2), pos(3))
    out(3) = phi_coor(pos(1), pos(2), pos(3))
  end subroutine ref_tip_coords

  subroutine ref_tip_normal(pos, out) bind(C, name='ref_tip_normal')
    real(c_double), intent(in) :: pos(3)
    real(c_double), intent(out) :: out(3)
    out = surface_normal(pos)
  end subroutine ref_tip_normal

  function ref_field_normal(pos, e) bind(C, name='ref_field_normal') result(f)
    real(c_double), intent(in) :: pos(3), e(3)
    real(c_double) :: f
    f = Field_normal(pos, e)
  end function ref_field_normal

  function ref_tip_area(xi1, xi2, phi1, phi2) bind(C, name='ref_tip_area') result(a)
    real(c_double), value :: xi1, xi2, phi1, phi2
    real(c_double) :: a
    a = Tip_Area(xi1, xi2, phi1, phi2)
  end function ref_tip_area

  subroutine ref_tip_vacuum(pos, out) bind(C, name='ref_tip_vacuum')
    real(c_double), intent(in) :: pos(3)
    real(c_double), intent(out) :: out(3)
    out = field_E_Hyperboloid(pos)
  end subroutine ref_tip_vacuum

  ! ---- Maxwell-Boltzmann emission velocity -------------------------------
  ! n draws of ptr_Get_Emission_Velocity after Init_Emission_Velocity(VELOCITY_MB)
  ! at temperature T, with the intrinsic generator seeded from `seed`.
  subroutine ref_mb_velocity(T, seed, n, out) bind(C, name='ref_mb_velocity')
    real(c_double), value :: T
    integer(c_int), value :: seed, n
    real(c_double), intent(out) :: out(3, n)
    integer :: k, i
    integer, allocatable :: sd(:)
    T_temp = T
    call Init_Emission_Velocity(VELOCITY_MB)
    call random_seed(size=k)
    allocate(sd(k))
    do i = 1, k
      sd(i) = seed + 7919 * i
    end do
    call random_seed(put=sd)
    do i = 1, n
      out(:, i) = ptr_Get_Emission_Velocity()
    end do
  end subroutine ref_mb_velocity
end module rd_ref_harness
