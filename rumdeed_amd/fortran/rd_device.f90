! rd_device -- the device context behind the drop-in modules (mod_verlet,
! mod_pair and the emission modules of this directory).
!
! The reference keeps the particle state in the mod_global arrays and, in its
! OpenACC build, caches a device copy (Particles_To_Device /
! Release_Device_Particles, src/mod_verlet.F90:28-111).  Here the state lives
! in HBM for the whole run and the host arrays are a mirror:
!   - after every Do_Emission / Update_Position / Remove_Particles the
!     counters of mod_global (nrPart, nrElec, nrIon, nrID, the removal
!     counters, nrPart_dropped), ramo_current and cur_time are updated;
!   - the particle arrays are copied back by RD_Sync_To_Host (or after every
!     step when rd_sync_each_step is set), and a host program that edits them
!     hands them back with RD_Put_From_Host.
! RD_Last_Record returns the record of the current step (the figures of
! emitted.dt, field.dt, absorbed.dt and ramo_current.dt).
module rd_device
  use, intrinsic :: iso_c_binding
  use rumdeed_hip
  use mod_global
  implicit none

  type(c_ptr) :: rd_ctx = c_null_ptr
  integer(c_int) :: rd_precision = RD_PREC_FP64
  integer(c_int64_t) :: rd_seed = 0_c_int64_t, rd_stream = 0_c_int64_t
  type(rd_step_record) :: rd_record
  logical :: rd_sync_each_step = .false.
  integer :: rd_ions = 0  ! ions appended through Add_Particle (the device emits electrons)
  integer :: rd_capacity = 0

contains

  ! FP64 (bit-reproducible, the default) or RD_PREC_MIXED.
  subroutine RD_Set_Precision(p)
    integer, intent(in) :: p
    rd_precision = int(p, c_int)
  end subroutine RD_Set_Precision

  ! Seed of the counter-based generator (the reference seeds its intrinsic
  ! generator from /dev/urandom, src/main.F90:747-793).
  subroutine RD_Use_Seed(seed, stream)
    integer(c_int64_t), intent(in) :: seed
    integer(c_int64_t), intent(in), optional :: stream
    rd_seed = seed
    if (present(stream)) rd_stream = stream
    if (c_associated(rd_ctx)) call rd_check(rd_set_seed(rd_ctx, rd_seed, rd_stream), 'rd_set_seed')
  end subroutine RD_Use_Seed

  ! Create the device context for the geometry in mod_global: planar (gap d,
  ! voltage V_s, image_charge, N_ic_max) or the hyperboloid tip
  ! (emitters_dim(1:3, 1) = d_tip, R_base, h_tip), removal height box_dim(3).
  ! Capacity: the size of the particle arrays, else MAX_PARTICLES.
  subroutine RD_Open(kind)
    integer(c_int), intent(in) :: kind
    type(rd_geometry) :: g
    integer :: cap
    if (c_associated(rd_ctx)) call RD_Close()
    cap = MAX_PARTICLES
    if (allocated(particles_cur_pos)) cap = size(particles_cur_pos, 2)
    rd_capacity = cap
    call rd_check(rd_ctx_create(0_c_int, int(cap, c_int64_t), rd_ctx), 'rd_ctx_create')
    g%kind = kind
    g%image_charge = merge(1_c_int, 0_c_int, image_charge)
    g%n_ic_max = int(N_ic_max, c_int)
    g%reserved = 0
    g%V = V_s
    g%d = d
    g%d_tip = 0.0d0
    g%R_base = 0.0d0
    g%h_tip = 0.0d0
    if (kind == RD_GEOM_TIP) then
      g%d_tip = emitters_dim(1, 1)
      g%R_base = emitters_dim(2, 1)
      g%h_tip = emitters_dim(3, 1)
    end if
    g%box_top = box_dim(3)
    call rd_check(rd_set_geometry(rd_ctx, g), 'rd_set_geometry')
    call rd_check(rd_set_time_step(rd_ctx, time_step), 'rd_set_time_step')
    ! two_time_step follows atom_time_interval (src/main.F90:391-395)
    two_time_step = atom_time_interval /= 0
    call rd_check(rd_set_two_time_step(rd_ctx, int(atom_time_interval, c_int)), &
                  'rd_set_two_time_step')
    call rd_check(rd_set_seed(rd_ctx, rd_seed, rd_stream), 'rd_set_seed')
    nrPart = 0
    nrElec = 0
    nrIon = 0
    nrID = 0
    rd_ions = 0
    if (.not. allocated(ramo_current)) allocate(ramo_current(1:nrSpecies))
    ramo_current = 0.0d0
    if (.not. allocated(life_time)) allocate(life_time(1:MAX_LIFE_TIME, 1:nrSpecies))
    life_time = 0
  end subroutine RD_Open

  ! The emitter of the simulation (emitters_pos/dim/delay(1) of mod_global):
  ! mode RD_EMIT_*, work-function table (rows top first, as in `work`).
  subroutine RD_Use_Emitter(mode, w_tip, wf, quad_n, quad_levels)
    integer(c_int), intent(in) :: mode
    double precision, intent(in) :: w_tip
    double precision, intent(in), target :: wf(:, :)  ! (nx, ny): wf(:, row)
    integer, intent(in), optional :: quad_n
    integer, intent(in), optional :: quad_levels  ! supply refinement (-1 none, 0 default)
    type(rd_emitter) :: em
    em%mode = mode
    em%delay = 0
    if (allocated(emitters_delay)) em%delay = int(emitters_delay(1), c_int)
    em%x0 = 0.0d0
    em%y0 = 0.0d0
    em%lx = 0.0d0
    em%ly = 0.0d0
    if (mode /= RD_EMIT_TIP_FE) then
      em%x0 = emitters_pos(1, 1)
      em%y0 = emitters_pos(2, 1)
      em%lx = emitters_dim(1, 1)
      em%ly = emitters_dim(2, 1)
    end if
    em%T = T_temp
    em%w_tip = w_tip
    em%wf_nx = int(size(wf, 1), c_int)
    em%wf_ny = int(size(wf, 2), c_int)
    em%wf_table = c_loc(wf)
    em%quad_n = 0
    if (present(quad_n)) em%quad_n = int(quad_n, c_int)
    em%ndim = 0
    em%quad_levels = 0
    if (present(quad_levels)) em%quad_levels = int(quad_levels, c_int)
    em%reserved = 0
    call rd_check(rd_set_emitter(rd_ctx, em), 'rd_set_emitter')
  end subroutine RD_Use_Emitter

  ! Read the checkerboard work function from the file `work` in the current
  ! directory (format of Read_work_function, src/mod_work_function.F90:53-92:
  ! type 1, then ny nx, then ny rows of nx values, top row first).
  subroutine RD_Read_Work(wf)
    double precision, allocatable, intent(out) :: wf(:, :)
    integer :: u, ios, wtype, nx, ny, i, j
    open(newunit=u, file='work', status='OLD', action='READ', iostat=ios)
    if (ios /= 0) then
      print '(a)', 'RUMDEED: Failed to open file work. ABORTING'
      error stop 1
    end if
    read(u, *) wtype
    if (wtype /= 1) then
      print '(a, i0)', 'RUMDEED: work function type not on the device path: ', wtype
      error stop 1
    end if
    read(u, *) ny, nx
    allocate(wf(1:nx, 1:ny))
    do i = 1, ny
      read(u, *) (wf(j, i), j = 1, nx)
    end do
    close(u)
  end subroutine RD_Read_Work

  ! Mirror the device counters into mod_global after a step routine.
  subroutine RD_Mirror(rec)
    type(rd_step_record), intent(in) :: rec
    integer(c_int), allocatable, target :: sp(:)
    integer(c_int) :: n
    call rd_check(rd_particle_count(rd_ctx, n), 'rd_particle_count')
    nrPart = n
    if (rd_ions > 0 .and. n > 0) then
      allocate(sp(n))
      call rd_check(rd_get_particles(rd_ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                                     c_loc(sp)), 'rd_get_particles')
      nrIon = count(sp == species_ion)
      deallocate(sp)
    else
      nrIon = 0
    end if
    nrElec = nrPart - nrIon
    nrPart_remove = rec%absorbed
    nrElec_remove = rec%absorbed
    nrPart_remove_top = rec%absorbed_top
    nrPart_remove_bot = rec%absorbed_bot
    nrElec_remove_top = rec%absorbed_top
    nrPart_dropped = rec%dropped
    ramo_current(1:3) = rec%ramo(1:3)
    rd_record = rec
  end subroutine RD_Mirror

  ! The record of the current step (complete after Remove_Particles).
  subroutine RD_Last_Record(rec)
    type(rd_step_record), intent(out) :: rec
    rec = rd_record
  end subroutine RD_Last_Record

  ! Copy the device particle state into the mod_global arrays (1:nrPart).
  subroutine RD_Sync_To_Host()
    integer(c_int), allocatable, target :: ints(:, :)
    double precision, allocatable, target :: q(:), m(:), v(:, :, :)
    integer(c_int) :: n
    call rd_check(rd_particle_count(rd_ctx, n), 'rd_particle_count')
    nrPart = n
    if (n == 0) return
    allocate(ints(5, n), q(n), m(n), v(3, n, 6))
    call rd_check(rd_get_state(rd_ctx, c_loc(v(1, 1, 1)), c_loc(v(1, 1, 2)), c_loc(v(1, 1, 3)), &
                  c_loc(v(1, 1, 4)), c_loc(v(1, 1, 5)), c_loc(v(1, 1, 6)), c_loc(ints)), &
                  'rd_get_state')
    call rd_check(rd_get_particles(rd_ctx, c_null_ptr, c_null_ptr, c_loc(q), c_loc(m), &
                                   c_null_ptr), 'rd_get_particles')
    particles_cur_pos(:, 1:n) = v(:, :, 1)
    particles_prev_pos(:, 1:n) = v(:, :, 2)
    particles_cur_vel(:, 1:n) = v(:, :, 3)
    particles_cur_accel(:, 1:n) = v(:, :, 4)
    particles_prev_accel(:, 1:n) = v(:, :, 5)
    particles_prev2_accel(:, 1:n) = v(:, :, 6)
    particles_charge(1:n) = q
    particles_mass(1:n) = m
    particles_species(1:n) = ints(1, :)
    particles_step(1:n) = ints(2, :)
    particles_emitter(1:n) = ints(3, :)
    particles_section(1:n) = ints(4, :)
    particles_id(1:n) = ints(5, :)
  end subroutine RD_Sync_To_Host

  ! Replace the device particle set by the mod_global arrays (1:nrPart):
  ! positions, velocities, charges, masses and species; the device's
  ! acceleration histories start from zero, as after rd_put_particles.
  subroutine RD_Put_From_Host()
    integer(c_int), allocatable, target :: sp(:)
    double precision, allocatable, target :: p(:, :), v(:, :), q(:), m(:)
    integer :: n
    n = max(nrPart, 1)
    allocate(sp(n), p(3, n), v(3, n), q(n), m(n))
    if (nrPart > 0) then
      sp(1:nrPart) = int(particles_species(1:nrPart), c_int)
      p(:, 1:nrPart) = particles_cur_pos(:, 1:nrPart)
      v(:, 1:nrPart) = particles_cur_vel(:, 1:nrPart)
      q(1:nrPart) = particles_charge(1:nrPart)
      m(1:nrPart) = particles_mass(1:nrPart)
    end if
    call rd_check(rd_put_particles(rd_ctx, int(nrPart, c_int), c_loc(p), c_loc(v), c_loc(q), &
                  c_loc(m), c_loc(sp)), 'rd_put_particles')
    rd_ions = 0
    if (nrPart > 0) rd_ions = count(sp(1:nrPart) == species_ion)
  end subroutine RD_Put_From_Host

  subroutine RD_Close()
    integer(c_int) :: r
    if (c_associated(rd_ctx)) r = rd_ctx_destroy(rd_ctx)
    rd_ctx = c_null_ptr
  end subroutine RD_Close

  subroutine RD_Require()
    if (.not. c_associated(rd_ctx)) then
      print '(a)', 'RUMDEED: no device context (call an Init_* emission routine or RD_Open)'
      error stop 1
    end if
  end subroutine RD_Require
end module rd_device
