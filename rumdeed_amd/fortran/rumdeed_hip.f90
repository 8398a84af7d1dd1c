! ISO_C_BINDING interface of librumdeed.so (include/rumdeed.h) and drop-in
! routines with the RUMDEED module interfaces on top of it.
!
!   module rumdeed_hip      bind(C) interfaces, derived types, constants
!   module rumdeed_dropin   routines named and shaped like the reference's:
!     Calculate_Acceleration_Particles   (src/mod_verlet.F90:597-620)
!     Calc_Field_at / Calc_Field_at_Batch (src/mod_verlet.F90:1466-1529, :1635-1911)
!     Update_Position                     (src/mod_verlet.F90:115-162)
!     Remove_Particles                    (src/mod_pair.F90:352-562)
!     Add_Particle                        (src/mod_pair.F90:29-159)
!     Do_Emission                         (ptr_Do_Emission, src/mod_global.F90:445-506)
!   The reference keeps its particle arrays in mod_global; the device keeps
!   them in HBM.  The host-array routines upload, compute and download (the
!   reference's OpenACC residency pattern); the step routines work on the
!   device-resident state, which is what a main loop calling Update_Position /
!   Remove_Particles / Do_Emission per step uses.
module rumdeed_hip
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: RD_OK = 0
  integer(c_int), parameter :: RD_GEOM_PLANAR = 1, RD_GEOM_TIP = 2
  integer(c_int), parameter :: RD_PREC_FP64 = 0, RD_PREC_MIXED = 1
  integer(c_int), parameter :: RD_EMIT_NONE = 0, RD_EMIT_TIP_FE = 3
  integer(c_int), parameter :: RD_EMIT_PLANAR_GTF = 9, RD_EMIT_PLANAR_FE = 10

  type, bind(C) :: rd_geometry
    integer(c_int) :: kind, image_charge, n_ic_max, reserved
    real(c_double) :: V, d, d_tip, R_base, h_tip, box_top
  end type rd_geometry

  type, bind(C) :: rd_emitter
    integer(c_int) :: mode, delay
    real(c_double) :: x0, y0, lx, ly, T, w_tip
    integer(c_int) :: wf_nx, wf_ny
    type(c_ptr)    :: wf_table
    integer(c_int) :: quad_n, ndim
  end type rd_emitter

  type, bind(C) :: rd_step_record
    integer(c_int) :: step, n, emitted, n_round
    integer(c_int) :: absorbed, absorbed_top, absorbed_bot, dropped
    integer(c_int) :: start_rounds, reserved0, reserved1, reserved2
    real(c_double) :: n_sup, ramo_current, mh_std, a_rate, df_avg
    real(c_double) :: f_avg(3), ramo(3), avg_elec_vel(3)
  end type rd_step_record

  interface
    function rd_last_error() bind(C, name='rd_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function rd_ctx_create(device, max_particles, ctx) bind(C, name='rd_ctx_create') result(r)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int), value :: device
      integer(c_int64_t), value :: max_particles
      type(c_ptr) :: ctx
      integer(c_int) :: r
    end function
    function rd_ctx_destroy(ctx) bind(C, name='rd_ctx_destroy') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: r
    end function
    function rd_set_geometry(ctx, g) bind(C, name='rd_set_geometry') result(r)
      import :: c_int, c_ptr, rd_geometry
      type(c_ptr), value :: ctx
      type(rd_geometry), intent(in) :: g
      integer(c_int) :: r
    end function
    function rd_set_time_step(ctx, dt) bind(C, name='rd_set_time_step') result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int) :: r
    end function
    function rd_set_seed(ctx, seed, stream) bind(C, name='rd_set_seed') result(r)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed, stream
      integer(c_int) :: r
    end function
    function rd_set_emitter(ctx, em) bind(C, name='rd_set_emitter') result(r)
      import :: c_int, c_ptr, rd_emitter
      type(c_ptr), value :: ctx
      type(rd_emitter), intent(in) :: em
      integer(c_int) :: r
    end function
    function rd_put_particles(ctx, n, pos, vel, q, m, species) bind(C, name='rd_put_particles') result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double), intent(in) :: pos(3, *), vel(3, *), q(*), m(*)
      integer(c_int), intent(in) :: species(*)
      integer(c_int) :: r
    end function
    function rd_particle_count(ctx, n) bind(C, name='rd_particle_count') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: n
      integer(c_int) :: r
    end function
    function rd_accel(ctx, precision) bind(C, name='rd_accel') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      integer(c_int) :: r
    end function
    function rd_get_accel(ctx, acc) bind(C, name='rd_get_accel') result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: acc(3, *)
      integer(c_int) :: r
    end function
    function rd_field_batch(ctx, m, pts, field, precision) bind(C, name='rd_field_batch') result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: m, precision
      real(c_double), intent(in) :: pts(3, *)
      real(c_double) :: field(3, *)
      integer(c_int) :: r
    end function
    function rd_add_particles(ctx, n, pos, vel, species, step, emit, sec) &
        bind(C, name='rd_add_particles') result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, step, emit
      real(c_double), intent(in) :: pos(3, *), vel(3, *)
      integer(c_int), intent(in) :: species(*), sec(*)
      integer(c_int) :: r
    end function
    function rd_update_position(ctx, step, precision, rec) bind(C, name='rd_update_position') result(r)
      import :: c_int, c_ptr, rd_step_record
      type(c_ptr), value :: ctx
      integer(c_int), value :: step, precision
      type(rd_step_record) :: rec
      integer(c_int) :: r
    end function
    function rd_remove_particles(ctx, step, rec) bind(C, name='rd_remove_particles') result(r)
      import :: c_int, c_ptr, rd_step_record
      type(c_ptr), value :: ctx
      integer(c_int), value :: step
      type(rd_step_record) :: rec
      integer(c_int) :: r
    end function
    function rd_do_emission(ctx, step, precision, rec) bind(C, name='rd_do_emission') result(r)
      import :: c_int, c_ptr, rd_step_record
      type(c_ptr), value :: ctx
      integer(c_int), value :: step, precision
      type(rd_step_record) :: rec
      integer(c_int) :: r
    end function
    function rd_get_state(ctx, pos, prev, vel, acc, a1, a2, ints) bind(C, name='rd_get_state') result(r)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: pos(3, *), prev(3, *), vel(3, *), acc(3, *), a1(3, *), a2(3, *)
      integer(c_int) :: ints(5, *)
      integer(c_int) :: r
    end function
  end interface

contains

  ! Stop with the library's message on failure (the reference stops on fatal
  ! conditions, e.g. src/main.F90:360-364).
  subroutine rd_check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (code == RD_OK) return
    call c_f_pointer(rd_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    print '(a, a, a, a)', 'RUMDEED-HIP: ', what, ': ', trim(text)
    error stop 1
  end subroutine rd_check
end module rumdeed_hip


module rumdeed_dropin
  use, intrinsic :: iso_c_binding
  use rumdeed_hip
  implicit none
  type(c_ptr) :: rd_ctx = c_null_ptr
  integer(c_int) :: rd_precision = RD_PREC_FP64

contains

  subroutine RD_Init(max_particles, geom, dt)
    integer, intent(in) :: max_particles
    type(rd_geometry), intent(in) :: geom
    double precision, intent(in) :: dt
    call rd_check(rd_ctx_create(0_c_int, int(max_particles, c_int64_t), rd_ctx), 'rd_ctx_create')
    call rd_check(rd_set_geometry(rd_ctx, geom), 'rd_set_geometry')
    call rd_check(rd_set_time_step(rd_ctx, dt), 'rd_set_time_step')
  end subroutine RD_Init

  subroutine RD_Finalize()
    integer(c_int) :: r
    r = rd_ctx_destroy(rd_ctx)
    rd_ctx = c_null_ptr
  end subroutine RD_Finalize

  ! Calculate_Acceleration_Particles on host arrays: the reference adds the
  ! result onto particles_cur_accel (zeroed by the position update).
  subroutine Calculate_Acceleration_Particles(n, cur_pos, charge, mass, species, cur_accel)
    integer, intent(in) :: n
    double precision, intent(in) :: cur_pos(3, n), charge(n), mass(n)
    integer, intent(in) :: species(n)
    double precision, intent(inout) :: cur_accel(3, n)
    double precision, allocatable :: vel(:, :), acc(:, :)
    integer(c_int), allocatable :: sp(:)
    allocate(vel(3, n), acc(3, n), sp(n))
    vel = 0.0d0
    sp = int(species, c_int)
    call rd_check(rd_put_particles(rd_ctx, int(n, c_int), cur_pos, vel, charge, mass, sp), &
                  'rd_put_particles')
    call rd_check(rd_accel(rd_ctx, rd_precision), 'rd_accel')
    call rd_check(rd_get_accel(rd_ctx, acc), 'rd_get_accel')
    cur_accel = cur_accel + acc
  end subroutine Calculate_Acceleration_Particles

  ! Calc_Field_at_Batch on the particles last put / resident on the device.
  subroutine Calc_Field_at_Batch(M, pos_in, field_out)
    integer, intent(in) :: M
    double precision, intent(in) :: pos_in(3, M)
    double precision, intent(out) :: field_out(3, M)
    call rd_check(rd_field_batch(rd_ctx, int(M, c_int), pos_in, field_out, rd_precision), &
                  'rd_field_batch')
  end subroutine Calc_Field_at_Batch

  function Calc_Field_at(pos_xyz) result(field)
    double precision, intent(in) :: pos_xyz(3)
    double precision :: field(3)
    double precision :: p(3, 1), f(3, 1)
    p(:, 1) = pos_xyz
    call Calc_Field_at_Batch(1, p, f)
    field = f(:, 1)
  end function Calc_Field_at

  ! Add_Particle (one particle, appended to the device-resident state).
  ! `life` (particles_life, src/mod_pair.F90:73) is read only by the ion
  ! collision model (src/mod_collisions.F90:123), which this path does not
  ! include; it is accepted for interface compatibility and not stored.
  subroutine Add_Particle(par_pos, par_vel, par_species, step, emit, life, opt_sec)
    double precision, intent(in) :: par_pos(3), par_vel(3)
    integer, intent(in) :: par_species, step, emit, life
    integer, intent(in), optional :: opt_sec
    integer(c_int) :: sp(1), sec(1)
    sp(1) = int(par_species, c_int)
    sec(1) = 1
    if (present(opt_sec)) sec(1) = int(opt_sec, c_int)
    call rd_check(rd_add_particles(rd_ctx, 1_c_int, par_pos, par_vel, sp, int(step, c_int), &
                  int(emit, c_int), sec), 'rd_add_particles')
  end subroutine Add_Particle

  subroutine Update_Position(step, rec)
    integer, intent(in) :: step
    type(rd_step_record), intent(out) :: rec
    call rd_check(rd_update_position(rd_ctx, int(step, c_int), rd_precision, rec), &
                  'rd_update_position')
  end subroutine Update_Position

  subroutine Remove_Particles(step, rec)
    integer, intent(in) :: step
    type(rd_step_record), intent(out) :: rec
    call rd_check(rd_remove_particles(rd_ctx, int(step, c_int), rec), 'rd_remove_particles')
  end subroutine Remove_Particles

  subroutine Do_Emission(step, rec)
    integer, intent(in) :: step
    type(rd_step_record), intent(out) :: rec
    call rd_check(rd_do_emission(rd_ctx, int(step, c_int), rd_precision, rec), 'rd_do_emission')
  end subroutine Do_Emission

  integer function RD_Count()
    integer(c_int) :: n
    call rd_check(rd_particle_count(rd_ctx, n), 'rd_particle_count')
    RD_Count = n
  end function RD_Count
end module rumdeed_dropin
