! ISO_C_BINDING interface of librumdeed.so (include/rumdeed.h): the bind(C)
! types, constants and entry points.  The drop-in modules of this directory
! (mod_verlet, mod_pair, mod_field_emission_v2, mod_field_thermo_emission,
! mod_emission_tip, over the context of rd_device) are built on it.
! Buffers the C ABI accepts as NULL are passed as type(c_ptr).
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
    integer(c_int) :: quad_n, ndim, quad_levels, reserved
  end type rd_emitter

  type, bind(C) :: rd_step_record
    integer(c_int) :: step, n, emitted, n_round
    integer(c_int) :: absorbed, absorbed_top, absorbed_bot, dropped
    integer(c_int) :: start_rounds, supply_level, supply_cells, supply_fail
    real(c_double) :: n_sup, ramo_current, mh_std, a_rate, df_avg
    real(c_double) :: f_avg(3), ramo(3), avg_elec_vel(3)
    real(c_double) :: supply_err
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
    function rd_set_two_time_step(ctx, k) bind(C, name='rd_set_two_time_step') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
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
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      type(c_ptr), value :: pos, vel, q, m, species
      integer(c_int) :: r
    end function
    function rd_particle_count(ctx, n) bind(C, name='rd_particle_count') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: n
      integer(c_int) :: r
    end function
    function rd_get_particles(ctx, pos, vel, q, m, species) bind(C, name='rd_get_particles') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pos, vel, q, m, species
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
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, step, emit
      type(c_ptr), value :: pos, vel, species, sec
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
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pos, prev, vel, acc, a1, a2, ints
      integer(c_int) :: r
    end function
    function rd_get_life_time(ctx, hist) bind(C, name='rd_get_life_time') result(r)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: hist(1000, 2)
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
