! mod_verlet -- the integrator and the force/field routines of RUMDEED's hot
! path, with the reference's interfaces (src/mod_verlet.F90), on the device.
!
!   Update_Position(step)              src/mod_verlet.F90:115-162: Beeman
!       position update with boundary marking, accelerations, Beeman velocity
!       update and Ramo current, all on the device-resident particles
!   Calculate_Acceleration_Particles() :597-620: the accelerations of the
!       device-resident particles, added onto particles_cur_accel(:, 1:nrPart)
!   Calc_Field_at(pos, org_pos, is_surface)  :1466-1529
!   Calc_Field_at_Batch(M, pos_in, field_out) :1635-1911
!   Particles_To_Device / Release_Device_Particles  :28-111 (the particles
!       are always device-resident: nothing to do)
!
! org_pos and is_surface only matter for the kd-tree tip geometries of the
! reference (their ptr_field_E); the planar and hyperboloid geometries on the
! device ignore them, as the reference's field_E_planar and
! field_E_Hyperboloid do.
module mod_verlet
  use, intrinsic :: iso_c_binding
  use mod_global
  use rumdeed_hip
  use rd_device
  implicit none
  private
  public :: Update_Position, Calculate_Acceleration_Particles, Calc_Field_at, Calc_Field_at_Batch
  public :: Particles_To_Device, Release_Device_Particles

contains

  subroutine Update_Position(step)
    integer, intent(in) :: step
    type(rd_step_record) :: rec
    call RD_Require()
    cur_time = time_step * step / time_scale
    call rd_check(rd_update_position(rd_ctx, int(step, c_int), rd_precision, rec), &
                  'rd_update_position')
    call RD_Mirror(rec)
  end subroutine Update_Position

  subroutine Calculate_Acceleration_Particles()
    double precision, allocatable :: acc(:, :)
    call RD_Require()
    if (nrPart < 1) return
    allocate(acc(3, nrPart))
    call rd_check(rd_accel(rd_ctx, rd_precision), 'rd_accel')
    call rd_check(rd_get_accel(rd_ctx, acc), 'rd_get_accel')
    particles_cur_accel(:, 1:nrPart) = particles_cur_accel(:, 1:nrPart) + acc
  end subroutine Calculate_Acceleration_Particles

  function Calc_Field_at(pos_xyz, org_pos, is_surface)
    double precision, dimension(1:3) :: Calc_Field_at
    double precision, dimension(1:3), intent(in) :: pos_xyz
    double precision, dimension(1:3), intent(in), optional :: org_pos
    logical, intent(in), optional :: is_surface
    double precision :: p(3, 1), f(3, 1)
    p(:, 1) = pos_xyz
    call Calc_Field_at_Batch(1, p, f)
    Calc_Field_at = f(:, 1)
  end function Calc_Field_at

  subroutine Calc_Field_at_Batch(M, pos_in, field_out)
    integer, intent(in) :: M
    double precision, dimension(1:3, 1:M), intent(in) :: pos_in
    double precision, dimension(1:3, 1:M), intent(out) :: field_out
    call RD_Require()
    if (M < 1) return
    call rd_check(rd_field_batch(rd_ctx, int(M, c_int), pos_in, field_out, rd_precision), &
                  'rd_field_batch')
  end subroutine Calc_Field_at_Batch

  subroutine Particles_To_Device()
  end subroutine Particles_To_Device

  subroutine Release_Device_Particles()
  end subroutine Release_Device_Particles
end module mod_verlet
