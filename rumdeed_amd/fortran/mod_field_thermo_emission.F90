! mod_field_thermo_emission -- planar thermal-field (GTF) emission with the
! reference's interface (src/mod_field_thermo_emission.F90:49-193), on the
! device: supply integral of J_GTF(F, T_temp, phi(x, y)), Poisson count,
! Metropolis-Hastings sites, Maxwell-Boltzmann emission velocities.
!
!   Init_Field_Thermo_Emission()   binds ptr_Do_Emission => Do_Field_Thermo_Emission
!   Do_Field_Thermo_Emission(step)
!   Clean_Up_Field_Thermo_Emission()
module mod_field_thermo_emission
  use, intrinsic :: iso_c_binding
  use mod_global
  use rumdeed_hip
  use rd_device
  implicit none
  private
  public :: Init_Field_Thermo_Emission, Clean_Up_Field_Thermo_Emission, Do_Field_Thermo_Emission

  double precision, allocatable, target :: wf_table(:, :)

contains

  subroutine Init_Field_Thermo_Emission()
    call RD_Open(RD_GEOM_PLANAR)
    call RD_Read_Work(wf_table)
    call RD_Use_Emitter(RD_EMIT_PLANAR_GTF, 4.7d0, wf_table)
    ptr_Do_Emission => Do_Field_Thermo_Emission
  end subroutine Init_Field_Thermo_Emission

  subroutine Do_Field_Thermo_Emission(step)
    integer, intent(in) :: step
    type(rd_step_record) :: rec
    call RD_Require()
    call rd_check(rd_do_emission(rd_ctx, int(step, c_int), rd_precision, rec), 'rd_do_emission')
    call RD_Mirror(rec)
  end subroutine Do_Field_Thermo_Emission

  subroutine Clean_Up_Field_Thermo_Emission()
    if (allocated(wf_table)) deallocate(wf_table)
    call RD_Close()
  end subroutine Clean_Up_Field_Thermo_Emission
end module mod_field_thermo_emission
