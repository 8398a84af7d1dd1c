! mod_field_emission_v2 -- planar Fowler-Nordheim emission with the
! reference's interface (src/mod_field_emission_v2.F90:113-205, :261-395),
! on the device: the surface supply integral, the Metropolis-Hastings site
! sampling of every candidate (the reference's mh_batch semantics), the
! escape test and Add_Particle run in HBM for the emitter of mod_global.
!
!   Init_Field_Emission_v2()    binds ptr_Do_Emission => Do_Field_Emission,
!                               creates the device context for the planar
!                               geometry and reads the work function `work`
!   Do_Field_Emission(step)     one emission step
!   Clean_Up_Field_Emission_v2()
module mod_field_emission_v2
  use, intrinsic :: iso_c_binding
  use mod_global
  use rumdeed_hip
  use rd_device
  implicit none
  private
  public :: Init_Field_Emission_v2, Clean_Up_Field_Emission_v2, Do_Field_Emission

  double precision, allocatable, target :: wf_table(:, :)

contains

  subroutine Init_Field_Emission_v2()
    call RD_Open(RD_GEOM_PLANAR)
    call RD_Read_Work(wf_table)
    call RD_Use_Emitter(RD_EMIT_PLANAR_FE, 4.7d0, wf_table)
    ptr_Do_Emission => Do_Field_Emission
  end subroutine Init_Field_Emission_v2

  subroutine Do_Field_Emission(step)
    integer, intent(in) :: step
    type(rd_step_record) :: rec
    call RD_Require()
    call rd_check(rd_do_emission(rd_ctx, int(step, c_int), rd_precision, rec), 'rd_do_emission')
    call RD_Mirror(rec)
  end subroutine Do_Field_Emission

  subroutine Clean_Up_Field_Emission_v2()
    if (allocated(wf_table)) deallocate(wf_table)
    call RD_Close()
  end subroutine Clean_Up_Field_Emission_v2
end module mod_field_emission_v2
