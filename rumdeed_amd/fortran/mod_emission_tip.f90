! mod_emission_tip -- field emission from the hyperboloid tip with the
! reference's interface (src/mod_emission_tip.f90:81-176, :417-534), on the
! device: the 100 x 100 (xi, phi) supply sum over the tip surface, the
! Metropolis-Hastings sites (Metro_algo_tip_v3), emission 1 nm along the
! surface normal.  The tip geometry comes from emitters_dim(1:3, 1) =
! d_tip, R_base, h_tip, as Init_Emission_Tip reads it.
!
! w_theta: the tip's work function (eV).  The reference fixes it at 4.7 eV
! (a parameter, :45); here it is a module variable with that default, so the
! 4.5 eV tip of the benchmark configurations can be set before Init.
module mod_emission_tip
  use, intrinsic :: iso_c_binding
  use mod_global
  use rumdeed_hip
  use rd_device
  implicit none
  private
  public :: Init_Emission_Tip, Clean_Up_Emission_Tip, Do_Emission_Tip, w_theta

  double precision :: w_theta = 4.7d0
  double precision, target :: wf_one(1, 1) = 4.7d0

contains

  subroutine Init_Emission_Tip()
    call RD_Open(RD_GEOM_TIP)
    wf_one = w_theta
    call RD_Use_Emitter(RD_EMIT_TIP_FE, w_theta, wf_one)
    ptr_Do_Emission => Do_Emission_Tip
  end subroutine Init_Emission_Tip

  subroutine Do_Emission_Tip(step)
    integer, intent(in) :: step
    type(rd_step_record) :: rec
    call RD_Require()
    call rd_check(rd_do_emission(rd_ctx, int(step, c_int), rd_precision, rec), 'rd_do_emission')
    call RD_Mirror(rec)
  end subroutine Do_Emission_Tip

  subroutine Clean_Up_Emission_Tip()
    call RD_Close()
  end subroutine Clean_Up_Emission_Tip
end module mod_emission_tip
