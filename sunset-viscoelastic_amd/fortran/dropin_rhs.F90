!! dropin_rhs.F90 -- drop-in replacement of the reference's module `rhs`
!! (source/base/rhs_lc.F90, rhs_di.F90): same module name and public routines
!! calc_all_rhs (rhs_lc.F90:47-117) and filter_variables (:818-893), for a
!! build that keeps the reference's own step module.  The reference's step
!! owns the state on the host, so every call sends the real nodes' registers
!! (ro, rou, rov, psi or c) and the control state (time, itime, iRKstep) to the
!! device and returns the six right-hand sides, or the filtered fields.
!!
!! calc_all_rhs covers node sets without boundary nodes (the device's
!! right-hand-side-only sweep is the interior branch); with boundary nodes the
!! step drop-in (dropin_step.F90) runs the whole step, boundaries included.
!! The pressure p = csq*ro is set on all np nodes, as calc_all_rhs does.
module rhs
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_vars
  use sv_capi
  implicit none

  private
  public calc_all_rhs, filter_variables

  logical, save :: on_device = .false.

contains
!! ------------------------------------------------------------------------------------------------
  subroutine calc_all_rhs
    integer(ikind) :: i
    if (nb /= 0) then
       write(6,*) "sunset_amd: calc_all_rhs drop-in covers node sets without boundary nodes;", &
                  " use the step drop-in (dropin_step.F90) with boundaries"
       stop
    end if
    if (.not. on_device) then
       call sv_upload_state(1_c_int32_t)
       on_device = .true.
    else
#ifdef di
       call sv_check(sv_set_state(sv_ctx, ro(1:npfb), rou(1:npfb), rov(1:npfb), cxx(1:npfb), &
                     cxy(1:npfb), cyy(1:npfb)), "sv_set_state")
#else
       call sv_check(sv_set_state(sv_ctx, ro(1:npfb), rou(1:npfb), rov(1:npfb), &
                     psixx(1:npfb), psixy(1:npfb), psiyy(1:npfb)), "sv_set_state")
#endif
       call sv_check(sv_set_control(sv_ctx, dt, time, int(itime, c_int32_t), emax_np1, &
                     emax_n, emax_nm1, 1_c_int32_t), "sv_set_control")
    end if
    rhs_row = zero
    rhs_xz = zero;rhs_yz = zero;rhs_zz = zero
    call sv_check(sv_calc_all_rhs(sv_ctx, int(iRKstep, c_int32_t), rhs_ro, rhs_rou, rhs_rov, &
                  rhs_xx, rhs_xy, rhs_yy), "sv_calc_all_rhs")
    do i=1,np
       p(i) = csq*ro(i)
    end do
  end subroutine calc_all_rhs
!! ------------------------------------------------------------------------------------------------
  subroutine filter_variables
#ifdef di
    call sv_check(sv_filter_variables(sv_ctx, ro, rou, rov, cxx, cxy, cyy), &
                  "sv_filter_variables")
#else
    call sv_check(sv_filter_variables(sv_ctx, ro, rou, rov, psixx, psixy, psiyy), &
                  "sv_filter_variables")
#endif
  end subroutine filter_variables
end module rhs
