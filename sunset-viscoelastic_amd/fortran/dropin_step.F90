!! dropin_step.F90 -- drop-in replacement of the reference's module `step`
!! (source/base/step.F90): same module name and public routines set_tstep
!! (:569-622), set_tstep_PID (:624-682), step_rk3_4S_2R_EE (:218-512) and
!! step_rk3_4S_2R (:38-216), called by the reference's own main loop
!! (sunset.F90:75-106).  The time step runs on the device as one fused loop
!! (sv_run): the state stays device-resident between steps.
!!
!! set_tstep and set_tstep_PID only record that they were called: the device
!! evaluates both at the start of its step, from the state the step starts
!! from, as the reference does between the calls.  The mode of the next step
!! follows the calls since the last one: set_tstep_PID (the main loop) -> PID
!! control; otherwise the dt in common_vars is kept (set_tstep alone only
!! computes dt_cfl and dt_parabolic: its dt line is commented out,
!! step.F90:610-611).
!!
!! After every step dt, time, the PID error history, dt_cfl and dt_parabolic
!! are copied back to common_vars.  The fields (ro, rou, rov, u, v, psi, c, p
!! on the real nodes, the mirror ghosts refilled by the reference's own
!! reapply_mirror_bcs, u, v and c of the ghosts derived as step.F90:514-567
!! does) are copied back after every step by default; with the environment
!! variable SUNSET_AMD_STEP_SYNC=outputs only before the main loop reads them:
!! when output_to_screen prints (itime a multiple of 100), when output_layer
!! or statistics_control will fire (time past n*dt_out or m*dt_out_stats, the
!! main loop's counters mirrored here) and when time passes time_end; builds
!! with -Dtracers or -Dvpid (which read the velocity every step) always sync.
!! alpha_out (the per-node error map for output) is not kept on the device.
!! Host changes to driving_force between steps (velocity_control) are sent
!! to the device before the next step.
module step
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_vars
  use mirror_boundaries
  use conf_transforms
  use sv_capi
  implicit none

  private
  public set_tstep, step_rk3_4S_2R, step_rk3_4S_2R_EE, set_tstep_PID

  logical, save :: on_device = .false.
  logical, save :: want_pid = .false.
  logical, save :: sync_every = .true.
  integer(c_int32_t), save :: dev_mode = -1
  integer(ikind), save :: n_out_next = 1, m_out_next = 1
  real(rkind), save :: dev_force(3) = 0.0d0

contains
!! ------------------------------------------------------------------------------------------------
  subroutine set_tstep
    !! dt_cfl and dt_parabolic come back from the device after the step
  end subroutine set_tstep
!! ------------------------------------------------------------------------------------------------
  subroutine set_tstep_PID
    want_pid = .true.
  end subroutine set_tstep_PID
!! ------------------------------------------------------------------------------------------------
  subroutine step_rk3_4S_2R
    write(6,*) "sunset_amd: step_rk3_4S_2R is not provided by the drop-in step module;", &
               " the main loop's step_rk3_4S_2R_EE is"
    stop
  end subroutine step_rk3_4S_2R
!! ------------------------------------------------------------------------------------------------
  subroutine step_rk3_4S_2R_EE
    integer(c_int32_t) :: mode, it
    real(c_double) :: em(3), dtc, dtp
    character(len=32) :: policy
    integer :: plen, pstat

    if (want_pid) then
       mode = 0
    else
       mode = 1
    end if
    want_pid = .false.
    if (.not. on_device) then
       call get_environment_variable("SUNSET_AMD_STEP_SYNC", policy, plen, pstat)
       sync_every = .not. (pstat == 0 .and. policy(1:plen) == "outputs")
#if defined(tracers) || defined(vpid)
       sync_every = .true.
#endif
       call sv_upload_state(mode)
       dev_mode = mode
       dev_force = driving_force
       on_device = .true.
    else
       if (any(driving_force /= dev_force)) then
          call sv_upload_phys()
          dev_force = driving_force
       end if
       if (mode /= dev_mode) then
          call sv_check(sv_set_control(sv_ctx, dt, time, int(itime, c_int32_t), emax_np1, &
                        emax_n, emax_nm1, mode), "sv_set_control")
          dev_mode = mode
       end if
    end if

    call sv_check(sv_run(sv_ctx, 1, 0), "sv_run")
    call sv_check(sv_get_control(sv_ctx, dt, time, it, em, dtc, dtp), "sv_get_control")
    emax_np1 = em(1);emax_n = em(2);emax_nm1 = em(3)
    dt_cfl = dtc;dt_parabolic = dtp

    if (sync_every) then
       call fetch_fields()
    else if (main_loop_reads_fields()) then
       call fetch_fields()
    end if
  end subroutine step_rk3_4S_2R_EE
!! ------------------------------------------------------------------------------------------------
  !! Will the main loop (sunset.F90:75-106) read the fields before the next
  !! step?  Called after the step, before the loop increments itime.
  logical function main_loop_reads_fields()
    main_loop_reads_fields = .false.
    if (mod(itime + 1, 100) == 0) main_loop_reads_fields = .true.       !! output_to_screen
    if (time > n_out_next*dt_out) then                                    !! output_layer
       main_loop_reads_fields = .true.
       n_out_next = n_out_next + 1
    end if
    if (time > m_out_next*dt_out_stats) then                              !! statistics_control
       main_loop_reads_fields = .true.
       m_out_next = m_out_next + 1
    end if
    if (time > time_end) main_loop_reads_fields = .true.                  !! the loop ends
  end function main_loop_reads_fields
!! ------------------------------------------------------------------------------------------------
  subroutine fetch_fields()
    integer(ikind) :: i
    real(rkind) :: ooro
    real(c_double), allocatable, target :: f(:,:)
    allocate(f(npfb, 11))
    call sv_check(sv_get_state(sv_ctx, c_loc(f(1,1)), c_loc(f(1,2)), c_loc(f(1,3)), &
                  c_loc(f(1,4)), c_loc(f(1,5)), c_loc(f(1,6)), c_loc(f(1,7)), c_loc(f(1,8)), &
                  c_loc(f(1,9)), c_loc(f(1,10)), c_loc(f(1,11))), "sv_get_state")
    ro(1:npfb) = f(:,1);rou(1:npfb) = f(:,2);rov(1:npfb) = f(:,3)
    u(1:npfb) = f(:,4);v(1:npfb) = f(:,5)
#ifndef di
    psixx(1:npfb) = f(:,6);psixy(1:npfb) = f(:,7);psiyy(1:npfb) = f(:,8)
#endif
    cxx(1:npfb) = f(:,9);cxy(1:npfb) = f(:,10);cyy(1:npfb) = f(:,11)
    deallocate(f)
    !! ghosts as the reference's step leaves them (step.F90:497-510)
    call reapply_mirror_bcs
    do i=npfb+1,np
       ooro = one/ro(i)
       u(i) = rou(i)*ooro
       v(i) = rov(i)*ooro
       w(i) = row(i)*ooro
#ifndef di
       call log_conf_c_from_psi(psixx(i),psixy(i),psiyy(i),psizz(i),cxx(i),cxy(i),cyy(i),czz(i))
#endif
    end do
    !! p = csq*ro of the returned state (the reference's p is left from the
    !! last stage's calc_all_rhs; only output reads it between steps)
    do i=1,np
       p(i) = csq*ro(i)
    end do
  end subroutine fetch_fields
end module step
