!! sv_capi.F90 -- iso_c_binding interfaces to libsunset_amd.so (the C ABI
!! declared in include/sunset_amd.h), and the device context shared by the
!! drop-in modules neighbours / labf / derivatives of this directory.
!!
!! The context is created on first use from the reference's own globals
!! (common_vars.F90): the IPART header (xmin..ymax, x/ybcond_*, smax, L_char)
!! and the real nodes 1..npfb (rp, s, h, h_small, node_type).  Orders come from
!! the same -Dmorder/-Dforder switches the reference is built with, the
!! conformation form from -Ddi/-Dlc.  Every nonzero status stops the program
!! with the library's message, as the reference stops on errors.
module sv_capi
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_vars
  implicit none

  type, bind(C) :: sv_config_t
     integer(c_int32_t) :: deriv_order, filt_order, conf_form, gpu
  end type sv_config_t

  type, bind(C) :: sv_phys_t
     real(c_double) :: L_char, U_char, Re, Wi, beta, epsPTT, Ma, Mdiff
     real(c_double) :: grav(3), driving_force(3), kol_amp, kol_k
  end type sv_phys_t

  integer(c_int32_t), parameter :: SV_OP_GRAD = 0, SV_OP_LAP = 1, SV_OP_GRAD_X = 2, &
       SV_OP_GRAD_Y = 3, SV_OP_FILTER = 4, SV_OP_FILTER_TERM = 5, SV_OP_LAP_TRANSVERSE = 6

  interface
     integer(c_int) function sv_create(cfg, out) bind(C, name='sv_create')
       import :: c_int, c_ptr, sv_config_t
       type(sv_config_t), intent(in) :: cfg
       type(c_ptr), intent(out) :: out
     end function sv_create
     subroutine sv_destroy(ctx) bind(C, name='sv_destroy')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine sv_destroy
     type(c_ptr) function sv_last_error(ctx) bind(C, name='sv_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function sv_last_error
     integer(c_int) function sv_set_domain(ctx, xmin, xmax, ymin, ymax, bc, smax, lchar) &
          bind(C, name='sv_set_domain')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), value :: xmin, xmax, ymin, ymax
       integer(c_int32_t), intent(in) :: bc(4)
       real(c_double), value :: smax, lchar
     end function sv_set_domain
     integer(c_int) function sv_set_nodes(ctx, n, x, y, s, h, hs, nt) bind(C, name='sv_set_nodes')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       real(c_double), intent(in) :: x(*), y(*), s(*), h(*), hs(*)
       integer(c_int32_t), intent(in) :: nt(*)
     end function sv_set_nodes
     integer(c_int) function sv_find_neighbours(ctx, nplink, cnt, cnts, link) &
          bind(C, name='sv_find_neighbours')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nplink
       integer(c_int32_t), intent(out) :: cnt(*), cnts(*), link(*)
     end function sv_find_neighbours
     integer(c_int) function sv_adapt_stencils(ctx, h, hs, reductions) &
          bind(C, name='sv_adapt_stencils')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, h, hs, reductions
     end function sv_adapt_stencils
     integer(c_int) function sv_calc_labf_weights(ctx, order_map) &
          bind(C, name='sv_calc_labf_weights')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, order_map
     end function sv_calc_labf_weights
     integer(c_int) function sv_calc_labf_sums(ctx) bind(C, name='sv_calc_labf_sums')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function sv_calc_labf_sums
     integer(c_int) function sv_filter_coefficients(ctx) bind(C, name='sv_filter_coefficients')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function sv_filter_coefficients
     integer(c_int) function sv_get_weights(ctx, nplink, wg, wl, wh, gs, ls, hs, vol) &
          bind(C, name='sv_get_weights')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nplink
       type(c_ptr), value :: wg, wl, wh, gs, ls, hs, vol
     end function sv_get_weights
     integer(c_int) function sv_get_ghosts(ctx, np, x, y, irel, vrel, cap) &
          bind(C, name='sv_get_ghosts')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: np
       type(c_ptr), value :: x, y, irel, vrel
       integer(c_int32_t), value :: cap
     end function sv_get_ghosts
     integer(c_int) function sv_apply(ctx, op, phi, out) bind(C, name='sv_apply')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: op
       real(c_double), intent(inout) :: phi(*)
       real(c_double), intent(out) :: out(*)
     end function sv_apply
     ! device-resident time loop, ensembles and reverse mode (sunset_amd.h)
     integer(c_int) function sv_set_phys(ctx, phys) bind(C, name='sv_set_phys')
       import :: c_int, c_ptr, sv_phys_t
       type(c_ptr), value :: ctx
       type(sv_phys_t), intent(in) :: phys
     end function sv_set_phys
     integer(c_int) function sv_set_initial(ctx, ro, u, v, cxx, cxy, cyy) &
          bind(C, name='sv_set_initial')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: ro(*), u(*), v(*), cxx(*), cxy(*), cyy(*)
     end function sv_set_initial
     integer(c_int) function sv_set_control(ctx, dt, time, itime, e1, e2, e3, dt_mode) &
          bind(C, name='sv_set_control')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, time, e1, e2, e3
       integer(c_int32_t), value :: itime, dt_mode
     end function sv_set_control
     integer(c_int) function sv_set_state(ctx, ro, rou, rov, xx, xy, yy) &
          bind(C, name='sv_set_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: ro(*), rou(*), rov(*), xx(*), xy(*), yy(*)
     end function sv_set_state
     integer(c_int) function sv_get_control(ctx, dt, time, itime, emax3, dt_cfl, dt_par) &
          bind(C, name='sv_get_control')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt, time, emax3(3), dt_cfl, dt_par
       integer(c_int32_t), intent(out) :: itime
     end function sv_get_control
     integer(c_int) function sv_calc_all_rhs(ctx, irk, r0, r1, r2, r3, r4, r5) &
          bind(C, name='sv_calc_all_rhs')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: irk
       real(c_double), intent(out) :: r0(*), r1(*), r2(*), r3(*), r4(*), r5(*)
     end function sv_calc_all_rhs
     integer(c_int) function sv_filter_variables(ctx, ro, rou, rov, xx, xy, yy) &
          bind(C, name='sv_filter_variables')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: ro(*), rou(*), rov(*), xx(*), xy(*), yy(*)
     end function sv_filter_variables
     integer(c_int) function sv_run(ctx, nsteps, use_graph) bind(C, name='sv_run')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps, use_graph
     end function sv_run
     integer(c_int) function sv_get_state(ctx, ro, rou, rov, u, v, pxx, pxy, pyy, cxx, cxy, &
          cyy) bind(C, name='sv_get_state')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, ro, rou, rov, u, v, pxx, pxy, pyy, cxx, cxy, cyy
     end function sv_get_state
     integer(c_int) function sv_ensemble_create(members, n, ens) &
          bind(C, name='sv_ensemble_create')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), intent(in) :: members(*)
       integer(c_int32_t), value :: n
       type(c_ptr), intent(out) :: ens
     end function sv_ensemble_create
     integer(c_int) function sv_ensemble_get_state(ens, m, ro, rou, rov, u, v, pxx, pxy, pyy, &
          cxx, cxy, cyy) bind(C, name='sv_ensemble_get_state')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: m
       type(c_ptr), value :: ro, rou, rov, u, v, pxx, pxy, pyy, cxx, cxy, cyy
     end function sv_ensemble_get_state
     integer(c_int) function sv_adj_setup(ctx, every) bind(C, name='sv_adj_setup')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: every
     end function sv_adj_setup
     integer(c_int) function sv_rollout_forward(ctx, nsteps, loss) &
          bind(C, name='sv_rollout_forward')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps
       real(c_double), intent(out) :: loss
     end function sv_rollout_forward
     integer(c_int) function sv_rollout_vjp(ctx, loss_bar, ubar_final, grad0) &
          bind(C, name='sv_rollout_vjp')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: loss_bar
       type(c_ptr), value :: ubar_final       ! c_null_ptr or c_loc of 6 c_ptr
       type(c_ptr), intent(in) :: grad0(6)    ! c_loc of the six output arrays
     end function sv_rollout_vjp
     integer(c_int) function sv_diagnostics(ctx, kind, member, out) &
          bind(C, name='sv_diagnostics')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kind, member    ! kind: 0 mass/energy, 1 conf, 2 Kolmogorov, 3 TG
       real(c_double), intent(out) :: out(*)
     end function sv_diagnostics
     integer(c_int) function sv_set_boundaries(ctx, rnx, rny) bind(C, name='sv_set_boundaries')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: rnx(*), rny(*)     ! rnorm(1:npfb,1), rnorm(1:npfb,2)
     end function sv_set_boundaries
     integer(c_int) function sv_set_inflow(ctx, inflow_type, velocity_control, u_start, u_end, &
          ramp_time) bind(C, name='sv_set_inflow')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: inflow_type, velocity_control
       real(c_double), value :: u_start, u_end, ramp_time
     end function sv_set_inflow
     integer(c_int) function sv_calc_boundary_weights(ctx) &
          bind(C, name='sv_calc_boundary_weights')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function sv_calc_boundary_weights
     integer(c_int) function sv_adj_param_grad(ctx, grad) bind(C, name='sv_adj_param_grad')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: grad(3)  ! d/dWi, d/dbeta, d/d(epsPTT)
     end function sv_adj_param_grad
  end interface

  type(c_ptr), save :: sv_ctx = c_null_ptr
  !! set by adapt_stencils: the device holds the lists rebuilt on the adapted
  !! h with the ghosts made from the original h, as the reference keeps the
  !! mirror particles of build_domain; the next find_neighbours only fetches
  !! them (re-uploading would remake the ghosts from the adapted h)
  logical, save :: sv_lists_adapted = .false.
  integer(ikind), save :: sv_ctx_npfb = -1

contains

  subroutine sv_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    if (rc == 0) return
    text = ''
    call c_f_pointer(sv_last_error(sv_ctx), msg, [512])
    do k = 1, 512
       if (msg(k) == c_null_char) exit
       text(k:k) = msg(k)
    end do
    write(6,*) "sunset_amd: ", where, " failed (", rc, "): ", trim(text)
    stop
  end subroutine sv_check

  !! Create the context if needed and upload the current real nodes.
  subroutine sv_upload_nodes()
    type(sv_config_t) :: cfg
    integer(c_int32_t) :: bc(4)
    integer(c_int) :: rc
    if (.not. c_associated(sv_ctx)) then
#if morder==4
       cfg%deriv_order = 4
#elif morder==6
       cfg%deriv_order = 6
#elif morder==10
       cfg%deriv_order = 10
#else
       cfg%deriv_order = 8
#endif
#if forder==4
       cfg%filt_order = 4
#elif forder==6
       cfg%filt_order = 6
#elif forder==10
       cfg%filt_order = 10
#else
       cfg%filt_order = 8
#endif
#ifdef di
       cfg%conf_form = 0
#else
       cfg%conf_form = 1
#endif
       cfg%gpu = 0
       rc = sv_create(cfg, sv_ctx)
       if (rc /= 0) then
          write(6,*) "sunset_amd: sv_create failed (", rc, "): no HIP device?"
          stop
       end if
    end if
    bc = (/ xbcond_L, xbcond_U, ybcond_L, ybcond_U /)
    call sv_check(sv_set_domain(sv_ctx, xmin, xmax, ymin, ymax, bc, smax, L_char), &
                  "sv_set_domain")
    call sv_check(sv_set_nodes(sv_ctx, npfb, rp(1:npfb,1), rp(1:npfb,2), s(1:npfb), &
                  h(1:npfb), h_small(1:npfb), node_type(1:npfb)), "sv_set_nodes")
    if (nb /= 0) call sv_check(sv_set_boundaries(sv_ctx, rnorm(1:npfb,1), rnorm(1:npfb,2)), &
                               "sv_set_boundaries")
    sv_ctx_npfb = npfb
  end subroutine sv_upload_nodes


  !! The physics of control.in as common_vars holds it (load_data.F90:21-159)
  subroutine sv_upload_phys()
    type(sv_phys_t) :: ph
    ph%L_char = L_char;ph%U_char = U_char;ph%Re = Re;ph%Wi = Wi;ph%beta = beta
    ph%epsPTT = epsPTT;ph%Ma = Ma;ph%Mdiff = Mdiff
    ph%grav = grav;ph%driving_force = driving_force
    ph%kol_amp = 0.0d0;ph%kol_k = 0.0d0
    call sv_check(sv_set_phys(sv_ctx, ph), "sv_set_phys")
  end subroutine sv_upload_phys

  !! The host state of the real nodes onto the device: sv_set_initial builds
  !! the boundary records and inflow velocities from the primitives, then the
  !! registers are replaced by the host's own values bit for bit; the control
  !! state (dt, time, itime, PID error history) follows.
  subroutine sv_upload_state(dt_mode)
    integer(c_int32_t), intent(in) :: dt_mode
    call sv_upload_phys()
    if (nb /= 0) call sv_check(sv_set_inflow(sv_ctx, inflow_type, inflow_velocity_control, &
                  u_inflow_start, u_inflow_end, ramp_time), "sv_set_inflow")
    call sv_check(sv_set_initial(sv_ctx, ro(1:npfb), u(1:npfb), v(1:npfb), cxx(1:npfb), &
                  cxy(1:npfb), cyy(1:npfb)), "sv_set_initial")
#ifdef di
    call sv_check(sv_set_state(sv_ctx, ro(1:npfb), rou(1:npfb), rov(1:npfb), cxx(1:npfb), &
                  cxy(1:npfb), cyy(1:npfb)), "sv_set_state")
#else
    call sv_check(sv_set_state(sv_ctx, ro(1:npfb), rou(1:npfb), rov(1:npfb), psixx(1:npfb), &
                  psixy(1:npfb), psiyy(1:npfb)), "sv_set_state")
#endif
    call sv_check(sv_set_control(sv_ctx, dt, time, int(itime, c_int32_t), emax_np1, emax_n, &
                  emax_nm1, dt_mode), "sv_set_control")
  end subroutine sv_upload_state
end module sv_capi
