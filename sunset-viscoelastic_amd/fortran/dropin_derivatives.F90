!! dropin_derivatives.F90 -- drop-in replacement of the reference's module
!! `derivatives` (source/base/derivatives.F90): same module name, same routines
!! and argument meaning (phi sized np with ghost values, outputs on 1..npfb,
!! third gradient column zero in 2D), same profiling buckets of
!! segment_time_local.  The sums run on the device (SELL-64 stencil apply);
!! neighbours and labf from this directory must have built the stencils.
module derivatives
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_vars
  use omp_lib
  use sv_capi
  implicit none

  real(rkind) :: segment_dtstart, segment_dtend

contains
  subroutine apply_op(op, phi, out)
    integer(c_int32_t), intent(in) :: op
    real(rkind), dimension(:), intent(in) :: phi
    real(c_double), dimension(:), intent(out) :: out
    real(c_double), allocatable :: work(:)
    allocate(work(np))
    work = phi(1:np)
    call sv_check(sv_apply(sv_ctx, op, work, out), "sv_apply")
  end subroutine apply_op

  subroutine calc_gradient(phi, gradphi)                 !! derivatives.F90:26-71
    real(rkind), dimension(:), intent(in) :: phi
    real(rkind), dimension(:,:), intent(inout) :: gradphi
    real(c_double), allocatable :: out(:)
    segment_dtstart = omp_get_wtime()
    allocate(out(2*npfb))
    call apply_op(SV_OP_GRAD, phi, out)
    gradphi(1:npfb,1) = out(1:npfb)
    gradphi(1:npfb,2) = out(npfb+1:2*npfb)
    gradphi(:,3) = zero
    segment_dtend = omp_get_wtime()
    segment_time_local(5) = segment_time_local(5) + segment_dtend - segment_dtstart
  end subroutine calc_gradient

  subroutine calc_laplacian(phi, lapphi)                 !! derivatives.F90:73-116
    real(rkind), dimension(:), intent(in) :: phi
    real(rkind), dimension(:), intent(inout) :: lapphi
    real(c_double), allocatable :: out(:)
    segment_dtstart = omp_get_wtime()
    allocate(out(npfb))
    call apply_op(SV_OP_LAP, phi, out)
    lapphi(1:npfb) = out
    segment_dtend = omp_get_wtime()
    segment_time_local(6) = segment_time_local(6) + segment_dtend - segment_dtstart
  end subroutine calc_laplacian

  subroutine calc_gradient_only(dim, phi, gradphi)       !! derivatives.F90:416-464
    integer(ikind), intent(in) :: dim
    real(rkind), dimension(:), intent(in) :: phi
    real(rkind), dimension(:,:), intent(inout) :: gradphi
    real(c_double), allocatable :: out(:)
    segment_dtstart = omp_get_wtime()
    if (dim <= 2) then
       allocate(out(npfb))
       if (dim == 1) call apply_op(SV_OP_GRAD_X, phi, out)
       if (dim == 2) call apply_op(SV_OP_GRAD_Y, phi, out)
       gradphi(1:npfb,dim) = out
    else
       gradphi(:,3) = zero
    end if
    segment_dtend = omp_get_wtime()
    segment_time_local(5) = segment_time_local(5) + segment_dtend - segment_dtstart
  end subroutine calc_gradient_only

  subroutine calc_filtered_var(phi)                      !! derivatives.F90:360-414
    real(rkind), dimension(:), intent(inout) :: phi
    real(c_double), allocatable :: work(:), dummy(:)
    segment_dtstart = omp_get_wtime()
    allocate(work(np), dummy(1))
    work = phi(1:np)
    call sv_check(sv_apply(sv_ctx, SV_OP_FILTER, work, dummy), "sv_apply")
    phi(1:npfb) = work(1:npfb)
    segment_dtend = omp_get_wtime()
    segment_time_local(7) = segment_time_local(7) + segment_dtend - segment_dtstart
  end subroutine calc_filtered_var

  subroutine calc_filter_term(phi, filtphi)              !! derivatives.F90:466-511
    real(rkind), dimension(:), intent(in) :: phi
    real(rkind), dimension(:), intent(out) :: filtphi
    real(c_double), allocatable :: out(:)
    segment_dtstart = omp_get_wtime()
    allocate(out(npfb))
    call apply_op(SV_OP_FILTER_TERM, phi, out)
    filtphi(1:npfb) = out
    segment_dtend = omp_get_wtime()
    segment_time_local(7) = segment_time_local(7) + segment_dtend - segment_dtstart
  end subroutine calc_filter_term

  !! derivatives.F90:306-358: the Laplacian on internal nodes, the tangential
  !! second derivative only on boundary nodes (their 1D LABFM weights from
  !! calc_boundary_weights, kept on the device)
  subroutine calc_laplacian_transverse_only_on_bound(phi, lapphi)
    real(rkind), dimension(:), intent(in) :: phi
    real(rkind), dimension(:), intent(inout) :: lapphi
    real(c_double), allocatable :: out(:)
    segment_dtstart = omp_get_wtime()
    allocate(out(npfb))
    call apply_op(SV_OP_LAP_TRANSVERSE, phi, out)
    lapphi(1:npfb) = out
    segment_dtend = omp_get_wtime()
    segment_time_local(6) = segment_time_local(6) + segment_dtend - segment_dtstart
  end subroutine calc_laplacian_transverse_only_on_bound

  !! derivatives.F90:119-189 (2D): div = d(phi1)/dx + d(phi2)/dy on the real
  !! nodes; on wall nodes, whose first-derivative weights are in the boundary
  !! frame (normal, tangent), n . (d(phi1)/dn, d(phi2)/dn) (:159-181)
  subroutine calc_divergence(phi1, phi2, divphi)
    real(rkind), dimension(:), intent(in) :: phi1, phi2
    real(rkind), dimension(:), intent(inout) :: divphi
    real(c_double), allocatable :: d1(:), d2(:), dn2(:)
    integer(ikind) :: jj, i
    segment_dtstart = omp_get_wtime()
    allocate(d1(npfb), d2(npfb))
    call apply_op(SV_OP_GRAD_X, phi1, d1)
    call apply_op(SV_OP_GRAD_Y, phi2, d2)
    divphi(1:npfb) = d1 + d2
    if (nb /= 0) then
       allocate(dn2(npfb))
       call apply_op(SV_OP_GRAD_X, phi2, dn2)
       do jj = 1, nb
          i = boundary_list(jj)
          if (node_type(i) == 0) divphi(i) = rnorm(i,1)*d1(i) + rnorm(i,2)*dn2(i)
       end do
    end if
    segment_dtend = omp_get_wtime()
    segment_time_local(5) = segment_time_local(5) + segment_dtend - segment_dtstart
  end subroutine calc_divergence

  !! The second-derivative boundary operators (derivatives.F90:191-304) serve
  !! only the compressible rhs.F90: without boundary nodes they do nothing (as
  !! the reference's loops over boundary_list), with them the drop-in stops.
  subroutine calc_grad2bound(phi, grad2phi)              !! derivatives.F90:191-238
    real(rkind), dimension(:), intent(in) :: phi
    real(rkind), dimension(:,:), intent(inout) :: grad2phi
    call no_boundary_operator("calc_grad2bound")
  end subroutine calc_grad2bound

  subroutine calc_grad2vecbound(phi1, phi2, phi3, g2phi) !! derivatives.F90:240-270
    real(rkind), dimension(:), intent(in) :: phi1, phi2, phi3
    real(rkind), dimension(:,:), intent(inout) :: g2phi
    call no_boundary_operator("calc_grad2vecbound")
  end subroutine calc_grad2vecbound

  subroutine calc_grad2crossbound(gradphi, g2phi)        !! derivatives.F90:272-304
    real(rkind), dimension(:,:), intent(in) :: gradphi
    real(rkind), dimension(:,:), intent(inout) :: g2phi
    call no_boundary_operator("calc_grad2crossbound")
  end subroutine calc_grad2crossbound

  subroutine no_boundary_operator(name)
    character(len=*), intent(in) :: name
    if (nb /= 0) then
       write(6,*) "sunset_amd: ", name, " with boundary nodes has no device implementation"
       stop
    end if
  end subroutine no_boundary_operator
end module derivatives
