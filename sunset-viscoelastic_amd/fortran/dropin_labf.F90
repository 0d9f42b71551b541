!! dropin_labf.F90 -- drop-in replacement of the reference's module `labf`
!! (source/base/labf.F90): same module name and public routines (:26-27).
!! calc_labf_weights (:34-231, with calc_filter_weights :233-394 and
!! calc_node_volumes :453-503), calc_labf_sums (:396-451) and
!! filter_coefficients (:1723-1811) run on the device and leave the same arrays
!! in common_vars as the reference does after each call (ij_w_grad, ij_w_lap,
!! ij_w_hyp, their sums, vol).  calc_boundary_weights (:505-1039) runs on the
!! device too and prunes the host lists of boundary nodes as the reference does;
!! ij_wb_grad2 stays on the device, where the tangential part (:755) serves the
!! drop-in calc_laplacian_transverse_only_on_bound (read by rhs_di.F90:399-408
!! and the compressible rhs.F90) and its normal part is not kept.
!! adapt_stencils (:1079-1721) runs on the device after find_neighbours,
!! writes the adapted h, h_small of the real nodes back to common_vars and
!! leaves the device lists rebuilt on them (ghosts kept from build_domain),
!! which the second find_neighbours (sunset.F90:52) fetches.
module labf
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_vars
  use sv_capi
  implicit none

  private
  public calc_labf_weights, adapt_stencils, calc_boundary_weights, calc_labf_sums, &
         filter_coefficients

contains
  subroutine fetch(with_sums)
    logical, intent(in) :: with_sums
    real(c_double), allocatable, target :: wg(:), wl(:), wh(:), gs(:), ls(:), hs(:), vo(:)
    allocate(wg(2*nplink*npfb), wl(nplink*npfb), wh(nplink*npfb), vo(npfb))
    allocate(gs(2*npfb), ls(npfb), hs(npfb))
    if (with_sums) then
       call sv_check(sv_get_weights(sv_ctx, nplink, c_loc(wg), c_loc(wl), c_loc(wh), &
                     c_loc(gs), c_loc(ls), c_loc(hs), c_loc(vo)), "sv_get_weights")
       ij_w_grad_sum = reshape(gs, (/ 2, npfb /))
       ij_w_lap_sum = ls
       ij_w_hyp_sum = hs
    else
       call sv_check(sv_get_weights(sv_ctx, nplink, c_loc(wg), c_loc(wl), c_loc(wh), &
                     c_null_ptr, c_null_ptr, c_null_ptr, c_loc(vo)), "sv_get_weights")
    end if
    ij_w_grad = reshape(wg, (/ 2, nplink, npfb /))
    ij_w_lap = reshape(wl, (/ nplink, npfb /))
    ij_w_hyp = reshape(wh, (/ nplink, npfb /))
    vol(1:npfb) = vo
  end subroutine fetch

  subroutine calc_labf_weights
    nplink = maxval(ij_count(1:npfb))                 !! labf.F90:69
    if (allocated(ij_w_grad)) deallocate(ij_w_grad)
    if (allocated(ij_w_lap)) deallocate(ij_w_lap)
    if (allocated(ij_w_hyp)) deallocate(ij_w_hyp)
    if (allocated(vol)) deallocate(vol)
    allocate(ij_w_grad(2,nplink,npfb), ij_w_lap(nplink,npfb), ij_w_hyp(nplink,npfb))
    allocate(vol(npfb))
    call sv_check(sv_calc_labf_weights(sv_ctx, c_null_ptr), "sv_calc_labf_weights")
    call fetch(.false.)
    write(6,*) "iproc", iproc, "LABFM weights calculated"
  end subroutine calc_labf_weights

  subroutine calc_labf_sums
    if (.not. allocated(ij_w_grad_sum)) allocate(ij_w_grad_sum(2,npfb))
    if (.not. allocated(ij_w_lap_sum)) allocate(ij_w_lap_sum(npfb))
    if (.not. allocated(ij_w_hyp_sum)) allocate(ij_w_hyp_sum(npfb))
    call sv_check(sv_calc_labf_sums(sv_ctx), "sv_calc_labf_sums")
    call fetch(.true.)
  end subroutine calc_labf_sums

  subroutine filter_coefficients
    call sv_check(sv_filter_coefficients(sv_ctx), "sv_filter_coefficients")
    call fetch(.true.)
    if (allocated(fd_parent)) deallocate(fd_parent)  !! as labf.F90:1812
  end subroutine filter_coefficients

  subroutine adapt_stencils
    real(c_double), allocatable, target :: ha(:), hsa(:)
    allocate(ha(npfb), hsa(npfb))
    call sv_check(sv_adapt_stencils(sv_ctx, c_loc(ha), c_loc(hsa), c_null_ptr), &
                  "sv_adapt_stencils")
    h(1:npfb) = ha
    h_small(1:npfb) = hsa
    !! the device rebuilt its lists on the adapted h; the host lists follow
    !! (refine_and_finalise_domain frees them and find_neighbours rebuilds them)
    call sv_check(sv_find_neighbours(sv_ctx, nplink, ij_count, ij_count_small, ij_link), &
                  "sv_find_neighbours")
    sv_lists_adapted = .true.
    write(6,*) iproc, "stencils adapted"
  end subroutine adapt_stencils

  subroutine calc_boundary_weights
    integer(ikind) :: jj, i, k, j, n, t
    integer(ikind), allocatable :: keep(:)
    call sv_check(sv_calc_boundary_weights(sv_ctx), "sv_calc_boundary_weights")
    !! the device pruned the lists of boundary nodes to their same-type
    !! neighbours and FD rows (labf.F90:541-566); the host lists follow it
    allocate(keep(nplink))
    do jj = 1, nb
       i = boundary_list(jj)
       t = node_type(i)
       n = 0
       do k = 1, ij_count(i)
          j = ij_link(k, i)
          if (node_type(j) == t .or. (j <= npfb .and. node_type(j) < 0 .and. &
              j + node_type(j) == i)) then
             n = n + 1
             keep(n) = j
          end if
       end do
       ij_link(:, i) = 0
       ij_link(1:n, i) = keep(1:n)
       ij_count(i) = n
       ij_count_small(i) = n
    end do
    call fetch(.false.)
    write(6,*) "finished boundary weights"
  end subroutine calc_boundary_weights
end module labf
