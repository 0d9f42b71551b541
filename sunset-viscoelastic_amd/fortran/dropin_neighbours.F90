!! dropin_neighbours.F90 -- drop-in replacement of the reference's module
!! `neighbours` (source/base/neighbours.F90): same module name, same public
!! routine find_neighbours (:32-73), same outputs in common_vars (ij_count,
!! ij_count_small, ij_link(nplink,npfb); lists sorted by distance, ghosts
!! numbered > npfb).  The search runs on the device (cell lists, bit-exact
!! distance tests) through the C ABI of libsunset_amd.so.
module neighbours
  use iso_c_binding
  use kind_parameters
  use common_parameter
  use common_vars
  use sv_capi
  implicit none

  private
  public :: find_neighbours

contains
  subroutine find_neighbours
    integer(c_int32_t) :: npdev
    if (.not. sv_lists_adapted) call sv_upload_nodes()
    sv_lists_adapted = .false.
    !! ghosts were made by create_mirror_particles; the device builds the same set
    call sv_check(sv_get_ghosts(sv_ctx, npdev, c_null_ptr, c_null_ptr, c_null_ptr, &
                  c_null_ptr, 0), "sv_get_ghosts")
    if (npdev /= np) then
       write(6,*) "sunset_amd: ghost count differs from create_mirror_particles", npdev, np
       stop
    end if
    if (allocated(ij_count_small)) deallocate(ij_count_small)
    if (allocated(ij_count)) deallocate(ij_count)
    if (allocated(ij_link)) deallocate(ij_link)
    allocate(ij_count(npfb), ij_count_small(npfb))
    allocate(ij_link(nplink, npfb))
    call sv_check(sv_find_neighbours(sv_ctx, nplink, ij_count, ij_count_small, ij_link), &
                  "sv_find_neighbours")
    write(6,*) iproc, "neighbours found"
  end subroutine find_neighbours
end module neighbours
