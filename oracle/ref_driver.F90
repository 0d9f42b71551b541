!! ref_driver.F90 -- harness that runs the UNMODIFIED reference hot-path modules
!! (compiled by oracle/Makefile from /root/reference/source) on a node set and
!! initial state read from a binary file, and writes neighbour lists, weights,
!! derivative-operator outputs and fields after N RK3 steps to a binary file.
!!
!! TEST INFRASTRUCTURE ONLY: this is the reference side of the oracle.  It is
!! this build's own code; it only sets the reference's module globals the way
!! build_domain/initial_setup/initial_solution do (setup_domain.F90:80-374,
!! setup_flow.F90:24-173) and calls the reference's own routines in the order of
!! the main program (sunset.F90:36-106).  Nothing here alters reference source.
!!
!! usage: sunset_ref <input.bin> <output.bin> [stencil.bin]
!!   input (stream, little endian):
!!     int32(12): npfb nsteps dt_mode(0 PID,1 fixed) nplink(0=default) flags
!!                xbcL xbcU ybcL ybcU itime0 0 0
!!     real64(24): xmin xmax ymin ymax smax L_char U_char Re Wi beta epsPTT Ma
!!                 Mdiff grav(3) drive(3) dt0 time0 emax_np1 emax_n emax_nm1
!!     real64(npfb) x y s h h_small ; int32(npfb) node_type ;
!!     real64(npfb) ro u v cxx cxy cyy
!!   flags: bit 1 (2) -- take neighbour lists, weights, sums and volumes from
!!          stencil.bin instead of find_neighbours/calc_labf_weights (timing runs
!!          of the RK3 loop at sizes where the serial weight setup would take
!!          minutes); layout: int32 nplink, ij_count(npfb), ij_count_small(npfb),
!!          ij_link(nplink,npfb); real64 ij_w_grad(2,nplink,npfb),
!!          ij_w_lap(nplink,npfb), ij_w_hyp(nplink,npfb), ij_w_grad_sum(2,npfb),
!!          ij_w_lap_sum(npfb), ij_w_hyp_sum(npfb), vol(npfb)
!!          bit 2 (4) -- write only the fields after the steps (no lists/weights)
!!          bit 3 (8) -- boundary block after the fields: real64 rnx(npfb),
!!          rny(npfb); int32 inflow_type, inflow_velocity_control; real64
!!          u_inflow_start, u_inflow_end, ramp_time.  Nodes of type 0/1/2 are
!!          boundary nodes, each followed by its FD rows (types -1..-4) as the
!!          IPART reader lays them out (setup_domain.F90:179-193)
!!          bit 5 (32) -- after the operators on ro, calc_laplacian_transverse_
!!          only_on_bound of ro and of cxx and calc_divergence of (rou, rov):
!!          real64 lt_ro(npfb), lt_cxx(npfb), div(npfb) (written only with the
!!          lists, bit 2 clear)
!!          bit 4 (16) -- call adapt_stencils after the first find_neighbours
!!          and find_neighbours again after it, as the main program does
!!          without restart (sunset.F90:36-52); the output then starts with
!!          int32 npfb; real64 h(npfb), h_small(npfb); int32 ij_count(npfb),
!!          ij_count_small(npfb) as adapt_stencils left them
!!   output: see write section below (mirrored by tests/refdrv.py)
program sunset_ref
  use kind_parameters
  use common_parameter
  use common_vars
  use neighbours
  use labf
  use mirror_boundaries
  use derivatives
  use conf_transforms
  use step
  use omp_lib
  implicit none
  character(len=512) :: fin,fout,fst
  integer(ikind) :: hdr(12),nsteps,dt_mode,flags,i,k,nm,itime0
  real(rkind) :: rh(24)
  real(rkind),allocatable :: tmp(:),tro(:),tu(:),tv(:),tcxx(:),tcxy(:),tcyy(:)
  real(rkind),allocatable :: g2(:,:),l1(:),f1(:),lt1(:),lt2(:),dv(:)
  real(rkind),allocatable :: ah(:),ahs(:)
  integer(ikind),allocatable :: itmp(:),ac(:),acs(:)
  integer(ikind) :: ib2(2)
  real(rkind) :: rb3(3)
  real(rkind) :: t0,t1

  call get_command_argument(1,fin)
  call get_command_argument(2,fout)
  call get_command_argument(3,fst)
  open(30,file=trim(fin),access='stream',form='unformatted',status='old')
  read(30) hdr
  read(30) rh
  npfb = hdr(1);nsteps = hdr(2);dt_mode = hdr(3);flags = hdr(5)
  xbcond_L = hdr(6);xbcond_U = hdr(7);ybcond_L = hdr(8);ybcond_U = hdr(9)
  itime0 = hdr(10)
  xmin = rh(1);xmax = rh(2);ymin = rh(3);ymax = rh(4)
  L_char = rh(6);U_char = rh(7);Time_char = L_char/U_char
  Re = rh(8);Wi = rh(9);beta = rh(10);epsPTT = rh(11);Ma = rh(12);Mdiff = rh(13)
  grav = rh(14:16);driving_force = rh(17:19)

  !! As initial_setup / build_domain (setup_domain.F90:50, :105, :158-202)
  nplink = 2*4*ceiling(ss*hovs)**2
  if(hdr(4).gt.0) nplink = hdr(4)
  nb = 0;iproc = 0;nprocs = 1;nz = 1;nz_global = 1
  segment_time_local = zero
  smax = rh(5);h0 = hovs*smax;sup_size = ss*h0;h2 = h0*h0;h3 = h2*h0
  nm = 12
  allocate(rp(nm*npfb,ithree),rnorm(nm*npfb,ithree),h(nm*npfb),s(nm*npfb))
  rp = zero;rnorm = zero
  allocate(node_type(nm*npfb),fd_parent(nm*npfb),global_index(nm*npfb))
  node_type = 0;fd_parent = 0
  allocate(h_small(nm*npfb))
  allocate(tmp(npfb),itmp(npfb))
  read(30) tmp;rp(1:npfb,1) = tmp
  read(30) tmp;rp(1:npfb,2) = tmp
  read(30) tmp;s(1:npfb) = tmp
  read(30) tmp;h(1:npfb) = tmp
  read(30) tmp;h_small(1:npfb) = tmp
  read(30) itmp;node_type(1:npfb) = itmp
  do i=1,npfb
     global_index(i) = i
  end do
  allocate(tro(npfb),tu(npfb),tv(npfb),tcxx(npfb),tcxy(npfb),tcyy(npfb))
  read(30) tro;read(30) tu;read(30) tv;read(30) tcxx;read(30) tcxy;read(30) tcyy
  if(iand(flags,8).ne.0) then
     read(30) tmp;rnorm(1:npfb,1) = tmp
     read(30) tmp;rnorm(1:npfb,2) = tmp
     read(30) ib2
     read(30) rb3
     inflow_type = ib2(1);inflow_velocity_control = ib2(2)
     u_inflow_start = rb3(1);u_inflow_end = rb3(2);ramp_time = rb3(3)
     do i=1,npfb
        if(node_type(i).lt.0) then
           fd_parent(i) = i + node_type(i)
           rnorm(i,:) = rnorm(fd_parent(i),:)
        end if
     end do
  end if
  close(30)
  npfb_layer = npfb;npfb_global = npfb

  call create_mirror_particles
  !! A corner between a periodic and a no-slip edge makes a ghost whose
  !! position and vrelation create_mirror_particles never writes
  !! (mirror_boundaries.F90:149-240 set them only for the periodic/symmetric
  !! pairs); the position then holds what the allocation held (zero here),
  !! the vrelation any value, and reapply_mirror_bcs leaves such a ghost's
  !! momentum unset.  Define it as a plain copy of the parent (vrelation 1),
  !! as the oracle reads it.
  do i=npfb+1,np
     if(vrelation(i).lt.1.or.vrelation(i).gt.4) vrelation(i) = 1
  end do
  !! boundary and internal lists as build_domain (setup_domain.F90:314-346)
  nb = 0
  do i=1,npfb
     if(node_type(i).ge.0.and.node_type(i).le.2) nb = nb + 1
  end do
  allocate(internal_list(npfb-nb))
  if(nb.ne.0) allocate(boundary_list(nb),znf_vdiff(nb),znf_vtdiff(nb))
  k = 0;nm = 0
  do i=1,npfb
     if(node_type(i).lt.0.or.node_type(i).eq.999.or.node_type(i).eq.998) then
        k = k + 1;internal_list(k) = i
     else
        nm = nm + 1;boundary_list(nm) = i
        znf_vdiff(nm) = node_type(i).eq.1
        znf_vtdiff(nm) = node_type(i).eq.2
     end if
  end do

  if(iand(flags,2).ne.0) then
     open(32,file=trim(fst),access='stream',form='unformatted',status='old')
     read(32) nplink
     allocate(ij_count(npfb),ij_count_small(npfb),ij_link(nplink,npfb))
     allocate(ij_w_grad(2,nplink,npfb),ij_w_lap(nplink,npfb),ij_w_hyp(nplink,npfb))
     allocate(ij_w_grad_sum(2,npfb),ij_w_lap_sum(npfb),ij_w_hyp_sum(npfb),vol(npfb))
     read(32) ij_count,ij_count_small,ij_link
     read(32) ij_w_grad,ij_w_lap,ij_w_hyp,ij_w_grad_sum,ij_w_lap_sum,ij_w_hyp_sum,vol
     close(32)
  else
     t0 = omp_get_wtime()
     call find_neighbours
     t1 = omp_get_wtime()
     write(6,*) "REFTIME find_neighbours",t1-t0
     if(iand(flags,16).ne.0) then
        t0 = omp_get_wtime()
        call adapt_stencils
        t1 = omp_get_wtime()
        write(6,*) "REFTIME adapt_stencils",t1-t0
        allocate(ah(npfb),ahs(npfb),ac(npfb),acs(npfb))
        ah = h(1:npfb);ahs = h_small(1:npfb)
        ac = ij_count(1:npfb);acs = ij_count_small(1:npfb)
        !! refine_and_finalise_domain (setup_domain.F90:373-386, non-MPI)
        !! frees the lists, then they are rebuilt with the new h
        !! (sunset.F90:49-52); the lists read only the real nodes' h
        deallocate(ij_link,ij_count)
        call find_neighbours
     end if
     t0 = omp_get_wtime()
     call calc_labf_weights
     t1 = omp_get_wtime()
     write(6,*) "REFTIME calc_labf_weights",t1-t0
     if(nb.ne.0) then
        call calc_boundary_weights
        !! calc_labf_weights skips boundary nodes and calc_boundary_weights
        !! writes ij_w_lap only for walls: the Laplacian weights of inflow and
        !! outflow nodes are whatever the allocation held.  Zero them, as a
        !! fresh zero-filled allocation would give.
        do k=1,nb
           i = boundary_list(k)
           if(node_type(i).ne.0) ij_w_lap(:,i) = zero
        end do
     end if
     call calc_labf_sums
     call filter_coefficients
  end if

  !! load_control_data_all derived quantities (load_data.F90:132-148)
  rho_char = one
  visc_total = rho_char*U_char*L_char/Re
  visc_solvent = beta*visc_total
  visc_polymeric = (one-beta)*visc_total
  lambda = Wi*L_char/U_char
  csq = (u_char/Ma)**two
  p_inflow = rho_char*csq;p_outflow = rho_char*csq      !! load_data.F90:147-148
  L_domain_x = (xmax - xmin)*L_char;L_domain_y = (ymax - ymin)*L_char   !! setup_domain.F90:98-99
  fenep_l2 = 1.0d30

  !! initial_solution (setup_flow.F90:24-173), fields from the input file
  allocate(rou(np),rov(np),row(np),ro(np))
  rou=zero;rov=zero;row=zero;ro=one
  allocate(psixx(np),psixy(np),psiyy(np),cxx(np),cxy(np),cyy(np))
  allocate(psixz(np),psiyz(np),psizz(np),cxz(np),cyz(np),czz(np))
  psixx=zero;psixy=zero;psiyy=zero;psixz=zero;psiyz=zero;psizz=zero
  allocate(p(np),u(np),v(np),w(np),alpha_out(np))
  p=zero;u=zero;v=zero;w=zero;alpha_out=zero
  cxx=one;cxy=zero;cyy=one;cxz=zero;cyz=zero;czz=one
  do i=1,npfb
     ro(i) = tro(i);u(i) = tu(i);v(i) = tv(i)
     cxx(i) = tcxx(i);cxy(i) = tcxy(i);cyy(i) = tcyy(i)
     p(i) = ro(i)*csq
#ifdef lc
     call log_conf_psi_from_c(cxx(i),cxy(i),cyy(i),czz(i),psixx(i),psixy(i),psiyy(i),psizz(i))
#endif
  end do
  do i=1,np
     rou(i) = ro(i)*u(i);rov(i) = ro(i)*v(i);row(i) = ro(i)*w(i)
  end do
  call reapply_mirror_bcs
#ifdef lc
  do i=1,np
     call log_conf_c_from_psi(psixx(i),psixy(i),psiyy(i),psizz(i),cxx(i),cxy(i),cyy(i),czz(i))
  end do
#endif
  do i=npfb+1,np
     u(i)=rou(i)/ro(i);v(i)=rov(i)/ro(i);w(i)=row(i)/ro(i)
  end do
  if(nb.ne.0) then   !! setup_flow.F90:48-51, :130-137
     allocate(u_inflow_local(nb));u_inflow_local = u_char
     do k=1,nb
        i = boundary_list(k)
        if(node_type(i).eq.1) u_inflow_local(k) = u(i)
     end do
  end if
  dt = rh(20);time = rh(21)
  emax_np1 = rh(22);emax_n = rh(23);emax_nm1 = rh(24)
  ero_norm = one;erou_norm = one/(one*u_char);exx_norm = one
  itime = itime0

  !! Derivative operators on the initial density (derivatives.F90:26-116, :466-511)
  allocate(g2(npfb,ithree),l1(npfb),f1(npfb),lt1(npfb),lt2(npfb),dv(npfb))
  call calc_gradient(ro,g2)
  call calc_laplacian(ro,l1)
  call calc_filter_term(ro,f1)
  if(iand(flags,32).ne.0) then   !! derivatives.F90:306-358
     call calc_laplacian_transverse_only_on_bound(ro,lt1)
     call calc_laplacian_transverse_only_on_bound(cxx,lt2)
     call calc_divergence(rou,rov,dv)         !! derivatives.F90:119-189
  end if

  open(31,file=trim(fout),access='stream',form='unformatted',status='replace')
  if(iand(flags,16).ne.0) then
     write(31) npfb
     write(31) ah,ahs
     write(31) ac,acs
  end if
  if(iand(flags,4).eq.0) then
  write(31) np,nplink
  write(31) rp(1:np,1),rp(1:np,2)
  write(31) irelation(npfb+1:np),vrelation(npfb+1:np)
  write(31) ij_count(1:npfb),ij_count_small(1:npfb)
  write(31) ij_link(1:nplink,1:npfb)
  write(31) ij_w_grad(1:2,1:nplink,1:npfb),ij_w_lap(1:nplink,1:npfb),ij_w_hyp(1:nplink,1:npfb)
  write(31) ij_w_grad_sum(1:2,1:npfb),ij_w_lap_sum(1:npfb),ij_w_hyp_sum(1:npfb),vol(1:npfb)
  write(31) g2(1:npfb,1),g2(1:npfb,2),l1(1:npfb),f1(1:npfb)
  if(iand(flags,32).ne.0) write(31) lt1(1:npfb),lt2(1:npfb),dv(1:npfb)
  end if

  !! Time loop (sunset.F90:75-106) without output/statistics
  t0 = omp_get_wtime()
  do k=1,nsteps
     t1 = omp_get_wtime()
     call set_tstep
     if(dt_mode.eq.0) call set_tstep_PID
     call step_rk3_4S_2R_EE
     itime = itime + 1
     write(6,*) "REFSTEP",k,omp_get_wtime()-t1
  end do
  t1 = omp_get_wtime()
  write(6,*) "REFTIME rk3_steps",nsteps,t1-t0

  write(31) ro(1:npfb),rou(1:npfb),rov(1:npfb),u(1:npfb),v(1:npfb)
  write(31) psixx(1:npfb),psixy(1:npfb),psiyy(1:npfb)
  write(31) cxx(1:npfb),cxy(1:npfb),cyy(1:npfb)
  write(31) dt,time,emax_np1,emax_n,emax_nm1,dt_cfl,dt_parabolic
  write(31) itime
  close(31)
end program sunset_ref
