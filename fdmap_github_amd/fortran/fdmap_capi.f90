! fdmap_capi.f90 -- iso_c_binding mirror of include/fdmap.h.
!
! The derived types below have the exact C layout of the structures in
! include/fdmap.h; the interfaces bind the library entry points.  This is
! the binding the Fortran driver uses to call the MI355X library in place
! of the reference's time_step_LS (time_step.f90:71-167).

module fdmap_capi

  use iso_c_binding
  implicit none

  integer(c_int32_t),parameter :: FDM_MAX_STAGES = 8, FDM_MAX_SOURCE_PLACES = 16

  integer(c_int),parameter :: FDM_OK = 0, FDM_EINVAL = -1, FDM_EDEVICE = -2, &
       FDM_ENOMEM = -3, FDM_ENEWTON = -4
  integer(c_int32_t),parameter :: FDM_FP64 = 0, FDM_FP32_BULK = 1
  integer,parameter :: FDM_SNAPSHOT_SLOTS = 4

  integer(c_int32_t),parameter :: FDM_BC_NONE = 0, FDM_BC_ABSORBING = 1, FDM_BC_FREE = 2, &
       FDM_BC_RIGID = 3, FDM_BC_ABSORBING_VELOCITY = 4, FDM_BC_FREE_VELOCITY = 5, &
       FDM_BC_ABSORBING_0 = 6, FDM_BC_FREE_0 = 7, FDM_BC_RIGID_0 = 8, FDM_BC_MMS_SIN_W = 9

  integer(c_int32_t),parameter :: FDM_COUPLING_LOCKED = 0, FDM_COUPLING_FRICTION = 1

  integer(c_int32_t),parameter :: FDM_LAW_FRICTIONLESS = 0, FDM_LAW_LOCKED = 1, FDM_LAW_SW = 2, &
       FDM_LAW_SL = 3, FDM_LAW_FL = 4, FDM_LAW_RSL = 5, FDM_LAW_RFL = 6, FDM_LAW_RSF = 7, &
       FDM_LAW_RAL = 8

  integer(c_int32_t),parameter :: FDM_STF_RAMP = 0, FDM_STF_SINUSOID = 1, FDM_STF_GAUSSIAN = 2, &
       FDM_STF_RICKER = 3

  integer(c_int32_t),parameter :: FDM_IF_D = 0, FDM_IF_PSI = 1, FDM_IF_DPSI = 2, FDM_IF_TRUP = 3, &
       FDM_IF_V = 4, FDM_IF_O = 5, FDM_IF_S = 6, FDM_IF_N = 7, FDM_IF_S0 = 8, FDM_IF_N0 = 9, &
       FDM_IF_DS = 10, FDM_IF_DN = 11, FDM_IF_DDS = 12, FDM_IF_DDN = 13, FDM_IF_W = 14, &
       FDM_IF_DW = 15, FDM_IF_NREG = 16, FDM_IF_DNREG = 17, FDM_IF_NFIELDS = 18

  type,bind(C) :: fdm_side
     integer(c_int32_t) :: bc, n
     type(c_ptr) :: normal, M, F0, x, y
  end type fdm_side

  type,bind(C) :: fdm_block
     integer(c_int32_t) :: mx, px, my, py, plastic, mms, id, pad_
     real(c_double) :: rho, G, M, gamma, nu
     real(c_double) :: mu, beta, b, h, eta, K
     real(c_double) :: s0(6)
     type(fdm_side) :: side(4)
     real(c_double) :: Cdiss = 0.0_c_double
     integer(c_int32_t) :: pmlx = 0, pmly = 0
     real(c_double) :: pml(7) = 0.0_c_double
  end type fdm_block

  type,bind(C) :: fdm_iface
     integer(c_int32_t) :: direction, blockm, blockp, m, p, coupling, law, reg, rup_field, pad_
     real(c_double) :: rup_threshold, skempton, flock, ld_T
     type(c_ptr) :: nhat, rs, sw, S0base, N0base, ldS, ldN, outside, Psi0
     integer(c_int32_t) :: force = 0, pad2_ = 0
     real(c_double) :: kn(8) = 0.0_c_double
     type(c_ptr) :: kn_x = c_null_ptr
  end type fdm_iface

  type,bind(C) :: fdm_source_place
     integer(c_int32_t) :: block, i0, j0, pad_
     real(c_double) :: dx(5), dy(5)
  end type fdm_source_place

  type,bind(C) :: fdm_source
     integer(c_int32_t) :: nplace, stf
     real(c_double) :: t_center, t_duration, Fx, Fy, Mxx, Mxy, Myy
     type(fdm_source_place) :: place(FDM_MAX_SOURCE_PLACES)
  end type fdm_source

  type,bind(C) :: fdm_problem
     integer(c_int32_t) :: nx, ny, nblocks, nifaces, precision, nstage
     real(c_double) :: rkA(FDM_MAX_STAGES), rkB(FDM_MAX_STAGES), rkC(FDM_MAX_STAGES+1)
     type(c_ptr) :: xq, xr, yq, yr, Ji, x, y
     type(c_ptr) :: blocks, ifaces
     type(fdm_source) :: source
     type(c_ptr) :: mat_rho = c_null_ptr, mat_cs = c_null_ptr, mat_cp = c_null_ptr
     type(c_ptr) :: J = c_null_ptr
     integer(c_int32_t) :: mode = 2, sbp = 4
  end type fdm_problem

  type,bind(C) :: fdm_profile_t
     real(c_double) :: bulk_ms
     integer(c_int64_t) :: bulk_launches
     real(c_double) :: edge_ms
     integer(c_int64_t) :: edge_launches, bulk_points
     real(c_double) :: band_ms
     integer(c_int64_t) :: band_launches
  end type fdm_profile_t

  interface

     function fdm_create(prob,ctx) bind(C,name='fdm_create') result(rc)
       import :: fdm_problem, c_ptr, c_int
       type(fdm_problem),intent(in) :: prob
       type(c_ptr),intent(out) :: ctx
       integer(c_int) :: rc
     end function fdm_create

     subroutine fdm_destroy(ctx) bind(C,name='fdm_destroy')
       import :: c_ptr
       type(c_ptr),value :: ctx
     end subroutine fdm_destroy

     function fdm_last_error(ctx) bind(C,name='fdm_last_error') result(msg)
       import :: c_ptr
       type(c_ptr),value :: ctx
       type(c_ptr) :: msg
     end function fdm_last_error

     function fdm_set_fields(ctx,F,DF) bind(C,name='fdm_set_fields') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx, F, DF
       integer(c_int) :: rc
     end function fdm_set_fields

     function fdm_get_fields(ctx,F,DF) bind(C,name='fdm_get_fields') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx, F, DF
       integer(c_int) :: rc
     end function fdm_get_fields

     function fdm_snapshot(ctx,mask,slot) bind(C,name='fdm_snapshot') result(rc)
       ! mask: uint32_t in C; the field masks used stay below 2**31
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr),value :: ctx
       integer(c_int32_t),value :: mask
       integer(c_int32_t) :: slot
       integer(c_int) :: rc
     end function fdm_snapshot

     function fdm_snapshot_read(ctx,slot,out) bind(C,name='fdm_snapshot_read') result(rc)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr),value :: ctx
       integer(c_int32_t),value :: slot
       type(c_ptr),value :: out
       integer(c_int) :: rc
     end function fdm_snapshot_read

     function fdm_init_edges(ctx,t) bind(C,name='fdm_init_edges') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr),value :: ctx
       real(c_double),value :: t
       integer(c_int) :: rc
     end function fdm_init_edges

     function fdm_step(ctx,t,dt,n_first,nsteps) bind(C,name='fdm_step') result(rc)
       import :: c_ptr, c_int, c_double, c_int64_t, c_int32_t
       type(c_ptr),value :: ctx
       real(c_double),value :: t, dt
       integer(c_int64_t),value :: n_first
       integer(c_int32_t),value :: nsteps
       integer(c_int) :: rc
     end function fdm_step

     function fdm_stage1(ctx,t,dt,n) bind(C,name='fdm_stage1') result(rc)
       import :: c_ptr, c_int, c_double, c_int64_t
       type(c_ptr),value :: ctx
       real(c_double),value :: t, dt
       integer(c_int64_t),value :: n
       integer(c_int) :: rc
     end function fdm_stage1

     function fdm_get_iface(ctx,iface,field,out) bind(C,name='fdm_get_iface') result(rc)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr),value :: ctx, out
       integer(c_int32_t),value :: iface, field
       integer(c_int) :: rc
     end function fdm_get_iface

     function fdm_get_iface_stage1(ctx,iface,field,out) bind(C,name='fdm_get_iface_stage1') result(rc)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr),value :: ctx, out
       integer(c_int32_t),value :: iface, field
       integer(c_int) :: rc
     end function fdm_get_iface_stage1

     function fdm_get_plastic(ctx,gammap,lambda) bind(C,name='fdm_get_plastic') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx, gammap, lambda
       integer(c_int) :: rc
     end function fdm_get_plastic

     function fdm_set_plastic(ctx,gammap,lambda) bind(C,name='fdm_set_plastic') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx, gammap, lambda
       integer(c_int) :: rc
     end function fdm_set_plastic

     function fdm_get_pml(ctx,Wx,Wy) bind(C,name='fdm_get_pml') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx, Wx, Wy
       integer(c_int) :: rc
     end function fdm_get_pml

     function fdm_set_pml(ctx,Wx,Wy) bind(C,name='fdm_set_pml') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx, Wx, Wy
       integer(c_int) :: rc
     end function fdm_set_pml

     function fdm_set_iface(ctx,iface,field,in) bind(C,name='fdm_set_iface') result(rc)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr),value :: ctx, in
       integer(c_int32_t),value :: iface, field
       integer(c_int) :: rc
     end function fdm_set_iface

     function fdm_max_slip_velocity(ctx,vmax) bind(C,name='fdm_max_slip_velocity') result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr),value :: ctx
       real(c_double),intent(out) :: vmax
       integer(c_int) :: rc
     end function fdm_max_slip_velocity

     function fdm_sync(ctx) bind(C,name='fdm_sync') result(rc)
       import :: c_ptr, c_int
       type(c_ptr),value :: ctx
       integer(c_int) :: rc
     end function fdm_sync

     function c_strlen(s) bind(C,name='strlen') result(n)
       import :: c_ptr, c_size_t
       type(c_ptr),value :: s
       integer(c_size_t) :: n
     end function c_strlen

  end interface

contains

  function fdm_error_message(ctx) result(msg)
    ! Fortran string of fdm_last_error(ctx)
    type(c_ptr),intent(in) :: ctx
    character(len=:),allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char),pointer :: chars(:)
    integer :: n,k
    p = fdm_last_error(ctx)
    if (.not.c_associated(p)) then
       msg = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p,chars,[n])
    allocate(character(len=n) :: msg)
    do k = 1,n
       msg(k:k) = chars(k)
    end do
  end function fdm_error_message

end module fdmap_capi
