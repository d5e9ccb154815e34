! fdmap_setup.f90 -- reads an FDMAP .in problem file and builds the
! device problem description (fdm_problem, include/fdmap.h).
!
! One process holds every block.  The namelists and their defaults are
! those of the reference (main.f90:29, domain.f90:84-117, grid.f90:51,
! boundaries.f90:32, material.f90:47, fields.f90:90, source.f90:41,
! interfaces.f90:44, friction.f90:79); sections are located with the same
! '!---BLOCKn---' / '!---IFACEn---' markers (io.f90:209-228).  Plane strain
! (mode 2) and antiplane shear (mode 3) with FDmethod SBP2, SBP4 or SBP6 are
! supported; an option the device path does not cover stops with a message
! naming it.

module fdmap_setup

  use iso_c_binding
  use fdmap_capi
  use fdmap_grid
  implicit none

  type :: fields_perturb
     character(256) :: shape
     real(dp) :: x0,y0,Lx,Ly,vx,vy,vz,sxx,sxy,sxz,syy,syz,szz
  end type fields_perturb

  type :: ratestate_constant
     real(dp) :: a,b,V0,f0,L,fw,Vw
  end type ratestate_constant

  type :: slipweak_constant
     real(dp) :: c0,fs,fd,Dc
  end type slipweak_constant

  type :: kinematic
     real(dp) :: c0,fd,dfdx,vL,vR,xL,xR,tmax
  end type kinematic

  type :: load
     character(256) :: shape
     real(dp) :: S,N,T,R,x0,y0
  end type load

  type :: side_data
     integer(c_int32_t) :: bc = FDM_BC_NONE
     real(dp),allocatable :: normal(:,:),M(:,:),F0(:,:),x(:),y(:)
  end type side_data

  type :: host_block
     type(block_grid) :: G
     character(256) :: response
     real(dp) :: rho,G_,cs,cp,M_,lambda,nu,K,gamma,Zs,Zp,mu,beta,b,h,eta
     logical :: het = .false.
     real(dp) :: Cdiss = 0.0_dp
     ! perfectly matched layer: pmlx, pmly and n, sigma, kappa, mx, px, my, py
     logical :: pmlx = .false., pmly = .false.
     real(dp) :: pml(7) = 0.0_dp
     real(dp) :: F0blk(9)
     character(256) :: problem
     type(side_data) :: side(4)
  end type host_block

  type :: host_iface
     character(24) :: coupling
     character(1) :: direction
     integer :: iblockm,iblockp,mg,pg,n
     character(256) :: friction_law
     real(dp) :: skempton,flock,rup_threshold,ld_T
     character(256) :: rup_field
     logical :: reg
     real(dp),allocatable :: nhat(:,:),x(:),y(:),rs(:,:),sw(:,:),S0b(:),N0b(:),ldS(:),ldN(:),Psi0(:)
     integer(c_int32_t),allocatable :: outside(:)
     ! kinematic forcing: c0, fd, dfdx, vL, vR, xL, xR, tmax, and the distance
     ! along the fault of every node
     logical :: force = .false.
     real(dp) :: kn(8) = 0.0_dp
     real(dp),allocatable :: knx(:)
  end type host_iface

  ! problem_list (main.f90:29)
  character(256) :: pb_name,RKmethod
  real(dp) :: t_init,CFL,dt,refine,tend,tnuc_min,tnuc_max,minV
  integer :: nt,ninfo,RKorder
  logical :: constant_dt,solid

  ! domain; pb_mode 2: plane strain (vx vy sxx sxy syy szz), 3: antiplane
  ! shear (vz sxz syz, stored in the first three field slots)
  integer :: nblocks_x,nblocks_y,nblocks,nifaces,gnx,gny
  integer :: pb_mode = 2
  type(host_block),allocatable,target :: HB(:)
  type(host_iface),allocatable,target :: HI(:)
  real(dp),allocatable,target :: gcx(:,:),gcy(:,:),gxq(:,:),gxr(:,:),gyq(:,:),gyr(:,:),gJ(:,:),gJi(:,:)
  real(dp),allocatable,target :: gF(:,:,:)
  ! heterogeneous medium (material.f90:226-321): rho, cs, cp at every point
  real(dp),allocatable,target :: gmrho(:,:),gmcs(:,:),gmcp(:,:)
  real(dp) :: hmin,hmax

  ! C description
  type(fdm_problem),target :: cprob
  type(fdm_block),allocatable,target :: cblk(:)
  type(fdm_iface),allocatable,target :: cif(:)

  character(512) :: setup_error = ''

contains

  subroutine fail(msg)
    character(*),intent(in) :: msg
    if (setup_error=='') setup_error = msg
  end subroutine fail


  subroutine seek_to_string(unit,str,found)
    ! io.f90:209-228 (reports instead of aborting)
    integer,intent(in) :: unit
    character(*),intent(in) :: str
    logical,intent(out) :: found
    character(256) :: temp
    rewind(unit)
    found = .true.
    do
       read(unit,*,end=100) temp
       if (temp==str) return
    end do
100 found = .false.
  end subroutine seek_to_string


  logical function mode3_bc(bc)
    character(*),intent(in) :: bc
    select case(trim(bc))
    case('none','absorbing','free','rigid')
       mode3_bc = .true.
    case default
       mode3_bc = .false.
    end select
  end function mode3_bc


  function bc_code(bc) result(code)
    character(*),intent(in) :: bc
    integer(c_int32_t) :: code
    select case(trim(bc))
    case('none');               code = FDM_BC_NONE
    case('absorbing');          code = FDM_BC_ABSORBING
    case('free');               code = FDM_BC_FREE
    case('rigid');              code = FDM_BC_RIGID
    case('absorbing-velocity'); code = FDM_BC_ABSORBING_VELOCITY
    case('free-velocity');      code = FDM_BC_FREE_VELOCITY
    case('absorbing-0');        code = FDM_BC_ABSORBING_0
    case('free-0');             code = FDM_BC_FREE_0
    case('rigid-0');            code = FDM_BC_RIGID_0
    case('mms-sin-w','inplane-fault-mms-w'); code = FDM_BC_MMS_SIN_W
    case default
       code = -1
       call fail("unsupported boundary condition '"//trim(bc)//"'")
    end select
  end function bc_code


  function law_code(law) result(code)
    character(*),intent(in) :: law
    integer(c_int32_t) :: code
    select case(trim(law))
    case('frictionless'); code = FDM_LAW_FRICTIONLESS
    case('locked');       code = FDM_LAW_LOCKED
    case('SW');           code = FDM_LAW_SW
    case('SL');           code = FDM_LAW_SL
    case('FL');           code = FDM_LAW_FL
    case('RSL');          code = FDM_LAW_RSL
    case('RFL');          code = FDM_LAW_RFL
    case('RSF');          code = FDM_LAW_RSF
    case('RAL');          code = FDM_LAW_RAL
    case default
       code = -1
       call fail("unsupported friction law '"//trim(law)//"'")
    end select
  end function law_code


  ! shape functions (utilities.f90:303-420)
  function boxcar(r,w,fin,fout) result(f)
    real(dp),intent(in) :: r,w,fin,fout
    real(dp) :: f
    real(dp),parameter :: tol = 100.0_dp*epsilon(1.0_dp)
    if (-w+tol<r.and.r<w-tol) then
       f = fin
    elseif (abs(-w-r)<=tol.or.abs(r-w)<=tol) then
       f = 0.5_dp*(fin+fout)
    else
       f = fout
    end if
  end function boxcar

  function gaussian(r,w,fin,fout) result(f)
    real(dp),intent(in) :: r,w,fin,fout
    real(dp) :: f
    f = fout+(fin-fout)*exp(-0.5_dp*(r/w)**2)
  end function gaussian

  function smooth(r,w,fin,fout) result(f)
    real(dp),intent(in) :: r,w,fin,fout
    real(dp) :: f
    real(dp),parameter :: tol = 100.0_dp*epsilon(1.0_dp)
    if (r>=w-tol) then
       f = fout
    else
       f = fout+(fin-fout)*exp(r**2/(r**2-w**2))
    end if
  end function smooth

  function triangle(r,w,fin,fout) result(f)
    real(dp),intent(in) :: r,w,fin,fout
    real(dp) :: f
    if (-w<r.and.r<w) then
       f = fout+(fin-fout)*(1.0_dp-r/w)
    else
       f = fout
    end if
  end function triangle

  function smooth_boxcar(r,w,l,fin,fout) result(f)
    real(dp),intent(in) :: r,w,l,fin,fout
    real(dp) :: f
    if (r<=w) then
       f = 1.0_dp
    elseif (r<w+l) then
       f = 0.5_dp*(1.0_dp+tanh(l/(r-w-l)+l/(r-w)))
    else
       f = 0.0_dp
    end if
    f = fout+(fin-fout)*f
  end function smooth_boxcar

  function decaying_step(x,x0,w,fin,fout) result(f)
    real(dp),intent(in) :: x,x0,w,fin,fout
    real(dp) :: f
    if (x==x0) then
       f = 0.5_dp
    elseif (x<x0) then
       f = exp((x-x0)/w)
    else
       f = 0.0_dp
    end if
    f = fout+(fin-fout)*f
  end function decaying_step


  function perturb_fields(x,y,P) result(A)
    ! fields.f90:900-934
    real(dp),intent(in) :: x,y
    type(fields_perturb),intent(in) :: P
    real(dp) :: A,r
    r = sqrt(((x-P%x0)/P%Lx)**2+((y-P%y0)/P%Ly)**2)
    select case(trim(P%shape))
    case('')
       A = 0.0_dp
    case('linear')
       A = (x-P%x0)/P%Lx+(y-P%y0)/P%Ly
    case('boxcar')
       A = boxcar(r,1.0_dp,1.0_dp,0.0_dp)
    case('triangle')
       A = triangle(r,1.0_dp,1.0_dp,0.0_dp)
    case('smooth')
       A = smooth(r,1.0_dp,1.0_dp,0.0_dp)
    case('gaussian')
       A = gaussian(r,1.0_dp,1.0_dp,0.0_dp)
    case('dy_gaussian')
       A = -(y-P%y0)/(P%Ly**2)*exp(-0.5_dp*(y-P%y0)**2/(P%Ly**2))
    case default
       A = 0.0_dp
       call fail("invalid perturbation shape '"//trim(P%shape)//"'")
    end select
  end function perturb_fields


  ! ------------------------------------------------------------------
  ! mms_sin (mms.f90:52-234): exact fields, used for initial conditions
  ! of 'mms-sin' problems
  ! ------------------------------------------------------------------
  function mms_sin(x,y,t,side,field) result(F)
    real(dp),intent(in) :: x,y,t
    integer,intent(in) :: side
    character(*),intent(in) :: field
    real(dp) :: F
    real(dp) :: G,cs,cp,rho,Zs,pi,kx,ky,mv,w,ft,mf,kf,ftpp,nx,ny,mx,my
    real(dp) :: vx,vy,syy,sxy,sxx
    G = 32.0_dp; cs = 3.0_dp; cp = 5.0_dp
    rho = G/cs**2
    Zs = rho*cs
    pi = 4.0_dp*atan(1.0_dp)
    kx = 2.0_dp*pi; ky = 2.0_dp*pi; mv = 2.0_dp
    w = (4.0_dp/0.46_dp)*pi
    ft = cos(w*t)
    mf = 0.1_dp; kf = 1.0_dp*pi
    ftpp = mf*kf*cos(kf*x)
    nx = ftpp/sqrt(1.0_dp+ftpp**2)
    ny = -1.0_dp/sqrt(1.0_dp+ftpp**2)
    mx = ny
    my = -nx
    vx = mv*(3.0_dp-2.0_dp*side)*(2.0_dp*12.6_dp/Zs)*cos(kx*x)*cos(ky*y)*ft
    vy = (2.0_dp*12.6_dp/Zs)*cos(kx*x)*cos(ky*y)*ft
    syy = -63.0_dp*cos(kx*x)*cos(ky*y)*ft-126.0_dp
    sxy = 0.6_dp*126.0_dp*cos(kx*x)*cos(ky*y)*ft
    sxx = (2.0_dp*side-3.0_dp)*0.6_dp*1260.0_dp*cos(kx*x)*cos(ky*y)*ft
    select case(field)
    case('vx');  F = ny*vx-my*vy
    case('vy');  F = mx*vy-nx*vx
    case('sxx'); F = syy*my**2+sxx*ny**2-2.0_dp*sxy*my*ny
    case('sxy'); F = -syy*my*mx-sxx*ny*nx+sxy*(ny*mx+nx*my)
    case('syy'); F = syy*mx**2+sxx*nx**2-2.0_dp*sxy*mx*nx
    case default; F = 0.0_dp
    end select
  end function mms_sin


  subroutine initial_fields(F,x,y,t,problem,iblock)
    ! fields.f90:937-1004 (mode 2 problems)
    real(dp),intent(inout) :: F(6)
    real(dp),intent(in) :: x,y,t
    character(*),intent(in) :: problem
    integer,intent(in) :: iblock
    select case(trim(problem))
    case('mms-sin','mms-sin-nostate','inplane-fault-mms','inplane-fault-mms-nostate')
       F(1) = F(1)+mms_sin(x,y,t,iblock,'vx')
       F(2) = F(2)+mms_sin(x,y,t,iblock,'vy')
       F(3) = F(3)+mms_sin(x,y,t,iblock,'sxx')
       F(4) = F(4)+mms_sin(x,y,t,iblock,'sxy')
       F(5) = F(5)+mms_sin(x,y,t,iblock,'syy')
       F(6) = F(6)+mms_sin(x,y,t,iblock,'szz')
    case('','uniform')
    case default
       call fail("unsupported fields problem '"//trim(problem)//"'")
    end select
  end subroutine initial_fields


  function search_floor(v,y,out) result(mid)
    ! search_binary(...,'floor',...) over an ordered list (utilities.f90:13-157)
    real(dp),intent(in) :: v(:),y
    logical,intent(out) :: out
    integer :: mid,n,lo,hi,m
    n = size(v)
    if (n==1) then
       mid = 1
       out = .not.(y==v(1))
       return
    end if
    if (y<v(1)) then
       mid = 1
       out = .true.
       return
    elseif (y>=v(n)) then
       mid = n-1
       out = .true.
       return
    end if
    lo = 1
    hi = n
    do while (hi-lo>1)
       m = (lo+hi)/2
       if (v(m)<=y) then
          lo = m
       else
          hi = m
       end if
    end do
    mid = lo
    out = .false.
  end function search_floor


  subroutine singular_source_weights(a,h,d)
    ! discrete delta at i0-2..i0+2 (source.f90:327-356)
    real(dp),intent(in) :: a,h
    real(dp),intent(out) :: d(5)
    real(dp) :: beta,gam,gap
    beta = (1.0_dp-2.0_dp*a)/4.0_dp+0.25_dp
    gam = (2.0_dp*a-1.0_dp)/16.0_dp
    gap = (2.0_dp*a+1.0_dp)/16.0_dp
    d(1) = (-gam)/h
    d(2) = (beta+3.0_dp*gam-gap)/h
    d(3) = (1.0_dp-a-2.0_dp*beta-3.0_dp*gam+3.0_dp*gap)/h
    d(4) = (a+beta+gam-3.0_dp*gap)/h
    d(5) = (gap)/h
  end subroutine singular_source_weights


  subroutine rk_coefficients(order,A,B,C,nstage)
    ! low-storage Runge-Kutta tables (time_step.f90:25-66)
    integer,intent(in) :: order
    real(dp),intent(out) :: A(8),B(8),C(9)
    integer,intent(out) :: nstage
    A = 0.0_dp; B = 0.0_dp; C = 0.0_dp
    select case(order)
    case(1)
       nstage = 1
       A(1) = 0.0_dp; B(1) = 1.0_dp; C(1:2) = (/ 0.0_dp,1.0_dp /)
    case(2)
       nstage = 2
       A(1:2) = (/ 0.0_dp,-1.0_dp /)
       B(1:2) = (/ 1.0_dp,1.0_dp/2.0_dp /)
       C(1:3) = (/ 0.0_dp,1.0_dp,1.0_dp /)
    case(3)
       nstage = 3
       A(1:3) = (/ 0.0_dp,-5.0_dp/9.0_dp,-153.0_dp/128.0_dp /)
       B(1:3) = (/ 1.0_dp/3.0_dp,15.0_dp/16.0_dp,8.0_dp/15.0_dp /)
       C(1:4) = (/ 0.0_dp,1.0_dp/3.0_dp,3.0_dp/4.0_dp,1.0_dp /)
    case(4)
       nstage = 5
       A(1:5) = (/ 0.0_dp,-567301805773.0_dp/1357537059087.0_dp, &
            -2404267990393.0_dp/2016746695238.0_dp,-3550918686646.0_dp/2091501179385.0_dp, &
            -1275806237668.0_dp/842570457699.0_dp /)
       B(1:5) = (/ 1432997174477.0_dp/9575080441755.0_dp,5161836677717.0_dp/13612068292357.0_dp, &
            1720146321549.0_dp/2090206949498.0_dp,3134564353537.0_dp/4481467310338.0_dp, &
            2277821191437.0_dp/14882151754819.0_dp /)
       C(1:6) = (/ 0.0_dp,1432997174477.0_dp/9575080441755.0_dp, &
            2526269341429.0_dp/6820363962896.0_dp,2006345519317.0_dp/3224310063776.0_dp, &
            2802321613138.0_dp/2924317926251.0_dp,1.0_dp /)
    case default
       nstage = 0
       call fail('invalid Runge-Kutta order')
    end select
  end subroutine rk_coefficients




  ! ------------------------------------------------------------------
  ! read the whole problem and build the host description
  ! ------------------------------------------------------------------
  subroutine setup_problem(infile,precision,ok)
    character(*),intent(in) :: infile
    integer(c_int32_t),intent(in) :: precision
    logical,intent(out) :: ok
    integer :: input,stat
    setup_error = ''
    ok = .false.
    call release_problem()
    open(newunit=input,file=infile,status='old',action='read',iostat=stat)
    if (stat/=0) then
       call fail('Error opening '//trim(infile))
       return
    end if
    call read_problem_list(input)
    if (setup_error=='') call read_domain(input)
    if (setup_error=='') call read_pml(input)
    if (setup_error=='') call build_grid()
    if (setup_error=='') call read_blocks(input)
    if (setup_error=='') call read_interfaces(input)
    if (setup_error=='') call set_dt()
    if (setup_error=='') call read_source(input)
    if (setup_error=='') call fill_cproblem(precision)
    close(input)
    ok = (setup_error=='')
  end subroutine setup_problem


  subroutine release_problem()
    if (allocated(HB)) deallocate(HB)
    if (allocated(HI)) deallocate(HI)
    if (allocated(gcx)) deallocate(gcx,gcy,gxq,gxr,gyq,gyr,gJ,gJi)
    if (allocated(gF)) deallocate(gF)
    if (allocated(gmrho)) deallocate(gmrho,gmcs,gmcp)
    if (allocated(cblk)) deallocate(cblk)
    if (allocated(cif)) deallocate(cif)
  end subroutine release_problem


  subroutine read_problem_list(input)
    ! main.f90:29,57-79
    integer,intent(in) :: input
    integer :: stat
    character(256) :: name
    real(dp) :: t
    namelist /problem_list/ name,t,nt,CFL,dt,refine,ninfo,RKmethod,RKorder, &
         tnuc_min,tnuc_max,minV,constant_dt,solid,tend
    name = 'default'
    t = 0.0_dp
    nt = 0
    CFL = 0.0_dp
    dt = 0.0_dp
    refine = 1.0_dp
    tend = 0.0_dp
    ninfo = 1
    constant_dt = .true.
    RKmethod = 'LS'
    RKorder = 3
    solid = .true.
    tnuc_min = huge(tnuc_min)
    tnuc_max = huge(tnuc_max)
    minV = 0.0_dp
    rewind(input)
    read(input,nml=problem_list,iostat=stat)
    if (stat>0) call fail('Error in problem_list')
    pb_name = name
    t_init = t
    if (trim(RKmethod)/='LS') call fail('Invalid Runge-Kutta method')
    if (.not.solid) call fail('solid = F is not supported')
  end subroutine read_problem_list


  subroutine read_domain(input)
    ! domain.f90:84-232, grid.f90:34-148, domain.f90:1538-1644
    integer,intent(in) :: input
    integer :: stat,mode,nx,ny,nprocs_x,nprocs_y,i,ix,iy
    integer :: nx_list(1000),ny_list(1000)
    integer,allocatable :: mgxa(:),mgya(:)
    character(10) :: FDmethod
    character(6) :: mpi_method
    logical :: operator_split,decomposition_info,energy_balance,displacement,exact_metric,peak, &
         acoustic_gravity_waves,basal_traction_plane_stress
    real(dp) :: Cdiss
    character(256) :: str
    logical :: found,ok
    integer :: iblock_x,iblock_y,mgx,mgy
    character(256) :: bndLfile,bndRfile,bndBfile,bndTfile
    real(dp) :: bndLmap,bndRmap,bndBmap,bndTmap
    type(point) :: LB,LT,RB,RT
    namelist /domain_list/ mode,FDmethod,nblocks_x,nblocks_y,nblocks,nifaces, &
         nx,ny,mpi_method,nprocs_x,nprocs_y,decomposition_info,operator_split, &
         energy_balance,displacement,exact_metric,peak,acoustic_gravity_waves, &
         basal_traction_plane_stress,nx_list,ny_list
    namelist /operator_list/ Cdiss
    namelist /grid_list/ iblock_x,iblock_y,nx,ny,mgx,mgy,LB,LT,RB,RT, &
         bndLfile,bndRfile,bndBfile,bndTfile,bndLmap,bndRmap,bndBmap,bndTmap

    mode = 2
    FDmethod = 'SBP6'
    nblocks_x = 0
    nblocks_y = 0
    nblocks = 0
    nifaces = 0
    exact_metric = .false.
    nx_list = 0
    ny_list = 0
    nx = 1
    ny = 1
    operator_split = .false.
    energy_balance = .false.
    displacement = .false.
    peak = .false.
    acoustic_gravity_waves = .false.
    basal_traction_plane_stress = .false.
    mpi_method = '2d'
    nprocs_x = 1
    nprocs_y = 1
    decomposition_info = .false.
    rewind(input)
    read(input,nml=domain_list,iostat=stat)
    if (stat>0) then
       call fail('Error in domain_list')
       return
    end if
    if (mode/=2.and.mode/=3) call fail('mode must be 2 (plane strain) or 3 (antiplane shear)')
    pb_mode = mode
    call init_op(trim(FDmethod),ok)
    if (.not.ok) call fail("FDmethod must be 'SBP2', 'SBP4' or 'SBP6' (not '"//trim(FDmethod)//"')")
    if (exact_metric) call fail('exact_metric is not supported')
    if (energy_balance.or.displacement.or.peak.or.acoustic_gravity_waves.or. &
         basal_traction_plane_stress.or.operator_split) &
         call fail('energy_balance, displacement, peak, acoustic_gravity_waves, '// &
         'basal_traction_plane_stress and operator_split are not supported')
    if (setup_error/='') return
    if (nblocks_x==0) nblocks_x = 1
    if (nblocks_y==0) nblocks_y = 1
    if (nblocks==0) nblocks = max(1,nblocks_x*nblocks_y)
    if (nblocks/=nblocks_x*nblocks_y) then
       call fail('nblocks must equal nblocks_x*nblocks_y')
       return
    end if
    nifaces = (nblocks_x-1)*nblocks_y+(nblocks_y-1)*nblocks_x
    allocate(HB(nblocks))

    do i = 1,nblocks
       Cdiss = 0.0_dp
       write(str,'(a,i0,a)') '!---BLOCK',i,'---'
       call seek_to_string(input,str,found)
       if (.not.found) then
          call fail("String '"//trim(str)//"' not found in file")
          return
       end if
       read(input,nml=operator_list,iostat=stat)
       if (stat>0) call fail('Error in operator_list')
       HB(i)%Cdiss = Cdiss
       ! the dissipation operators are built for SBP4 only
       if (Cdiss>0.0_dp.and.sbp_order/=4) &
            call fail("artificial dissipation (Cdiss > 0) needs FDmethod = 'SBP4'")
    end do

    do i = 1,nblocks
       iblock_x = 0; iblock_y = 0; nx = 1; ny = 1; mgx = 1; mgy = 1
       LB = point(huge(1.0_dp),huge(1.0_dp)); LT = LB; RB = LB; RT = LB
       bndLfile = ''; bndRfile = ''; bndBfile = ''; bndTfile = ''
       bndLmap = 0.0_dp; bndRmap = 0.0_dp; bndBmap = 0.0_dp; bndTmap = 0.0_dp
       write(str,'(a,i0,a)') '!---BLOCK',i,'---'
       call seek_to_string(input,str,found)
       read(input,nml=grid_list,iostat=stat)
       if (stat>0) call fail('Error in grid_list')
       HB(i)%G%nx = nx
       HB(i)%G%ny = ny
       HB(i)%G%LB = LB; HB(i)%G%LT = LT; HB(i)%G%RB = RB; HB(i)%G%RT = RT
       HB(i)%G%bndLfile = bndLfile; HB(i)%G%bndRfile = bndRfile
       HB(i)%G%bndBfile = bndBfile; HB(i)%G%bndTfile = bndTfile
       HB(i)%G%bndLmap = bndLmap; HB(i)%G%bndRmap = bndRmap
       HB(i)%G%bndBmap = bndBmap; HB(i)%G%bndTmap = bndTmap
    end do

    if (nx_list(1)/=0.and.ny_list(1)/=0) then
       do ix = 1,nblocks_x
          do iy = 1,nblocks_y
             HB(ix+(iy-1)*nblocks_x)%G%nx = nx_list(ix)
             HB(ix+(iy-1)*nblocks_x)%G%ny = ny_list(iy)
          end do
       end do
    end if

    ! set_grid_indices (domain.f90:1538-1579)
    allocate(mgxa(nblocks_x),mgya(nblocks_y))
    mgxa(1) = 1
    mgya(1) = 1
    do ix = 2,nblocks_x
       mgxa(ix) = mgxa(ix-1)+HB(ix-1)%G%nx
    end do
    do iy = 2,nblocks_y
       mgya(iy) = mgya(iy-1)+HB(1+(iy-2)*nblocks_x)%G%ny
    end do
    do ix = 1,nblocks_x
       do iy = 1,nblocks_y
          i = ix+(iy-1)*nblocks_x
          HB(i)%G%mgx = mgxa(ix)
          HB(i)%G%mgy = mgya(iy)
          HB(i)%G%iblock_x = ix
          HB(i)%G%iblock_y = iy
       end do
    end do

    ! set_refine (grid.f90:110-148)
    do i = 1,nblocks
       associate(G => HB(i)%G)
         if (G%nx/=1) G%nx = ceiling(dble(G%nx-1)*refine)+1
         if (G%ny/=1) G%ny = ceiling(dble(G%ny-1)*refine)+1
         G%mgx = ceiling(dble(G%mgx-G%iblock_x)*refine)+G%iblock_x
         G%mgy = ceiling(dble(G%mgy-G%iblock_y)*refine)+G%iblock_y
         G%pgx = G%mgx+G%nx-1
         G%pgy = G%mgy+G%ny-1
         if (G%nx<max(8,2*nbst,nbnd).or.G%ny<max(8,2*nbst,nbnd)) then
            write(str,'(a,i0,a,a)') 'every block needs at least ',max(8,2*nbst,nbnd), &
                 ' points in x and y for ',trim(FDmethod)
            call fail(trim(str))
         end if
       end associate
    end do

    ! set_grid (domain.f90:1583-1607)
    gnx = 0
    gny = 0
    do ix = 1,nblocks_x
       gnx = gnx+HB(ix)%G%nx
    end do
    do iy = 1,nblocks_y
       gny = gny+HB(1+(iy-1)*nblocks_x)%G%ny
    end do
  end subroutine read_domain


  subroutine read_pml(input)
    ! init_pml (material.f90:152-223): pml_list after each block's marker.
    ! The reference leaves n unset when the list does not give it; here it
    ! defaults to 0, the constant-damping layer.
    integer,intent(in) :: input
    integer :: i,stat
    character(256) :: str
    logical :: found
    logical :: pmlx,pmly
    real(dp) :: n,sigma,kappa,mx,px,my,py
    namelist /pml_list/ pmlx,pmly,n,sigma,kappa,mx,px,my,py
    do i = 1,nblocks
       pmlx = .false.
       pmly = .false.
       n = 0.0_dp
       sigma = 0.0_dp
       kappa = 0.0_dp
       mx = -huge(0.0_dp)
       px = huge(0.0_dp)
       my = -huge(0.0_dp)
       py = huge(0.0_dp)
       write(str,'(a,i0,a)') '!---BLOCK',i,'---'
       call seek_to_string(input,str,found)
       read(input,nml=pml_list,iostat=stat)
       if (stat>0) then
          call fail('Error in pml_list')
          return
       end if
       HB(i)%pmlx = pmlx
       HB(i)%pmly = pmly
       HB(i)%pml = (/ n,sigma,kappa,mx,px,my,py /)
       if (pb_mode==3.and.(pmlx.or.pmly)) then
          ! set_rates has no layer in mode 3 (time_step.f90:522-547)
          call fail('pml_list: the perfectly matched layer is available in mode 2 only')
          return
       end if
       if ((pmlx.or.pmly).and.sbp_order/=4) then
          call fail("pml_list: the perfectly matched layer needs FDmethod = 'SBP4'")
          return
       end if
    end do
  end subroutine read_pml


  subroutine build_grid()
    ! init_grid + init_grid_partials for every block (grid.f90:188-618)
    integer :: i
    logical :: ok
    character(256) :: msg
    allocate(gcx(gnx,gny),gcy(gnx,gny),gxq(gnx,gny),gxr(gnx,gny),gyq(gnx,gny),gyr(gnx,gny), &
         gJ(gnx,gny),gJi(gnx,gny))
    gcx = 1d40
    gcy = 1d40
    gxq = 1.0_dp
    gxr = 0.0_dp
    gyq = 0.0_dp
    gyr = 1.0_dp
    gJ = 1.0_dp
    gJi = 1.0_dp
    hmin = 1d40
    hmax = 0.0_dp
    do i = 1,nblocks
       if (HB(i)%G%mgx<1.or.gnx<HB(i)%G%pgx.or.HB(i)%G%mgy<1.or.gny<HB(i)%G%pgy) then
          call fail('block indices too small or large')
          return
       end if
       call build_block_coords(HB(i)%G,gcx,gcy,ok,msg)
       if (.not.ok) then
          call fail(msg)
          return
       end if
    end do
    do i = 1,nblocks
       call block_metrics(HB(i)%G,gcx,gcy,gxq,gxr,gyq,gyr,gJ,gJi,hmin,hmax)
    end do
  end subroutine build_grid


  subroutine read_blocks(input)
    ! init_material, init_boundaries, init_fields, init_edges per block
    ! (material.f90:29-149, boundaries.f90:15-71, fields.f90:57-278,
    ! domain.f90:1349-1459)
    integer,intent(in) :: input
    integer :: ib,stat,s,k,i,j,n,l
    character(256) :: str
    logical :: found
    ! material_list
    character(256) :: response,filename
    real(dp) :: rho,G,cs,cp,beta,mu,b,h,eta
    logical :: heterogeneous,plastic_strain_tensor,plastic_work
    ! boundaries_list
    character(24) :: bcL,bcR,bcB,bcT
    real(dp) :: rhog
    character(256) :: problem
    ! fields_list
    real(dp) :: Psi,vx0,vy0,vz0,sxx0,sxy0,sxz0,syy0,syz0,szz0
    type(fields_perturb) :: P1,P2,P3,P4,P5
    logical :: prestress_from_file,initial_condition_from_file
    character(256) :: prestress_filename,initial_condition_filename
    real(dp) :: F0(6),FP(6,5),A(5),xx,yy,Fv(6),Ks,Kp,nrm,erho,ecs,ecp
    namelist /material_list/ response,heterogeneous,filename,rho,G,cs,cp,beta,mu,b,h,eta, &
         plastic_strain_tensor,plastic_work
    namelist /boundaries_list/ bcL,bcR,bcB,bcT,rhog,problem
    namelist /fields_list/ problem,Psi,vx0,vy0,vz0,sxx0,sxy0,sxz0,syy0,syz0,szz0, &
         P1,P2,P3,P4,P5,prestress_from_file,prestress_filename, &
         initial_condition_from_file,initial_condition_filename

    allocate(gF(gnx,gny,6))
    gF = 1d40
    do ib = 1,nblocks
       write(str,'(a,i0,a)') '!---BLOCK',ib,'---'
       associate(HBk => HB(ib), GB => HB(ib)%G)
         ! ---- material ----
         response = 'elastic'
         heterogeneous = .false.
         plastic_strain_tensor = .false.
         plastic_work = .false.
         filename = ''
         rho = 0.0_dp; G = 0.0_dp; cs = 0.0_dp; cp = 0.0_dp
         beta = 0.0_dp; mu = 1.0_dp; b = 0.0_dp; h = 0.0_dp; eta = 0.0_dp
         call seek_to_string(input,str,found)
         read(input,nml=material_list,iostat=stat)
         if (stat>0) call fail('Error in material_list')
         if (trim(response)/='elastic'.and.trim(response)/='plastic') &
              call fail('Invalid material response: '//trim(response))
         ! a plastic block with these flags accumulates the plastic strain
         ! tensor Ep and plastic work Wp (plastic.f90:51-54) and writes them
         ! into the field checkpoint (fields.f90:194-200, 381-389); neither
         ! is built, so the run stops rather than dropping them
         if (trim(response)=='plastic'.and.(plastic_strain_tensor.or.plastic_work)) &
              call fail('plastic_strain_tensor and plastic_work are not supported')
         ! the first heterogeneous block reads the medium of the whole grid
         ! (material.f90:80-81)
         if (heterogeneous.and.(.not.allocated(gmrho))) call read_medium(filename)
         if (cs==0.0_dp.and.(.not.heterogeneous)) call fail('Must specify S-wave speed')
         if (setup_error/='') return
         if (cp==0.0_dp) cp = sqrt(3.0_dp)*cs
         if (rho==0.0_dp) rho = G/cs**2
         if (G==0.0_dp) G = rho*cs**2
         HBk%response = response
         HBk%het = heterogeneous
         HBk%rho = rho
         HBk%G_ = G
         HBk%cs = cs
         HBk%cp = cp
         HBk%beta = beta
         HBk%mu = mu
         HBk%b = b
         HBk%h = h
         HBk%eta = eta
         HBk%M_ = rho*cp**2
         HBk%lambda = HBk%M_-2.0_dp*G
         HBk%nu = 0.5_dp*HBk%lambda/(HBk%lambda+G)
         HBk%K = HBk%lambda+2.0_dp*G/3.0_dp
         HBk%gamma = 1.0_dp-2.0_dp*(cs/cp)**2
         HBk%Zs = rho*cs
         HBk%Zp = rho*cp

         ! ---- boundaries ----
         bcL = 'none'; bcR = 'none'; bcB = 'none'; bcT = 'none'
         rhog = 0.0_dp
         problem = ''
         call seek_to_string(input,str,found)
         read(input,nml=boundaries_list,iostat=stat)
         if (stat>0) call fail(trim(str)//' :: Error in boundaries_list')
         ! set_bc_mode3 knows absorbing, free and rigid (boundaries.f90:186-200)
         if (pb_mode==3) then
            if (.not.(mode3_bc(bcL).and.mode3_bc(bcR).and.mode3_bc(bcB).and.mode3_bc(bcT))) &
                 call fail("mode 3 boundary conditions are 'absorbing', 'free' or 'rigid'")
         end if
         HBk%side(1)%bc = bc_code(bcL)
         HBk%side(2)%bc = bc_code(bcR)
         HBk%side(3)%bc = bc_code(bcB)
         HBk%side(4)%bc = bc_code(bcT)
         if (setup_error/='') return

         ! ---- fields ----
         problem = ''
         Psi = 0.0_dp
         vx0 = 0.0_dp; vy0 = 0.0_dp; vz0 = 0.0_dp
         sxx0 = 0.0_dp; sxy0 = 0.0_dp; sxz0 = 0.0_dp; syy0 = 0.0_dp; syz0 = 0.0_dp; szz0 = 0.0_dp
         P1 = fields_perturb('',0.0_dp,0.0_dp,1.0_dp,1.0_dp,0.0_dp,0.0_dp,0.0_dp,0.0_dp, &
              0.0_dp,0.0_dp,0.0_dp,0.0_dp,0.0_dp)
         P2 = P1; P3 = P1; P4 = P1; P5 = P1
         prestress_from_file = .false.
         prestress_filename = ''
         initial_condition_from_file = .false.
         initial_condition_filename = ''
         call seek_to_string(input,str,found)
         read(input,nml=fields_list,iostat=stat)
         if (stat>0) call fail('Error in fields_list')
         if (prestress_from_file.or.initial_condition_from_file) &
              call fail('prestress/initial condition files are not supported')
         if (setup_error/='') return
         HBk%problem = problem
         if (trim(problem)=='uniform') then
            if (Psi==45.0_dp) then
               sxx0 = syy0
            else
               sxx0 = (1.0_dp-2.0_dp*sxy0/(syy0*tan(2.0_dp*(0.017453292519943_dp*Psi))))*syy0
            end if
            szz0 = 0.5_dp*(sxx0+syy0)
         end if
         HBk%F0blk = (/ vx0,vy0,vz0,sxx0,sxy0,sxz0,syy0,syz0,szz0 /)
         if (pb_mode==3) then
            ! fields.f90:167-174, 803-809
            if (index(problem,'mms')>0.or.index(problem,'inplane')>0) &
                 call fail("fields problem '"//trim(problem)//"' is not available in mode 3")
            if (trim(response)=='plastic') &
                 call fail('mode 3 covers elastic blocks (no plastic response)')
            if (setup_error/='') return
            F0 = (/ vz0,sxz0,syz0,0.0_dp,0.0_dp,0.0_dp /)
            FP(:,1) = (/ P1%vz,P1%sxz,P1%syz,0.0_dp,0.0_dp,0.0_dp /)
            FP(:,2) = (/ P2%vz,P2%sxz,P2%syz,0.0_dp,0.0_dp,0.0_dp /)
            FP(:,3) = (/ P3%vz,P3%sxz,P3%syz,0.0_dp,0.0_dp,0.0_dp /)
            FP(:,4) = (/ P4%vz,P4%sxz,P4%syz,0.0_dp,0.0_dp,0.0_dp /)
            FP(:,5) = (/ P5%vz,P5%sxz,P5%syz,0.0_dp,0.0_dp,0.0_dp /)
         else
            F0 = (/ vx0,vy0,sxx0,sxy0,syy0,szz0 /)
            FP(:,1) = (/ P1%vx,P1%vy,P1%sxx,P1%sxy,P1%syy,P1%szz /)
            FP(:,2) = (/ P2%vx,P2%vy,P2%sxx,P2%sxy,P2%syy,P2%szz /)
            FP(:,3) = (/ P3%vx,P3%vy,P3%sxx,P3%sxy,P3%syy,P3%szz /)
            FP(:,4) = (/ P4%vx,P4%vy,P4%sxx,P4%sxy,P4%syy,P4%szz /)
            FP(:,5) = (/ P5%vx,P5%vy,P5%sxx,P5%sxy,P5%syy,P5%szz /)
         end if

         ! sides: curve coordinates, normals, F0, M (init_fields_side, init_edges)
         do s = 1,4
            if (s<=2) then
               n = GB%ny
            else
               n = GB%nx
            end if
            allocate(HBk%side(s)%normal(n,2),HBk%side(s)%M(n,5),HBk%side(s)%F0(n,6), &
                 HBk%side(s)%x(n),HBk%side(s)%y(n))
            do k = 1,n
               select case(s)
               case(1)
                  j = GB%mgy+k-1; i = GB%mgx
                  xx = GB%bndL%x(j); yy = GB%bndL%y(j)
                  HBk%side(s)%normal(k,:) = GB%bndL%n(j,:)
                  nrm = sqrt(gxr(i,j)**2+gyr(i,j)**2)
               case(2)
                  j = GB%mgy+k-1; i = GB%pgx
                  xx = GB%bndR%x(j); yy = GB%bndR%y(j)
                  HBk%side(s)%normal(k,:) = GB%bndR%n(j,:)
                  nrm = sqrt(gxr(i,j)**2+gyr(i,j)**2)
               case(3)
                  i = GB%mgx+k-1; j = GB%mgy
                  xx = GB%bndB%x(i); yy = GB%bndB%y(i)
                  HBk%side(s)%normal(k,:) = GB%bndB%n(i,:)
                  nrm = sqrt(gxq(i,j)**2+gyq(i,j)**2)
               case default
                  i = GB%mgx+k-1; j = GB%pgy
                  xx = GB%bndT%x(i); yy = GB%bndT%y(i)
                  HBk%side(s)%normal(k,:) = GB%bndT%n(i,:)
                  nrm = sqrt(gxq(i,j)**2+gyq(i,j)**2)
               end select
               HBk%side(s)%x(k) = xx
               HBk%side(s)%y(k) = yy
               A(1) = perturb_fields(xx,yy,P1)
               A(2) = perturb_fields(xx,yy,P2)
               A(3) = perturb_fields(xx,yy,P3)
               A(4) = perturb_fields(xx,yy,P4)
               Fv = F0+A(1)*FP(:,1)+A(2)*FP(:,2)+A(3)*FP(:,3)+A(4)*FP(:,4)
               call initial_fields(Fv,xx,yy,t_init,problem,ib)
               HBk%side(s)%F0(k,:) = Fv
               ! init_edges: the medium at the node for a heterogeneous block
               ! (domain.f90:1379-1456)
               if (HBk%het) then
                  erho = gmrho(i,j); ecs = gmcs(i,j); ecp = gmcp(i,j)
               else
                  erho = rho; ecs = cs; ecp = cp
               end if
               Ks = H00i*ecs*nrm/gJ(i,j)
               Kp = H00i*ecp*nrm/gJ(i,j)
               HBk%side(s)%M(k,1) = erho*ecs
               HBk%side(s)%M(k,2) = erho*ecp
               HBk%side(s)%M(k,3) = 1.0_dp-2.0_dp*(ecs/ecp)**2
               HBk%side(s)%M(k,4) = Ks
               HBk%side(s)%M(k,5) = Kp
            end do
         end do

         ! interior (init_fields_interior, fields.f90:774-823)
         do j = GB%mgy,GB%pgy
            do i = GB%mgx,GB%pgx
               A(1) = perturb_fields(gcx(i,j),gcy(i,j),P1)
               A(2) = perturb_fields(gcx(i,j),gcy(i,j),P2)
               A(3) = perturb_fields(gcx(i,j),gcy(i,j),P3)
               A(4) = perturb_fields(gcx(i,j),gcy(i,j),P4)
               A(5) = perturb_fields(gcx(i,j),gcy(i,j),P5)
               Fv = F0+A(1)*FP(:,1)+A(2)*FP(:,2)+A(3)*FP(:,3)+A(4)*FP(:,4)+A(5)*FP(:,5)
               call initial_fields(Fv,gcx(i,j),gcy(i,j),t_init,problem,ib)
               do l = 1,6
                  gF(i,j,l) = Fv(l)
               end do
            end do
         end do
       end associate
       if (setup_error/='') return
    end do
  end subroutine read_blocks


  subroutine read_interfaces(input)
    ! init_iface, init_iface_blocks, init_friction (interfaces.f90:26-278,
    ! friction.f90:55-317), load_stress shapes (friction.f90:1335-1374)
    integer,intent(in) :: input
    integer :: ii,ix,iy,off,stat,n,k,im,ip,iblk
    integer,allocatable :: Im_(:),Ip_(:)
    character(1),allocatable :: dir(:)
    character(256) :: str
    logical :: found,read_params
    ! interface_list
    character(24) :: coupling
    character(1) :: direction
    integer :: iblockm,iblockp,mg,pg
    ! friction_list
    logical :: force,friction_file,stress_file,strength_file,state_file,reg
    character(256) :: friction_law,problem,rup_field,filename,stress_filename,strength_filename, &
         state_filename
    real(dp) :: Psi0,angle,rup_threshold,uni_x0,uni_dSdx,xlockm,xlockp,flock,skempton,Sf0
    type(ratestate_constant) :: rs
    type(slipweak_constant) :: sw
    type(kinematic) :: kn
    type(load) :: ld
    real(dp) :: xx,r,Av,ang
    namelist /interface_list/ coupling,iblockm,iblockp,direction,mg,pg
    namelist /friction_list/ force,friction_law,problem,friction_file, &
         filename,stress_file,stress_filename, &
         state_file,state_filename,Psi0, &
         angle,rs,sw,kn,ld,rup_field,rup_threshold,uni_x0, &
         uni_dSdx,xlockm,xlockp,flock,skempton,Sf0, &
         strength_file,strength_filename,reg

    allocate(HI(nifaces))
    if (nifaces==0) return
    allocate(Im_(nifaces),Ip_(nifaces),dir(nifaces))
    ! get_iface_neighbors (domain.f90:1648-1682)
    do ix = 1,nblocks_x-1
       do iy = 1,nblocks_y
          Im_(ix+(iy-1)*(nblocks_x-1)) = ix+nblocks_x*(iy-1)
          Ip_(ix+(iy-1)*(nblocks_x-1)) = ix+1+nblocks_x*(iy-1)
       end do
    end do
    off = (nblocks_x-1)*nblocks_y
    dir(1:off) = 'x'
    do ix = 1,nblocks_x
       do iy = 1,nblocks_y-1
          Im_(off+ix+(iy-1)*nblocks_x) = ix+nblocks_x*(iy-1)
          Ip_(off+ix+(iy-1)*nblocks_x) = ix+nblocks_x*iy
       end do
    end do
    dir(off+1:nifaces) = 'y'

    do ii = 1,nifaces
       coupling = 'locked'
       iblockm = Im_(ii)
       iblockp = Ip_(ii)
       direction = dir(ii)
       mg = 0
       pg = 0
       write(str,'(a,i0,a)') '!---IFACE',ii,'---'
       call seek_to_string(input,str,read_params)
       if (read_params) then
          read(input,nml=interface_list,iostat=stat)
          if (stat>0) call fail('Error in interface_list')
       end if
       associate(I => HI(ii), Gm => HB(iblockm)%G)
         if (mg==0.or.pg==0) then
            if (dir(ii)=='x') then
               mg = Gm%mgy; pg = Gm%pgy
            else
               mg = Gm%mgx; pg = Gm%pgx
            end if
         else
            if (direction=='x') then
               iblk = Gm%iblock_y
            else
               iblk = Gm%iblock_x
            end if
            if (mg==pg) then
               mg = ceiling(dble(mg-iblk)*refine)+iblk
               pg = mg
            else
               mg = ceiling(dble(mg-iblk)*refine)+iblk
               pg = ceiling(dble(pg-iblk)*refine)+iblk
            end if
         end if
         I%coupling = coupling
         I%direction = direction
         I%iblockm = iblockm
         I%iblockp = iblockp
         I%mg = mg
         I%pg = pg
         n = pg-mg+1
         I%n = n
         allocate(I%nhat(n,2),I%x(n),I%y(n))
         if (direction=='x') then
            I%nhat = Gm%bndR%n(mg:pg,:)
            I%x = Gm%bndR%x(mg:pg)
            I%y = Gm%bndR%y(mg:pg)
         else
            I%nhat = Gm%bndT%n(mg:pg,:)
            I%x = Gm%bndT%x(mg:pg)
            I%y = Gm%bndT%y(mg:pg)
         end if
         I%friction_law = 'frictionless'
         I%skempton = 0.0_dp
         I%flock = 1d10
         I%rup_threshold = 0.0_dp
         I%rup_field = 'V'
         I%reg = .false.
         I%ld_T = 0.0_dp
         if (trim(coupling)=='locked') cycle
         if (trim(coupling)/='friction') then
            call fail("unsupported interface coupling '"//trim(coupling)//"'")
            return
         end if

         ! friction_list (friction.f90:88-121)
         force = .false.
         friction_law = 'frictionless'
         problem = ''
         friction_file = .false.
         stress_file = .false.
         strength_file = .false.
         state_file = .false.
         filename = ''
         stress_filename = ''
         strength_filename = ''
         state_filename = ''
         Psi0 = 0.0_dp
         angle = 0.0_dp
         rs = ratestate_constant(0.01_dp,0.014_dp,1d-6,0.6_dp,0.4_dp,0.2_dp,1d20)
         sw = slipweak_constant(0.0_dp,1.0_dp,0.0_dp,1.0_dp)
         kn = kinematic(0.0_dp,0.0_dp,1.0_dp,0.0_dp,0.0_dp,0.0_dp,0.0_dp,huge(1.0_dp))
         ld = load('smooth',0.0_dp,0.0_dp,0.0_dp,1.0_dp,0.0_dp,0.0_dp)
         rup_field = 'V'
         rup_threshold = 0.0_dp
         uni_x0 = 0.0_dp
         uni_dSdx = 0.0_dp
         xlockm = -1d10
         xlockp = 1d10
         flock = 1d10
         skempton = 0.0_dp
         Sf0 = 0.0_dp
         reg = .false.
         call seek_to_string(input,str,found)
         if (.not.found) then
            call fail("String '"//trim(str)//"' not found in file")
            return
         end if
         read(input,nml=friction_list,iostat=stat)
         if (stat>0) call fail('Error in friction_list')
         if (trim(problem)/='') call fail("friction problem '"//trim(problem)//"' is not supported")
         if (strength_file) call fail('strength_file is not supported')
         if (trim(rup_field)/='V'.and.trim(rup_field)/='D') &
              call fail('Invalid field to define rupture front')
         if (law_code(friction_law)<0) return
         if (setup_error/='') return
         I%friction_law = friction_law
         I%skempton = skempton
         I%flock = flock
         I%rup_threshold = rup_threshold
         I%rup_field = rup_field
         I%reg = reg
         I%ld_T = ld%T
         I%force = force
         I%kn = (/ kn%c0,kn%fd,kn%dfdx,kn%vL,kn%vR,kn%xL,kn%xR,kn%tmax /)
         allocate(I%rs(n,7),I%sw(n,4),I%S0b(n),I%N0b(n),I%ldS(n),I%ldN(n),I%Psi0(n),I%outside(n), &
              I%knx(n))
         I%rs(:,1) = rs%a; I%rs(:,2) = rs%b; I%rs(:,3) = rs%V0; I%rs(:,4) = rs%f0
         I%rs(:,5) = rs%L; I%rs(:,6) = rs%fw; I%rs(:,7) = rs%Vw
         I%sw(:,1) = sw%c0; I%sw(:,2) = sw%fs; I%sw(:,3) = sw%fd; I%sw(:,4) = sw%Dc
         I%Psi0 = Psi0
         I%S0b = 0.0_dp
         I%N0b = 0.0_dp
         if (friction_file) then
            call read_arrays(filename,n,7,I%rs)
            if (setup_error/='') return
         end if
         if (stress_file) then
            call read_pair(stress_filename,n,I%S0b,I%N0b)
            if (setup_error/='') return
         end if
         if (state_file) then
            call read_arrays(state_filename,n,1,I%Psi0)
            if (setup_error/='') return
         end if
         ang = 0.017453292519943_dp*angle
         do k = 1,n
            xx = I%x(k)*cos(ang)+I%y(k)*sin(ang)
            I%knx(k) = xx
            r = abs((xx-ld%x0)/ld%R)
            select case(trim(ld%shape))
            case('')
               Av = 0.0_dp
            case('uniform')
               Av = 1.0_dp
            case('linear')
               Av = (xx-ld%x0)/ld%R
            case('boxcar')
               Av = boxcar(r,1.0_dp,1.0_dp,0.0_dp)
            case('smooth_boxcar')
               Av = smooth_boxcar(r,1.0_dp,0.1_dp,1.0_dp,0.0_dp)
            case('triangle')
               Av = triangle(r,1.0_dp,1.0_dp,0.0_dp)
            case('smooth')
               Av = smooth(r,1.0_dp,1.0_dp,0.0_dp)
            case('gaussian')
               Av = gaussian(r,1.0_dp,1.0_dp,0.0_dp)
            case('decaying_step')
               Av = decaying_step(I%x(k),ld%x0,ld%R,1.0_dp,0.0_dp)
            case default
               call fail('Invalid load shape')
               return
            end select
            I%ldS(k) = ld%S*Av
            I%ldN(k) = ld%N*Av
            if (xx<xlockm.or.xx>xlockp) then
               I%outside(k) = 1
            else
               I%outside(k) = 0
            end if
         end do
       end associate
    end do
  end subroutine read_interfaces


  subroutine read_arrays(fname,n,ncomp,a)
    ! consecutive FP64 arrays of length n (friction.f90:373-419, 469-489)
    character(*),intent(in) :: fname
    integer,intent(in) :: n,ncomp
    real(dp),intent(inout) :: a(n,ncomp)
    integer :: u,stat
    open(newunit=u,file=fname,access='stream',form='unformatted',status='old', &
         action='read',iostat=stat)
    if (stat/=0) then
       call fail('cannot open '//trim(fname))
       return
    end if
    read(u,iostat=stat) a
    if (stat/=0) call fail('cannot read '//trim(fname))
    close(u)
  end subroutine read_arrays


  subroutine read_medium(fname)
    ! mediumIO('read') (material.f90:278-321): rho, cs, cp of the whole grid,
    ! each an FP64 array [gnx, gny], x fastest
    character(*),intent(in) :: fname
    integer :: u,stat
    allocate(gmrho(gnx,gny),gmcs(gnx,gny),gmcp(gnx,gny))
    open(newunit=u,file=fname,access='stream',form='unformatted',status='old', &
         action='read',iostat=stat)
    if (stat/=0) then
       call fail('cannot open medium file '//trim(fname))
       return
    end if
    read(u,iostat=stat) gmrho
    if (stat==0) read(u,iostat=stat) gmcs
    if (stat==0) read(u,iostat=stat) gmcp
    if (stat/=0) call fail('cannot read medium file '//trim(fname))
    close(u)
  end subroutine read_medium


  subroutine read_pair(fname,n,a,b)
    ! S0 then N0 (friction.f90:422-443)
    character(*),intent(in) :: fname
    integer,intent(in) :: n
    real(dp),intent(inout) :: a(n),b(n)
    integer :: u,stat
    open(newunit=u,file=fname,access='stream',form='unformatted',status='old', &
         action='read',iostat=stat)
    if (stat/=0) then
       call fail('cannot open '//trim(fname))
       return
    end if
    read(u,iostat=stat) a
    if (stat==0) read(u,iostat=stat) b
    if (stat/=0) call fail('cannot read '//trim(fname))
    close(u)
  end subroutine read_pair


  subroutine set_dt()
    ! domain.f90:1462-1535 and the nt rule of main.f90:112-116
    integer :: i
    real(dp) :: cs,cp,c
    if (hmin<=0.0_dp) then
       call fail('Minimum grid spacing must be positive')
       return
    end if
    if (allocated(gmrho)) then ! heterogeneous properties
       cs = maxval(gmcs)
       cp = maxval(gmcp)
    else
       cs = 0.0_dp
       cp = 0.0_dp
       do i = 1,nblocks
          cs = max(cs,HB(i)%cs)
          cp = max(cp,HB(i)%cp)
       end do
    end if
    c = sqrt(cs**2+cp**2)
    dt = dt/refine
    if (CFL==0.0_dp) then
       if (dt==0.0_dp) then
          CFL = 0.5_dp
          dt = CFL*hmin/c
       else
          CFL = c*dt/hmin
       end if
    else
       if (dt==0.0_dp) then
          dt = CFL*hmin/c
       else
          CFL = c*dt/hmin
       end if
    end if
    if (nt==0) then
       nt = ceiling(tend/dt)
    else
       nt = ceiling(dble(nt-1)*refine)+1
    end if
  end subroutine set_dt


  subroutine read_source(input)
    ! init_source and the point-source placement of set_source
    ! (source.f90:21-254)
    integer,intent(in) :: input
    integer :: stat,ib,i0,j0,np
    logical :: forcing_from_file,outx,outy
    real(dp) :: gx,gy,gz,x0,y0,t_duration,t_center,Fx,Fy,Fz,Mxx,Mxy,Myy,Mzx,Mzy
    real(dp) :: gauss_amp,gauss_sigma,gauss_sigmat,hx,hy,ax,ay
    character(256) :: source_time_function,problem,forcing_filename
    namelist /source_list/ problem,gauss_amp,gauss_sigma,gauss_sigmat, &
         gx,gy,gz,x0,y0,t_duration,t_center,Fx,Fy,Fz,Mxx,Mxy,Myy,Mzx,Mzy,source_time_function, &
         forcing_from_file,forcing_filename
    gx = 0.0_dp; gy = 0.0_dp; gz = 0.0_dp
    x0 = 0.0_dp; y0 = 0.0_dp
    t_duration = 1.0_dp; t_center = 0.0_dp
    Fx = 0.0_dp; Fy = 0.0_dp; Fz = 0.0_dp
    Mxx = 0.0_dp; Mxy = 0.0_dp; Myy = 0.0_dp; Mzx = 0.0_dp; Mzy = 0.0_dp
    source_time_function = 'gaussian'
    problem = ''
    gauss_amp = 0.0_dp; gauss_sigma = 0.0_dp; gauss_sigmat = 0.0_dp
    forcing_from_file = .false.
    forcing_filename = ''
    rewind(input)
    read(input,nml=source_list,iostat=stat)
    if (stat>0) then
       call fail('Error in source_list')
       return
    end if
    if (forcing_from_file) call fail('forcing_from_file is not supported')
    if (gx/=0.0_dp.or.gy/=0.0_dp) call fail('gravity body force (gx, gy) is not supported')
    cprob%source%nplace = 0
    cprob%source%t_center = t_center
    cprob%source%t_duration = t_duration
    cprob%source%Fx = Fx
    cprob%source%Fy = Fy
    cprob%source%Mxx = Mxx
    cprob%source%Mxy = Mxy
    cprob%source%Myy = Myy
    select case(trim(source_time_function))
    case('ramp');     cprob%source%stf = FDM_STF_RAMP
    case('sinusoid'); cprob%source%stf = FDM_STF_SINUSOID
    case('gaussian'); cprob%source%stf = FDM_STF_GAUSSIAN
    case('Ricker');   cprob%source%stf = FDM_STF_RICKER
    case default
       call fail('Invalid source time function')
       return
    end select
    if (.not.((Fx/=0.0_dp.or.Fy/=0.0_dp.or.Fz/=0.0_dp).or. &
         (Mxx/=0.0_dp.or.Mxy/=0.0_dp.or.Myy/=0.0_dp.or.Mzx/=0.0_dp.or.Mzy/=0.0_dp))) return
    if (pb_mode==3) then
       call fail('point sources are not available in mode 3')
       return
    end if
    np = 0
    do ib = 1,nblocks
       associate(G => HB(ib)%G)
         i0 = search_floor(gcx(G%mgx:G%pgx,G%mgy),x0,outx)+G%mgx-1
         if (outx) cycle
         j0 = search_floor(gcy(G%mgx,G%mgy:G%pgy),y0,outy)+G%mgy-1
         if (outy) cycle
         hx = gcx(G%mgx+1,G%mgy)-gcx(G%mgx,G%mgy)
         hy = gcy(G%mgx,G%mgy+1)-gcy(G%mgx,G%mgy)
         ax = (x0-gcx(i0,G%mgy))/hx
         ay = (y0-gcy(G%mgx,j0))/hy
         np = np+1
         if (np>FDM_MAX_SOURCE_PLACES) then
            call fail('point source found in too many blocks')
            return
         end if
         cprob%source%place(np)%block = ib-1
         cprob%source%place(np)%i0 = i0-1
         cprob%source%place(np)%j0 = j0-1
         call singular_source_weights(ax,hx,cprob%source%place(np)%dx)
         call singular_source_weights(ay,hy,cprob%source%place(np)%dy)
       end associate
    end do
    cprob%source%nplace = np
  end subroutine read_source


  subroutine fill_cproblem(precision)
    ! the fdm_problem description handed to fdm_create
    integer(c_int32_t),intent(in) :: precision
    integer :: ib,s,ii,nstage
    logical :: mms
    character(256) :: last_problem
    real(dp) :: A(8),B(8),C(9)
    allocate(cblk(max(1,nblocks)),cif(max(1,nifaces)))
    call rk_coefficients(RKorder,A,B,C,nstage)
    if (setup_error/='') return
    cprob%nx = gnx
    cprob%ny = gny
    cprob%nblocks = nblocks
    cprob%nifaces = nifaces
    cprob%precision = precision
    cprob%nstage = nstage
    cprob%rkA = A
    cprob%rkB = B
    cprob%rkC = C
    cprob%xq = c_loc(gxq(1,1))
    cprob%xr = c_loc(gxr(1,1))
    cprob%yq = c_loc(gyq(1,1))
    cprob%yr = c_loc(gyr(1,1))
    cprob%Ji = c_loc(gJi(1,1))
    cprob%x = c_loc(gcx(1,1))
    cprob%y = c_loc(gcy(1,1))
    cprob%J = c_loc(gJ(1,1))
    cprob%mode = pb_mode
    cprob%sbp = sbp_order
    if (allocated(gmrho)) then
       cprob%mat_rho = c_loc(gmrho(1,1))
       cprob%mat_cs = c_loc(gmcs(1,1))
       cprob%mat_cp = c_loc(gmcp(1,1))
    else
       cprob%mat_rho = c_null_ptr
       cprob%mat_cs = c_null_ptr
       cprob%mat_cp = c_null_ptr
    end if
    ! the fields problem of the last block selects the body source
    ! (fields_type%problem is shared, fields.f90:132; source.f90:256-296)
    last_problem = HB(nblocks)%problem
    mms = (trim(last_problem)=='mms-sin'.or.trim(last_problem)=='inplane-fault-mms')
    do ib = 1,nblocks
       associate(cb => cblk(ib), HBk => HB(ib))
         cb%mx = HBk%G%mgx-1
         cb%px = HBk%G%pgx-1
         cb%my = HBk%G%mgy-1
         cb%py = HBk%G%pgy-1
         if (trim(HBk%response)=='plastic') then
            cb%plastic = 1
         else
            cb%plastic = 0
         end if
         if (mms) then
            cb%mms = 1
         else
            cb%mms = 0
         end if
         cb%id = ib
         cb%pad_ = 0
         cb%rho = HBk%rho
         cb%G = HBk%G_
         cb%M = HBk%M_
         cb%gamma = HBk%gamma
         cb%nu = HBk%nu
         cb%mu = HBk%mu
         cb%beta = HBk%beta
         cb%b = HBk%b
         cb%h = HBk%h
         cb%eta = HBk%eta
         cb%K = HBk%K
         ! artificial dissipation scaled as (wave speed)/(grid spacing)
         ! (domain.f90:316-320)
         cb%Cdiss = HBk%Cdiss*CFL/dt
         cb%pmlx = merge(1,0,HBk%pmlx)
         cb%pmly = merge(1,0,HBk%pmly)
         cb%pml = HBk%pml
         ! initial_stress (fields.f90:1007-1058), default problem: the stress
         ! components that are not fields; in mode 3 the in-plane prestress
         ! that frictional interfaces resolve onto the fault
         if (pb_mode==3) then
            cb%s0 = (/ HBk%F0blk(4),HBk%F0blk(5),0.0_dp,HBk%F0blk(7),0.0_dp,HBk%F0blk(9) /)
         else
            cb%s0 = (/ 0.0_dp,0.0_dp,HBk%F0blk(6),0.0_dp,HBk%F0blk(8),0.0_dp /)
         end if
         do s = 1,4
            cb%side(s)%bc = HBk%side(s)%bc
            cb%side(s)%n = size(HBk%side(s)%x)
            cb%side(s)%normal = c_loc(HBk%side(s)%normal(1,1))
            cb%side(s)%M = c_loc(HBk%side(s)%M(1,1))
            cb%side(s)%F0 = c_loc(HBk%side(s)%F0(1,1))
            cb%side(s)%x = c_loc(HBk%side(s)%x(1))
            cb%side(s)%y = c_loc(HBk%side(s)%y(1))
         end do
       end associate
    end do
    do ii = 1,nifaces
       associate(ci => cif(ii), I => HI(ii))
         if (I%direction=='x') then
            ci%direction = 0
         else
            ci%direction = 1
         end if
         ci%blockm = I%iblockm-1
         ci%blockp = I%iblockp-1
         ci%m = I%mg-1
         ci%p = I%pg-1
         ci%pad_ = 0
         ci%nhat = c_loc(I%nhat(1,1))
         if (trim(I%coupling)=='locked') then
            ci%coupling = FDM_COUPLING_LOCKED
            ci%law = FDM_LAW_FRICTIONLESS
            ci%reg = 0
            ci%rup_field = 0
            ci%rup_threshold = 0.0_dp
            ci%skempton = 0.0_dp
            ci%flock = 1d10
            ci%ld_T = 0.0_dp
            ci%rs = c_null_ptr
            ci%sw = c_null_ptr
            ci%S0base = c_null_ptr
            ci%N0base = c_null_ptr
            ci%ldS = c_null_ptr
            ci%ldN = c_null_ptr
            ci%outside = c_null_ptr
            ci%Psi0 = c_null_ptr
         else
            ci%coupling = FDM_COUPLING_FRICTION
            ci%law = law_code(I%friction_law)
            if (I%reg) then
               ci%reg = 1
            else
               ci%reg = 0
            end if
            if (trim(I%rup_field)=='D') then
               ci%rup_field = 1
            else
               ci%rup_field = 0
            end if
            ci%rup_threshold = I%rup_threshold
            ci%skempton = I%skempton
            ci%flock = I%flock
            ci%ld_T = I%ld_T
            ci%rs = c_loc(I%rs(1,1))
            ci%sw = c_loc(I%sw(1,1))
            ci%S0base = c_loc(I%S0b(1))
            ci%N0base = c_loc(I%N0b(1))
            ci%ldS = c_loc(I%ldS(1))
            ci%ldN = c_loc(I%ldN(1))
            ci%outside = c_loc(I%outside(1))
            ci%Psi0 = c_loc(I%Psi0(1))
            ci%force = 0
            if (I%force) ci%force = 1
            ci%kn = I%kn
            ci%kn_x = c_loc(I%knx(1))
         end if
       end associate
    end do
    cprob%blocks = c_loc(cblk(1))
    cprob%ifaces = c_loc(cif(1))
  end subroutine fill_cproblem

end module fdmap_setup
