! fdmap_grid.f90 -- block grids, boundary curves, transfinite interpolation
! and SBP metrics on the host (FP64).
!
! Restates grid.f90:188-618, geometry.f90:70-200 and fd.f90:42-80 for one
! process holding every block: the metrics computed here are the inputs
! the device kernels consume (fdm_problem.xq/xr/yq/yr/Ji).

module fdmap_grid

  use iso_c_binding, only : c_double
  implicit none

  integer,parameter :: dp = c_double
  ! the first-derivative operator of the problem (FDmethod, fd_coeff.f90:
  ! 41-248), set by init_op: points of a boundary stencil, boundary
  ! stencils, interior half-width, DI(-pI:pI), DL(0:nbnd-1,0:nbst-1) and the
  ! SAT weight H00i; DR(-s,-r) = -DL(s,r) (:242-246)
  integer :: nbnd = 6, nbst = 4, pI = 2, sbp_order = 4
  real(dp) :: DI(-3:3) = 0.0_dp
  real(dp) :: DL(0:8,0:5) = 0.0_dp
  real(dp) :: H00i = 2.823529411764706_dp

  type :: point
     real(dp) :: x,y
  end type point

  type :: curve
     real(dp),allocatable :: x(:),y(:),n(:,:)
  end type curve

  ! one block: global 1-based index limits and its four edge curves
  type :: block_grid
     integer :: nx,ny,mgx,pgx,mgy,pgy,iblock_x,iblock_y
     character(256) :: bndLfile,bndRfile,bndBfile,bndTfile
     real(dp) :: bndLmap,bndRmap,bndBmap,bndTmap
     type(point) :: LB,LT,RB,RT
     type(curve) :: bndL,bndR,bndB,bndT
  end type block_grid

contains

  subroutine init_op(method,ok)
    ! init_fd (fd_coeff.f90:41-248) for 'SBP2', 'SBP4' and 'SBP6': the
    ! reference's 15-digit literals
    character(*),intent(in) :: method
    logical,intent(out) :: ok
    ok = .true.
    DI = 0.0_dp
    DL = 0.0_dp
    select case(method)
    case('SBP2')
       sbp_order = 2; nbnd = 2; nbst = 1; pI = 1
       H00i = 2.0_dp
       DI(-1:1) = (/ -0.5_dp, 0.0_dp, 0.5_dp /)
       DL(0:1,0) = (/ -1.0_dp, 1.0_dp /)
    case('SBP4')
       sbp_order = 4; nbnd = 6; nbst = 4; pI = 2
       H00i = 2.823529411764706_dp
       DI(-2:2) = (/ 0.083333333333333_dp, -0.666666666666667_dp, &
            0.000000000000000_dp, 0.666666666666667_dp, -0.083333333333333_dp /)
       DL(0:5,0) = (/ -1.411764705882353_dp,  1.735294117647059_dp, -0.235294117647059_dp, &
            -0.088235294117647_dp,  0.000000000000000_dp,  0.000000000000000_dp /)
       DL(0:5,1) = (/ -0.500000000000000_dp,  0.000000000000000_dp,  0.500000000000000_dp, &
            0.000000000000000_dp,  0.000000000000000_dp,  0.000000000000000_dp /)
       DL(0:5,2) = (/ 0.093023255813953_dp, -0.686046511627907_dp,  0.000000000000000_dp, &
            0.686046511627907_dp, -0.093023255813953_dp,  0.000000000000000_dp /)
       DL(0:5,3) = (/ 0.030612244897959_dp,  0.000000000000000_dp, -0.602040816326531_dp, &
            0.000000000000000_dp,  0.653061224489796_dp, -0.081632653061224_dp /)
    case('SBP6')
       sbp_order = 6; nbnd = 9; nbst = 6; pI = 3
       H00i = 3.165067037878233_dp
       DI = (/ -0.016666666666667_dp, 0.150000000000000_dp, -0.750000000000000_dp, &
            0.000000000000000_dp, 0.750000000000000_dp, -0.150000000000000_dp, &
            0.016666666666667_dp /)
       DL(:,0) = (/ -1.582533518939116_dp, 2.033378678700676_dp, -0.141512858744873_dp, &
            -0.450398306578272_dp, 0.104488069284042_dp, 0.036577936277544_dp, &
            0.000000000000000_dp, 0.000000000000000_dp, 0.000000000000000_dp /)
       DL(:,1) = (/ -0.462059195631158_dp, 0.000000000000000_dp, 0.287258622978251_dp, &
            0.258816087376832_dp, -0.069112065532624_dp, -0.014903449191300_dp, &
            0.000000000000000_dp, 0.000000000000000_dp, 0.000000000000000_dp /)
       DL(:,2) = (/ 0.071247104721830_dp, -0.636451095137907_dp, 0.000000000000000_dp, &
            0.606235523609147_dp, -0.022902190275815_dp, -0.018129342917256_dp, &
            0.000000000000000_dp, 0.000000000000000_dp, 0.000000000000000_dp /)
       DL(:,3) = (/ 0.114713313798970_dp, -0.290087484386815_dp, -0.306681191361148_dp, &
            0.000000000000000_dp, 0.520262285050482_dp, -0.051642265516119_dp, &
            0.013435342414630_dp, 0.000000000000000_dp, 0.000000000000000_dp /)
       DL(:,4) = (/ -0.036210680656541_dp, 0.105400944933782_dp, 0.015764336127392_dp, &
            -0.707905442575989_dp, 0.000000000000000_dp, 0.769199413962647_dp, &
            -0.164529643265203_dp, 0.018281071473911_dp, 0.000000000000000_dp /)
       DL(:,5) = (/ -0.011398193015050_dp, 0.020437334208704_dp, 0.011220896474665_dp, &
            0.063183694641876_dp, -0.691649024426814_dp, 0.000000000000000_dp, &
            0.739709139060752_dp, -0.147941827812150_dp, 0.016437980868017_dp /)
    case default
       ok = .false.
    end select
  end subroutine init_op


  function DR(s,r) result(c)
    ! DR(s,r) for s = -(nbnd-1)..0, r = -(nbst-1)..0
    integer,intent(in) :: s,r
    real(dp) :: c
    c = -DL(-s,-r)
  end function DR


  subroutine diff(mg,pg,f,Df)
    ! first derivative with unit spacing over one block, both block ends
    ! inside the range (fd.f90:42-80)
    integer,intent(in) :: mg,pg
    real(dp),intent(in) :: f(mg:pg)
    real(dp),intent(out) :: Df(mg:pg)
    integer :: i,k
    real(dp) :: c(0:8)
    if (mg==pg) then
       Df = 0.0_dp
       return
    end if
    do i = 0,nbst-1
       Df(mg+i) = dot_product(DL(0:nbnd-1,i),f(mg:mg+nbnd-1))
    end do
    do i = 0,nbst-1
       do k = 0,nbnd-1
          c(k) = DR(k-(nbnd-1),-i)
       end do
       Df(pg-i) = dot_product(c(0:nbnd-1),f(pg-nbnd+1:pg))
    end do
    do i = mg+nbst,pg-nbst
       Df(i) = dot_product(DI(-pI:pI),f(i-pI:i+pI))
    end do
  end subroutine diff


  subroutine line_map(C,P1,P2,m,p,t,t1,t2,B)
    ! straight edge from P1 to P2 with optional tanh clustering
    ! (geometry.f90:70-106)
    type(curve),intent(inout) :: C
    type(point),intent(in) :: P1,P2
    integer,intent(in) :: m,p
    real(dp),intent(in) :: t(m:p),t1,t2,B
    real(dp) :: L,w(m:p)
    if (m>p) return
    w = (t-t1)/(t2-t1)
    if (B>0.0_dp) then
       w = 1.0_dp+tanh(B*(w-1.0_dp))/tanh(B)
    elseif (B<0.0_dp) then
       w = tanh(B*w)/tanh(B)
    end if
    C%x(m:p) = P1%x+(P2%x-P1%x)*w
    C%y(m:p) = P1%y+(P2%y-P1%y)*w
    L = sqrt((P2%x-P1%x)**2+(P2%y-P1%y)**2)
    if (L/=0) then
       C%n(m:p,1) = (P2%y-P1%y)/L
       C%n(m:p,2) = (P1%x-P2%x)/L
    end if
  end subroutine line_map


  subroutine read_bnd(C,filename,m,p,n,ok)
    ! binary FP64 curve file: x(1:n), y(1:n), n1(1:n), n2(1:n)
    ! (grid.f90:646-702, format of matlab/fdmap_write_profile.m)
    type(curve),intent(inout) :: C
    character(*),intent(in) :: filename
    integer,intent(in) :: m,p,n
    logical,intent(out) :: ok
    integer :: u,stat
    real(dp) :: buf(n)
    ok = .false.
    open(newunit=u,file=filename,access='stream',form='unformatted',status='old', &
         action='read',iostat=stat)
    if (stat/=0) return
    read(u,iostat=stat) buf
    if (stat/=0) return
    C%x(m:p) = buf
    read(u,iostat=stat) buf
    if (stat/=0) return
    C%y(m:p) = buf
    read(u,iostat=stat) buf
    if (stat/=0) return
    C%n(m:p,1) = buf
    read(u,iostat=stat) buf
    if (stat/=0) return
    C%n(m:p,2) = buf
    close(u)
    ok = .true.
  end subroutine read_bnd


  subroutine transfinite_interp(mx,px,my,py,qmin,qmax,rmin,rmax,q,r,bndL,bndR,bndB,bndT, &
       LB,LT,RB,RT,x,y)
    ! Boyd's transfinite interpolation of the block interior from its four
    ! edges (geometry.f90:153-200)
    integer,intent(in) :: mx,px,my,py
    real(dp),intent(in) :: qmin,qmax,rmin,rmax,q(mx:px),r(my:py)
    type(curve),intent(in) :: bndL,bndR,bndB,bndT
    type(point),intent(in) :: LB,LT,RB,RT
    real(dp),intent(inout) :: x(:,:),y(:,:)
    integer :: i,j
    real(dp) :: Wq,Wr,aq,bq,ar,br
    Wq = qmax-qmin
    Wr = rmax-rmin
    if (Wq==0.0_dp) Wq = 1.0_dp
    if (Wr==0.0_dp) Wr = 1.0_dp
    do j = my,py
       ar = rmax-r(j)
       br = r(j)-rmin
       do i = mx,px
          aq = qmax-q(i)
          bq = q(i)-qmin
          x(i,j) = aq*bndL%x(j)/Wq+bq*bndR%x(j)/Wq+ar*bndB%x(i)/Wr+br*bndT%x(i)/Wr- &
               aq*ar*LB%x/(Wq*Wr)-aq*br*LT%x/(Wq*Wr)-bq*ar*RB%x/(Wq*Wr)-bq*br*RT%x/(Wq*Wr)
          y(i,j) = aq*bndL%y(j)/Wq+bq*bndR%y(j)/Wq+ar*bndB%y(i)/Wr+br*bndT%y(i)/Wr- &
               aq*ar*LB%y/(Wq*Wr)-aq*br*LT%y/(Wq*Wr)-bq*ar*RB%y/(Wq*Wr)-bq*br*RT%y/(Wq*Wr)
       end do
    end do
  end subroutine transfinite_interp


  subroutine build_block_coords(B,x,y,ok,msg)
    ! edges (straight lines or curve files) and interior of one block
    ! (grid.f90:391-462); x,y are the global 1-based coordinate arrays
    type(block_grid),intent(inout) :: B
    real(dp),intent(inout) :: x(:,:),y(:,:)
    logical,intent(out) :: ok
    character(*),intent(out) :: msg
    integer :: i,j
    real(dp) :: q(B%mgx:B%pgx),r(B%mgy:B%pgy)
    logical :: rok
    ok = .true.
    msg = ''
    q = (/ (dble(i), i=B%mgx,B%pgx) /)
    r = (/ (dble(j), j=B%mgy,B%pgy) /)
    allocate(B%bndL%x(B%mgy:B%pgy),B%bndL%y(B%mgy:B%pgy),B%bndL%n(B%mgy:B%pgy,2))
    allocate(B%bndR%x(B%mgy:B%pgy),B%bndR%y(B%mgy:B%pgy),B%bndR%n(B%mgy:B%pgy,2))
    allocate(B%bndB%x(B%mgx:B%pgx),B%bndB%y(B%mgx:B%pgx),B%bndB%n(B%mgx:B%pgx,2))
    allocate(B%bndT%x(B%mgx:B%pgx),B%bndT%y(B%mgx:B%pgx),B%bndT%n(B%mgx:B%pgx,2))
    B%bndL%n = 1d40
    B%bndR%n = 1d40
    B%bndB%n = 1d40
    B%bndT%n = 1d40
    if (B%bndLfile=='') then
       call line_map(B%bndL,B%LB,B%LT,B%mgy,B%pgy,r,dble(B%mgy),dble(B%pgy),B%bndLmap)
    else
       call read_bnd(B%bndL,trim(B%bndLfile),B%mgy,B%pgy,B%ny,rok)
       if (.not.rok) then
          ok = .false.; msg = 'cannot read curve file '//trim(B%bndLfile); return
       end if
    end if
    B%bndL%n = -B%bndL%n
    if (B%bndRfile=='') then
       call line_map(B%bndR,B%RB,B%RT,B%mgy,B%pgy,r,dble(B%mgy),dble(B%pgy),B%bndRmap)
    else
       call read_bnd(B%bndR,trim(B%bndRfile),B%mgy,B%pgy,B%ny,rok)
       if (.not.rok) then
          ok = .false.; msg = 'cannot read curve file '//trim(B%bndRfile); return
       end if
    end if
    if (B%bndBfile=='') then
       call line_map(B%bndB,B%LB,B%RB,B%mgx,B%pgx,q,dble(B%mgx),dble(B%pgx),B%bndBmap)
    else
       call read_bnd(B%bndB,trim(B%bndBfile),B%mgx,B%pgx,B%nx,rok)
       if (.not.rok) then
          ok = .false.; msg = 'cannot read curve file '//trim(B%bndBfile); return
       end if
    end if
    if (B%bndTfile=='') then
       call line_map(B%bndT,B%LT,B%RT,B%mgx,B%pgx,q,dble(B%mgx),dble(B%pgx),B%bndTmap)
    else
       call read_bnd(B%bndT,trim(B%bndTfile),B%mgx,B%pgx,B%nx,rok)
       if (.not.rok) then
          ok = .false.; msg = 'cannot read curve file '//trim(B%bndTfile); return
       end if
    end if
    B%bndT%n = -B%bndT%n
    call transfinite_interp(B%mgx,B%pgx,B%mgy,B%pgy,dble(B%mgx),dble(B%pgx),dble(B%mgy), &
         dble(B%pgy),q,r,B%bndL,B%bndR,B%bndB,B%bndT,B%LB,B%LT,B%RB,B%RT,x,y)
  end subroutine build_block_coords


  subroutine block_metrics(B,x,y,xq,xr,yq,yr,J,Ji,hmin,hmax)
    ! SBP derivatives of the coordinates, Jacobian and effective grid
    ! spacings (grid.f90:465-565)
    type(block_grid),intent(in) :: B
    real(dp),intent(in) :: x(:,:),y(:,:)
    real(dp),intent(inout) :: xq(:,:),xr(:,:),yq(:,:),yr(:,:),J(:,:),Ji(:,:)
    real(dp),intent(inout) :: hmin,hmax
    integer :: i,k
    real(dp) :: dx,dy
    real(dp),allocatable :: f(:),Df(:)
    allocate(f(B%mgx:B%pgx),Df(B%mgx:B%pgx))
    do k = B%mgy,B%pgy
       f = x(B%mgx:B%pgx,k)
       call diff(B%mgx,B%pgx,f,Df)
       xq(B%mgx:B%pgx,k) = Df
       f = y(B%mgx:B%pgx,k)
       call diff(B%mgx,B%pgx,f,Df)
       yq(B%mgx:B%pgx,k) = Df
    end do
    deallocate(f,Df)
    allocate(f(B%mgy:B%pgy),Df(B%mgy:B%pgy))
    do i = B%mgx,B%pgx
       f = x(i,B%mgy:B%pgy)
       call diff(B%mgy,B%pgy,f,Df)
       xr(i,B%mgy:B%pgy) = Df
       f = y(i,B%mgy:B%pgy)
       call diff(B%mgy,B%pgy,f,Df)
       yr(i,B%mgy:B%pgy) = Df
    end do
    J(B%mgx:B%pgx,B%mgy:B%pgy) = xq(B%mgx:B%pgx,B%mgy:B%pgy)*yr(B%mgx:B%pgx,B%mgy:B%pgy)- &
         xr(B%mgx:B%pgx,B%mgy:B%pgy)*yq(B%mgx:B%pgx,B%mgy:B%pgy)
    Ji(B%mgx:B%pgx,B%mgy:B%pgy) = 1.0_dp/J(B%mgx:B%pgx,B%mgy:B%pgy)
    do k = B%mgy,B%pgy
       do i = B%mgx,B%pgx
          dx = abs(J(i,k)/sqrt(xr(i,k)**2+yr(i,k)**2))
          dy = abs(J(i,k)/sqrt(xq(i,k)**2+yq(i,k)**2))
          hmin = min(hmin,dx,dy)
          hmax = max(hmax,dx,dy)
       end do
    end do
  end subroutine block_metrics

end module fdmap_grid
