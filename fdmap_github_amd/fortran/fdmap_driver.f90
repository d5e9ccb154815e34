! fdmap_driver.f90 -- the Fortran driver: output items, the time-step
! loop and the C-callable entry points of libfdmap_host.so.
!
! The loop keeps the reference's semantics (main.f90:163-259): nt
! iterations, D%t = t+(n-1)*dt, the last iteration computes stage-1 values
! only, outputs are written at stage 1 of each output step with the step
! index of the previous iteration (time_step.f90:129, output.f90:1011-1074),
! nucleation check and the final 'timing:' line.  Each time_step_LS call is
! replaced by fdm_step / fdm_stage1 of the MI355X library.

module fdmap_driver

  use iso_c_binding
  use fdmap_capi
  use fdmap_grid, only : dp
  use fdmap_setup
  use fdmap_checkpoint
  implicit none

  type :: out_item
     character(256) :: name,location
     character(6) :: field
     integer :: iface = -1, block = -1
     integer :: nmin,nmax,st,nx,ny
     real(dp) :: tmin,tmax
     integer :: ufile = -1, utime = -1
  end type out_item

  type(out_item),allocatable :: items(:)
  integer :: nitems = 0
  type(c_ptr) :: dev_ctx = c_null_ptr
  character(kind=c_char,len=513),target :: c_errbuf

contains

  subroutine message(str)
    character(*),intent(in) :: str
    write(6,'(a)') trim(str)
    flush(6)
  end subroutine message


  function walltime() result(t)
    real(dp) :: t
    integer(8) :: count,rate
    call system_clock(count,rate)
    t = dble(count)/dble(rate)
  end function walltime


  subroutine word_split(str,words,nw)
    character(*),intent(in) :: str
    character(6),intent(out) :: words(:)
    integer,intent(out) :: nw
    integer :: k,start,n
    n = len_trim(str)
    nw = 0
    k = 1
    do while (k<=n)
       do while (k<=n)
          if (str(k:k)/=' ') exit
          k = k+1
       end do
       if (k>n) exit
       start = k
       do while (k<=n)
          if (str(k:k)==' ') exit
          k = k+1
       end do
       nw = nw+1
       if (nw<=size(words)) words(nw) = str(start:k-1)
    end do
  end subroutine word_split


  subroutine read_output(infile)
    ! output list between '!---begin:output_list---' and
    ! '!---end:output_list---' (output.f90:67-294); locations body,
    ! block<k> and iface<k>
    character(*),intent(in) :: infile
    integer :: u,stat,nw,k,st
    character(256) :: name,loc,temp
    character(6) :: flds(64)
    real(dp) :: tmin,tmax
    type(out_item),allocatable :: tmp(:)
    type(out_item) :: it
    allocate(items(0))
    nitems = 0
    open(newunit=u,file=infile,status='old',action='read',iostat=stat)
    if (stat/=0) return
    do
       read(u,'(a)',iostat=stat) temp
       if (stat/=0) then
          close(u)
          return
       end if
       if (temp=='!---begin:output_list---') exit
    end do
    do
       read(u,'(a)',iostat=stat) name
       if (stat/=0) then
          call fail("Error reading list 'output' in input file")
          exit
       end if
       if (name=='!---end:output_list---') exit
       read(u,'(a)') loc
       read(u,'(a)') temp
       call word_split(temp,flds,nw)
       read(u,*) tmin,tmax,st
       it%iface = -1
       it%block = -1
       if (index(loc,'iface')==1) then
          read(loc(6:),'(i6)') it%iface
          if (it%iface<1.or.it%iface>nifaces) call fail('Invalid iface number')
          it%location = 'iface'
       elseif (index(loc,'block')==1) then
          read(loc(6:),'(i6)') it%block
          if (it%block<1.or.it%block>nblocks) call fail('Invalid block number')
          it%location = 'block'
       elseif (trim(loc)=='body') then
          it%location = 'body'
       else
          call fail('Invalid output location: '//trim(loc))
       end if
       if (setup_error/='') exit
       it%tmin = tmin
       it%tmax = tmax
       it%st = st
       ! output time-step window (output.f90:412-418)
       it%nmin = ceiling((tmin-t_init)/dt)
       if ((tmax-t_init)/dt>dble(huge(1))) then
          it%nmax = huge(1)
       else
          it%nmax = floor((tmax-t_init)/dt)
       end if
       select case(trim(it%location))
       case('body')
          it%nx = gnx
          it%ny = gny
       case('block')
          it%nx = HB(it%block)%G%nx
          it%ny = HB(it%block)%G%ny
       case('iface')
          it%nx = HI(it%iface)%n
          it%ny = 1
       end select
       do k = 1,nw
          it%name = trim(name)//'_'//trim(flds(k))
          it%field = flds(k)
          allocate(tmp(nitems+1))
          tmp(1:nitems) = items
          tmp(nitems+1) = it
          call move_alloc(tmp,items)
          nitems = nitems+1
       end do
    end do
    close(u)
  end subroutine read_output


  subroutine open_outputs(begin)
    ! <name>_<item>.dat (FP32 data) and <name>_<item>.t (FP32 times)
    ! (output.f90:916-987).  A run resumed after step begin keeps the files
    ! and writes from the record of that step on (or the next recorded one)
    integer,intent(in) :: begin
    integer :: k,noffset
    integer(8) :: off
    character(16) :: st
    do k = 1,nitems
       st = 'replace'
       noffset = 0
       if (begin>0) then
          st = 'unknown'
          if (begin>=items(k)%nmin) then
             noffset = (begin-items(k)%nmin)/items(k)%st
             if (mod(begin-items(k)%nmin,items(k)%st)/=0) noffset = noffset+1
          end if
          if (noffset>items(k)%nmax) noffset = 0
       end if
       open(newunit=items(k)%ufile,file=trim(pb_name)//'_'//trim(items(k)%name)//'.dat', &
            access='stream',form='unformatted',status=trim(st),action='write')
       open(newunit=items(k)%utime,file=trim(pb_name)//'_'//trim(items(k)%name)//'.t', &
            access='stream',form='unformatted',status=trim(st),action='write')
       if (begin>0) then
          off = int(noffset,8)*int(items(k)%nx,8)*int(items(k)%ny,8)*4_8
          write(items(k)%ufile,pos=off+1_8)
          write(items(k)%utime,pos=int(noffset,8)*4_8+1_8)
       end if
    end do
  end subroutine open_outputs


  subroutine close_outputs()
    integer :: k
    do k = 1,nitems
       if (items(k)%ufile>0) close(items(k)%ufile)
       if (items(k)%utime>0) close(items(k)%utime)
    end do
  end subroutine close_outputs


  function item_due(it,m) result(due)
    type(out_item),intent(in) :: it
    integer,intent(in) :: m
    logical :: due
    due = .false.
    if (m<it%nmin.or.m>it%nmax) return
    due = (mod(m-it%nmin,it%st)==0)
  end function item_due


  subroutine write_echo()
    ! MATLAB echo <name>.m, read back by matlab/init.m (io.f90:618-760)
    integer :: u,k,stat
    open(newunit=u,file=trim(pb_name)//'.m',status='replace',action='write',iostat=stat)
    if (stat/=0) return
    call wm_c(u,'endian','l')
    call wm_i(u,'nt',nt)
    call wm_r(u,'refine',refine)
    call wm_c(u,'method','LS','RK')
    call wm_i(u,'order',RKorder,'RK')
    call wm_i(u,'mode',pb_mode)
    call wm_c(u,'FDmethod',merge('SBP2',merge('SBP4','SBP6',sbp_order==4),sbp_order==2))
    call wm_i(u,'nblocks',nblocks)
    call wm_i(u,'nifaces',nifaces)
    call wm_i(u,'nx',gnx)
    call wm_i(u,'ny',gny)
    call wm_r(u,'dt',dt)
    call wm_r(u,'CFL',CFL)
    call wm_r(u,'hmin',hmin)
    call wm_r(u,'hmax',hmax)
    call wm_r(u,'h',0.5_dp*(hmin+hmax))
    do k = 1,nitems
       call wm_c(u,'name',items(k)%name,items(k)%name)
       call wm_c(u,'field',items(k)%field,items(k)%name)
       if (trim(items(k)%location)=='iface') then
          if (HI(items(k)%iface)%direction=='x') then
             call wm_c(u,'location','ifacex',items(k)%name)
          else
             call wm_c(u,'location','ifacey',items(k)%name)
          end if
       else
          call wm_c(u,'location',items(k)%location,items(k)%name)
       end if
       call wm_r(u,'tmin',items(k)%tmin,items(k)%name)
       call wm_r(u,'tmax',items(k)%tmax,items(k)%name)
       call wm_i(u,'stride_t',items(k)%st,items(k)%name)
       call wm_i(u,'nx',items(k)%nx,items(k)%name)
       call wm_i(u,'ny',items(k)%ny,items(k)%name)
       if (items(k)%block/=-1) call wm_i(u,'block',items(k)%block,items(k)%name)
       if (items(k)%iface/=-1) call wm_i(u,'iface',items(k)%iface,items(k)%name)
    end do
    close(u)
  contains
    subroutine wm_c(u,name,str,struct)
      integer,intent(in) :: u
      character(*),intent(in) :: name,str
      character(*),intent(in),optional :: struct
      if (present(struct)) then
         write(u,'(a)') trim(adjustl(struct))//'.'//trim(adjustl(name))//" = '"//trim(adjustl(str))//"';"
      else
         write(u,'(a)') trim(adjustl(name))//" = '"//trim(adjustl(str))//"';"
      end if
    end subroutine wm_c
    subroutine wm_i(u,name,v,struct)
      integer,intent(in) :: u,v
      character(*),intent(in) :: name
      character(*),intent(in),optional :: struct
      if (present(struct)) then
         write(u,'(a,i12,a)') trim(adjustl(struct))//'.'//trim(adjustl(name))//' = ',v,';'
      else
         write(u,'(a,i12,a)') trim(adjustl(name))//' = ',v,';'
      end if
    end subroutine wm_i
    subroutine wm_r(u,name,v,struct)
      integer,intent(in) :: u
      real(dp),intent(in) :: v
      character(*),intent(in) :: name
      character(*),intent(in),optional :: struct
      if (present(struct)) then
         write(u,'(a,e30.15,a)') trim(adjustl(struct))//'.'//trim(adjustl(name))//' = ',v,';'
      else
         write(u,'(a,e30.15,a)') trim(adjustl(name))//' = ',v,';'
      end if
    end subroutine wm_r
  end subroutine write_echo


  subroutine write_item(it,t,F,gp,lam)
    ! one output record of one item (domain.f90:502-894)
    type(out_item),intent(in) :: it
    real(dp),intent(in) :: t
    real(dp),intent(in) :: F(:,:,:),gp(:,:),lam(:,:)
    integer :: mx,px,my,py,ii,field,n,rc
    real(dp),allocatable,target :: buf(:)
    real(c_float),allocatable :: rec(:,:)
    write(it%utime) real(t,c_float)
    select case(trim(it%location))
    case('body','block')
       if (trim(it%location)=='body') then
          mx = 1; px = gnx; my = 1; py = gny
       else
          mx = HB(it%block)%G%mgx; px = HB(it%block)%G%pgx
          my = HB(it%block)%G%mgy; py = HB(it%block)%G%pgy
       end if
       allocate(rec(mx:px,my:py))
       select case(trim(it%field))
       case('x');   rec = real(gcx(mx:px,my:py),c_float)
       case('y');   rec = real(gcy(mx:px,my:py),c_float)
       case('xq');  rec = real(gxq(mx:px,my:py),c_float)
       case('xr');  rec = real(gxr(mx:px,my:py),c_float)
       case('yq');  rec = real(gyq(mx:px,my:py),c_float)
       case('yr');  rec = real(gyr(mx:px,my:py),c_float)
       case('J');   rec = real(gJ(mx:px,my:py),c_float)
       case('vx','vy','sxx','sxy','syy','szz','vz','sxz','syz')
          ! the fields of this mode, in their order (fields.f90:158-175)
          if (pb_mode==3) then
             field = index('vz  sxz syz ',trim(it%field)//' ')
          else
             field = index('vx  vy  sxx sxy syy szz ',trim(it%field)//' ')
          end if
          if (field>0) then
             rec = real(F(mx:px,my:py,(field-1)/4+1),c_float)
          else
             rec = 0.0
          end if
       case('gammap'); rec = real(gp(mx:px,my:py),c_float)
       case('lambda'); rec = real(lam(mx:px,my:py),c_float)
       case default
          rec = 0.0
       end select
       write(it%ufile) rec
    case('iface')
       ii = it%iface
       n = HI(ii)%n
       allocate(buf(n))
       field = -1
       select case(trim(it%field))
       case('x'); buf = HI(ii)%x
       case('y'); buf = HI(ii)%y
       case('V'); field = FDM_IF_V
       case('O'); field = FDM_IF_O
       case('S'); field = FDM_IF_S
       case('N'); field = FDM_IF_N
       case('W'); field = FDM_IF_W
       case('S0'); field = FDM_IF_S0
       case('N0'); field = FDM_IF_N0
       case('Ds'); field = FDM_IF_DS
       case('Dn'); field = FDM_IF_DN
       case('D'); field = FDM_IF_D
       case('Psi'); field = FDM_IF_PSI
       case('trup'); field = FDM_IF_TRUP
       case('Nreg'); field = FDM_IF_NREG
       case('a','b','V0','f0','L','fw','Vw')
          if (allocated(HI(ii)%rs)) then
             select case(trim(it%field))
             case('a');  buf = HI(ii)%rs(:,1)
             case('b');  buf = HI(ii)%rs(:,2)
             case('V0'); buf = HI(ii)%rs(:,3)
             case('f0'); buf = HI(ii)%rs(:,4)
             case('L');  buf = HI(ii)%rs(:,5)
             case('fw'); buf = HI(ii)%rs(:,6)
             case('Vw'); buf = HI(ii)%rs(:,7)
             end select
          else
             buf = 0.0_dp
          end if
       case default
          buf = 0.0_dp
       end select
       if (field>=0) rc = fdm_get_iface_stage1(dev_ctx,ii-1,field,c_loc(buf(1)))
       write(it%ufile) real(buf,c_float)
    end select
  end subroutine write_item


  function run_problem(infile,precision) result(rc)
    ! main.f90:32-271 with time_step_LS on the device
    character(*),intent(in) :: infile
    integer(c_int32_t),intent(in) :: precision
    integer :: rc
    logical :: ok,final_step,due,slipping,need_body,need_plastic,abort,deferred
    integer :: n,k,nnext,nrun,stat,nstart
    integer(c_int64_t) :: n64
    real(dp) :: t,wstart,w1,w2,wave,vmax,wprog
    real(dp),allocatable,target :: F(:,:,:),gp(:,:),lam(:,:),Fs(:,:,:)
    character(512) :: str
    ! body and block outputs in flight: the snapshot slot, step and time of
    ! each, oldest first (written when read back)
    integer :: npend
    integer(c_int32_t) :: slot
    integer :: pslot(FDM_SNAPSHOT_SLOTS),pn(FDM_SNAPSHOT_SLOTS)
    real(dp) :: pt(FDM_SNAPSHOT_SLOTS)

    rc = 0
    wprog = walltime()
    call message('Reading input file '//trim(adjustl(infile)))
    call message('Number of processors = 1')
    call setup_problem(infile,precision,ok)
    if (.not.ok) then
       call message('Error: '//trim(setup_error))
       rc = 1
       return
    end if
    call message('Starting problem: '//trim(adjustl(pb_name)))
    call read_output(infile)
    if (setup_error=='') call init_checkpoint(infile)
    if (setup_error/='') then
       call message('Error: '//trim(setup_error))
       rc = 1
       return
    end if
    stat = fdm_create(cprob,dev_ctx)
    if (stat/=FDM_OK) then
       call message('Error: fdm_create: '//fdm_error_message(c_null_ptr))
       rc = 1
       return
    end if
    stat = fdm_set_fields(dev_ctx,c_loc(gF(1,1,1)),c_null_ptr)
    if (stat==FDM_OK) stat = fdm_init_edges(dev_ctx,t_init)
    if (stat/=FDM_OK) then
       call message('Error: '//fdm_error_message(dev_ctx))
       rc = 1
       return
    end if
    ! resume: fields, rates, plastic strain and fault state of step begin
    ! (main.f90:150-154)
    call ckpt_read(dev_ctx,ok)
    if (.not.ok) then
       if (setup_error/='') then
          call message('Error: '//trim(setup_error))
       else
          call message('Error: '//fdm_error_message(dev_ctx))
       end if
       rc = 1
       return
    end if
    nstart = ck%begin+1
    call write_echo()
    call open_outputs(ck%begin)
    allocate(F(gnx,gny,6),gp(gnx,gny),lam(gnx,gny),Fs(gnx,gny,6))
    gp = 0.0_dp
    lam = 0.0_dp
    npend = 0
    call message('Initialization complete')

    wstart = walltime()
    wave = 0.0_dp
    abort = .false.
    n = nstart
    do while (n<=nt)
       w1 = walltime()
       t = t_init+dble(n-1)*dt
       final_step = (n==nt)
       ! outputs of this iteration use the step index n-1 (D%n is updated
       ! after time_step_LS, main.f90:189)
       due = .false.
       need_body = .false.
       need_plastic = .false.
       do k = 1,nitems
          if (item_due(items(k),n-1)) then
             due = .true.
             if (trim(items(k)%location)/='iface') need_body = .true.
             if (trim(items(k)%field)=='gammap'.or.trim(items(k)%field)=='lambda') need_plastic = .true.
          end if
       end do
       ! body and block outputs stream: a device snapshot of the fields now,
       ! copied to the host while the next steps run, written when read
       ! back (at most FDM_SNAPSHOT_SLOTS-1 in flight).  Plastic outputs
       ! take the synchronous path, after the outputs in flight.
       deferred = .false.
       if (need_body.and.(.not.need_plastic)) then
          if (npend==FDM_SNAPSHOT_SLOTS-1) call flush_outputs(1)
          stat = fdm_snapshot(dev_ctx,int(2**nF_mode()-1,c_int32_t),slot)
          if (stat==FDM_OK) then
             npend = npend+1
             pslot(npend) = slot
             pn(npend) = n-1
             pt(npend) = t
             deferred = .true.
          end if
       end if
       if ((need_body.or.need_plastic).and.(.not.deferred)) then
          call flush_outputs(npend)
          if (need_body) stat = fdm_get_fields(dev_ctx,c_loc(F(1,1,1)),c_null_ptr)
          if (need_plastic) stat = fdm_get_plastic(dev_ctx,c_loc(gp(1,1)),c_loc(lam(1,1)))
       end if
       if (final_step) then
          stat = fdm_stage1(dev_ctx,t_init,dt,int(n,c_int64_t))
          nrun = 1
       elseif (due.or.within(tnuc_min,t+dt,tnuc_max).or.mod(n-1,ninfo)==0.or.n>=nt-1) then
          stat = fdm_step(dev_ctx,t_init,dt,int(n,c_int64_t),1_c_int32_t)
          nrun = 1
       else
          ! run ahead to the next step that needs host attention
          nnext = n+1
          do while (nnext<nt-1)
             if (mod(nnext-1,ninfo)==0) exit
             if (within(tnuc_min,t_init+dble(nnext)*dt,tnuc_max)) exit
             ok = .false.
             do k = 1,nitems
                if (item_due(items(k),nnext-1)) ok = .true.
             end do
             if (ok) exit
             ! a checkpoint is due after step nnext-1
             if (ck%ckpt) then
                if (mod(nnext-1,ck%interval)==0) exit
             end if
             nnext = nnext+1
          end do
          nrun = nnext-n
          n64 = n
          stat = fdm_step(dev_ctx,t_init,dt,n64,int(nrun,c_int32_t))
       end if
       if (stat/=FDM_OK) then
          call message('Error: '//fdm_error_message(dev_ctx))
          rc = 1
          exit
       end if
       if (due) then
          do k = 1,nitems
             if (.not.item_due(items(k),n-1)) cycle
             if (deferred.and.trim(items(k)%location)/='iface') cycle
             call write_item(items(k),t,F,gp,lam)
          end do
       end if
       n = n+nrun
       ! abort request and checkpoint after step n-1 (main.f90:191-198); an
       ! abort is seen at the end of a run of steps
       abort = abort_now(walltime()-wprog)
       if (abort) call message('Aborting now...')
       ! a checkpoint records the output files as they stand
       if (ck%ckpt.or.abort) call flush_outputs(npend)
       call ckpt_write_delete(dev_ctx,n-1,abort,ok)
       if (.not.ok) then
          if (setup_error/='') then
             call message('Error: '//trim(setup_error))
          else
             call message('Error: '//fdm_error_message(dev_ctx))
          end if
          rc = 1
          exit
       end if
       ! status update (main.f90:202-211), for the last step of the run
       if (mod(n-2,ninfo)==0) then
          stat = fdm_sync(dev_ctx)
          w2 = walltime()
          write(str,'(a,i0,a,i0,a,f0.6,a,f0.6,a,f0.6,a)') 'time step ',n-1,' of ',nt,': t = ', &
               t_init+dble(n-2)*dt,' (wall clock: ',w2-wstart,' s, ',(w2-w1)/dble(nrun),' s/step)'
          call message(str)
       end if
       ! nucleation check (main.f90:215-221) on D%t after the step
       if (.not.final_step) then
          if (within(tnuc_min,t_init+dble(n-2)*dt+dt,tnuc_max)) then
             stat = fdm_max_slip_velocity(dev_ctx,vmax)
             slipping = (vmax>=minV)
             if (.not.slipping) then
                call message('Fault slip velocity too small (nucleation failed) -- terminating run...')
                exit
             end if
          end if
       end if
       if (abort) exit
       if (n-1==nt-1) then
          stat = fdm_sync(dev_ctx)
          wave = (walltime()-wstart)/dble(max(1,nt-nstart))
       end if
    end do
    call flush_outputs(npend)
    stat = fdm_sync(dev_ctx)
    ! the last checkpoint goes once the run has finished (main.f90:233-235)
    if (rc==0.and.(.not.abort).and.ck%del_final) call ckpt_delete(.true.)
    call close_outputs()
    write(str,'(a,f0.8,a,i0,a,f0.4,a)') 'total: ',walltime()-wstart,' s for ',nt, &
         ' time steps (',wave,' s/step)'
    call message(str)
    write(str,'(a,i0,a,f0.8)') 'timing: ',1,'  ',wave
    call message(str)
    if (n-1==nt.and.rc==0) then
       call message('Finished problem: '//trim(adjustl(pb_name)))
    else
       call message('Early termination of problem: '//trim(adjustl(pb_name)))
    end if
    call fdm_destroy(dev_ctx)
    dev_ctx = c_null_ptr
  contains
    function within(a,x,b) result(w)
      real(dp),intent(in) :: a,x,b
      logical :: w
      w = (a<=x.and.x<=b)
    end function within

    function nF_mode() result(m)
      integer :: m
      m = merge(3,6,pb_mode==3)
    end function nF_mode

    subroutine flush_outputs(m)
      ! read back the m oldest snapshots and write their body and block items
      integer,intent(in) :: m
      integer :: r,kk,st
      do r = 1,m
         st = fdm_snapshot_read(dev_ctx,int(pslot(r),c_int32_t),c_loc(Fs(1,1,1)))
         if (st/=FDM_OK) then
            call message('Error: '//fdm_error_message(dev_ctx))
            cycle
         end if
         do kk = 1,nitems
            if (.not.item_due(items(kk),pn(r))) cycle
            if (trim(items(kk)%location)=='iface') cycle
            call write_item(items(kk),pt(r),Fs,gp,lam)
         end do
      end do
      if (m>0) then
         pslot(1:npend-m) = pslot(m+1:npend)
         pn(1:npend-m) = pn(m+1:npend)
         pt(1:npend-m) = pt(m+1:npend)
         npend = npend-m
      end if
    end subroutine flush_outputs
  end function run_problem


  ! ---------------------------------------------------------------
  ! C entry points (Python binding, tests, bench)
  ! ---------------------------------------------------------------

  function c_to_f_string(s) result(f)
    character(kind=c_char),intent(in) :: s(*)
    character(len=:),allocatable :: f
    integer :: n,k
    n = 0
    do while (s(n+1)/=c_null_char)
       n = n+1
    end do
    allocate(character(len=n) :: f)
    do k = 1,n
       f(k:k) = s(k)
    end do
  end function c_to_f_string


  function fdmap_host_setup(path,precision) bind(C,name='fdmap_host_setup') result(rc)
    ! parse an .in file and build the fdm_problem description
    character(kind=c_char),intent(in) :: path(*)
    integer(c_int32_t),value :: precision
    integer(c_int) :: rc
    logical :: ok
    call setup_problem(c_to_f_string(path),precision,ok)
    rc = 0
    if (.not.ok) rc = -1
  end function fdmap_host_setup


  function fdmap_host_error() bind(C,name='fdmap_host_error') result(p)
    type(c_ptr) :: p
    c_errbuf = trim(setup_error)//c_null_char
    p = c_loc(c_errbuf)
  end function fdmap_host_error


  function fdmap_host_problem() bind(C,name='fdmap_host_problem') result(p)
    type(c_ptr) :: p
    p = c_loc(cprob)
  end function fdmap_host_problem


  function fdmap_host_fields() bind(C,name='fdmap_host_fields') result(p)
    ! initial fields [6][ny][nx]
    type(c_ptr) :: p
    p = c_loc(gF(1,1,1))
  end function fdmap_host_fields


  subroutine fdmap_host_info(t,dt_out,nt_out,cfl_out,hmin_out,nx,ny) bind(C,name='fdmap_host_info')
    real(c_double),intent(out) :: t,dt_out,cfl_out,hmin_out
    integer(c_int32_t),intent(out) :: nt_out,nx,ny
    t = t_init
    dt_out = dt
    nt_out = nt
    cfl_out = CFL
    hmin_out = hmin
    nx = gnx
    ny = gny
  end subroutine fdmap_host_info


  function fdmap_host_coords(which) bind(C,name='fdmap_host_coords') result(p)
    ! 0: x, 1: y, 2: xq, 3: xr, 4: yq, 5: yr, 6: J, 7: Ji ([ny][nx])
    integer(c_int32_t),value :: which
    type(c_ptr) :: p
    select case(which)
    case(0); p = c_loc(gcx(1,1))
    case(1); p = c_loc(gcy(1,1))
    case(2); p = c_loc(gxq(1,1))
    case(3); p = c_loc(gxr(1,1))
    case(4); p = c_loc(gyq(1,1))
    case(5); p = c_loc(gyr(1,1))
    case(6); p = c_loc(gJ(1,1))
    case default; p = c_loc(gJi(1,1))
    end select
  end function fdmap_host_coords


  function fdmap_host_run(path,precision) bind(C,name='fdmap_host_run') result(rc)
    ! the whole driver (what the fdmap executable runs)
    character(kind=c_char),intent(in) :: path(*)
    integer(c_int32_t),value :: precision
    integer(c_int) :: rc
    rc = run_problem(c_to_f_string(path),precision)
  end function fdmap_host_run

end module fdmap_driver
