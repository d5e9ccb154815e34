! fdmap_checkpoint.f90 -- checkpoint, resume and abort for the Fortran
! driver (checkpoint.f90 of the reference), with the device as the holder of
! the state.
!
! Files, all FP64 in native byte order with no headers, as the reference
! writes them through MPI-IO on one process:
!   <name>.ckpt                    number of the latest checkpoint (list-
!                                  directed integer; -1 once a run finished)
!   <name>F.ckpt<n>                F(nF planes), DF(nF planes), then gammap
!                                  and lambda when block 1 is plastic, then
!                                  the PML fields Wx (6 planes) when a block
!                                  sets pmlx and Wy when one sets pmly, each
!                                  plane nx*ny, x fastest (fields.f90:335-429);
!                                  nF = 6 in plane strain, 3 in antiplane shear
!   <name>block<i><L|R|B|T>.ckpt<n>  per side: F, Fhat, F0 of each of the nF
!                                  components, U and DU of nU components (2,
!                                  or 1 in antiplane shear), E, DE
!                                  (fields.f90:158-175, 432-644, 848-860)
!   <name>iface<k>side<1|2>.ckpt<n>  the 18 fault arrays of a frictional
!                                  interface in checkpoint order (friction.f90:
!                                  491-552; interfaces.f90:329-398); empty for
!                                  a locked one
!
! The side files hold values that no resume reads: F and Fhat are recomputed
! by the edge stage before any use (time_step.f90:94-99), F0 is rebuilt by
! the setup, U stays 0 and DU, E, DE keep their initial values without
! displacement or energy_balance, which this host refuses.  They are written
! with those initial values (U = E = 0, DU = DE = 1d40) and, for F and Fhat,
! with the reference's own unset fill 1d40, so the reference can resume from
! them; a resume here reads the field and interface files only.

module fdmap_checkpoint

  use iso_c_binding
  use fdmap_capi
  use fdmap_grid, only : dp
  use fdmap_setup
  implicit none

  type :: ckpt_type
     logical :: ckpt = .false., resume = .false., del_final = .true.
     integer :: begin = 0, interval = 1, previous = 0
     real(dp) :: max_time = 1d40
     character(512) :: filename = ''
  end type ckpt_type

  type(ckpt_type) :: ck

  real(dp),parameter :: unset = 1d40

  ! PML auxiliary fields [nx,ny,6] between the device and the files
  real(dp),allocatable,target :: ckWx(:,:,:),ckWy(:,:,:)

contains

  function ckname(suffix,number) result(f)
    character(*),intent(in) :: suffix
    integer,intent(in) :: number
    character(len=:),allocatable :: f
    character(32) :: num
    write(num,'(i0)') number
    f = trim(adjustl(pb_name))//suffix//'.ckpt'//trim(num)
  end function ckname


  function itoa(k) result(s)
    integer,intent(in) :: k
    character(len=:),allocatable :: s
    character(32) :: b
    write(b,'(i0)') k
    s = trim(b)
  end function itoa


  subroutine write_number(number)
    ! rewind(C%unit); write(C%unit,*) number (checkpoint.f90:205-208)
    integer,intent(in) :: number
    integer :: u,stat
    open(newunit=u,file=trim(ck%filename),status='replace',action='write',iostat=stat)
    if (stat/=0) then
       call fail('Error opening '//trim(ck%filename))
       return
    end if
    write(u,*) number
    close(u)
  end subroutine write_number


  subroutine init_checkpoint(infile)
    ! init_checkpoint (checkpoint.f90:30-120): the checkpoint_list namelist
    ! and the number file; resume takes the number from the file
    character(*),intent(in) :: infile
    logical :: ckpt,resume,del_final,adjoint,exists
    integer :: begin,interval,u,stat
    real(dp) :: max_time
    namelist /checkpoint_list/ ckpt,resume,begin,interval,max_time,del_final,adjoint

    ckpt = .false.
    resume = .false.
    del_final = .true.
    adjoint = .false.
    begin = 0
    interval = 1
    max_time = 1d40
    open(newunit=u,file=infile,status='old',action='read',iostat=stat)
    if (stat/=0) then
       call fail('Error opening '//trim(infile))
       return
    end if
    read(u,nml=checkpoint_list,iostat=stat)
    close(u)
    if (stat>0) then
       call fail('Error in checkpoint_list')
       return
    end if
    if (adjoint) then
       call fail('checkpoint_list: adjoint checkpoints are not supported')
       return
    end if
    if (interval<1) then
       call fail('checkpoint_list: interval must be positive')
       return
    end if
    ck%ckpt = ckpt
    ck%resume = resume
    ck%del_final = del_final
    ck%begin = begin
    ck%interval = interval
    ck%max_time = max_time
    ck%filename = trim(adjustl(pb_name))//'.ckpt'

    if (ck%resume) then
       inquire(file=trim(ck%filename),exist=exists)
       if (exists) then
          open(newunit=u,file=trim(ck%filename),status='old',action='read',iostat=stat)
          if (stat==0) read(u,*,iostat=stat) ck%begin
          if (stat/=0) then
             call fail('Error reading '//trim(ck%filename))
             return
          end if
          close(u)
          if (ck%begin==-1) then
             call fail('Problem successfully completed, no need to restart. Delete '// &
                  trim(ck%filename)//' to solve problem again')
             return
          end if
       else
          call write_number(0)
       end if
    else
       ! a run that does not resume starts an empty number file
       open(newunit=u,file=trim(ck%filename),status='replace',action='write',iostat=stat)
       if (stat/=0) then
          call fail('Error opening '//trim(ck%filename))
          return
       end if
       close(u)
    end if
    ck%begin = max(ck%begin,0)
    ck%previous = ck%begin
  end subroutine init_checkpoint


  function abort_now(elapsed) result(abort)
    ! abort_now (checkpoint.f90:123-150): <name>.abort holding T, or the
    ! wall-clock limit max_time (minutes) reached
    real(dp),intent(in) :: elapsed
    logical :: abort
    integer :: u,stat
    abort = .false.
    open(newunit=u,file=trim(adjustl(pb_name))//'.abort',status='old',action='read',iostat=stat)
    if (stat==0) then
       read(u,'(l1)',iostat=stat) abort
       if (stat/=0) abort = .false.
       close(u)
    end if
    abort = (abort.or.(elapsed>=ck%max_time*60.0_dp))
  end function abort_now


  logical function any_pmlx()
    ! Wx exists when some block sets pmlx (fields.f90:203-206)
    integer :: i
    any_pmlx = .false.
    do i = 1,nblocks
       any_pmlx = any_pmlx.or.HB(i)%pmlx
    end do
  end function any_pmlx


  logical function any_pmly()
    integer :: i
    any_pmly = .false.
    do i = 1,nblocks
       any_pmly = any_pmly.or.HB(i)%pmly
    end do
  end function any_pmly


  subroutine alloc_w()
    if (any_pmlx().and..not.allocated(ckWx)) allocate(ckWx(gnx,gny,6))
    if (any_pmly().and..not.allocated(ckWy)) allocate(ckWy(gnx,gny,6))
    if (allocated(ckWx)) ckWx = 0.0_dp
    if (allocated(ckWy)) ckWy = 0.0_dp
  end subroutine alloc_w


  logical function block1_plastic()
    ! gammap and lambda exist when the first block is plastic (fields.f90:194)
    block1_plastic = (trim(HB(1)%response)=='plastic')
  end function block1_plastic


  integer function nfric_nodes()
    integer :: k
    nfric_nodes = 0
    do k = 1,nifaces
       if (trim(HI(k)%coupling)=='friction') nfric_nodes = nfric_nodes+HI(k)%n
    end do
  end function nfric_nodes


  subroutine open_ckpt(u,name,ok)
    integer,intent(out) :: u
    character(*),intent(in) :: name
    logical,intent(out) :: ok
    integer :: stat
    open(newunit=u,file=name,access='stream',form='unformatted',status='replace', &
         action='write',iostat=stat)
    ok = (stat==0)
    if (.not.ok) call fail('Error opening '//name)
  end subroutine open_ckpt


  subroutine files_write(number,F,DF,gp,lam,IF)
    ! checkpoint_op('write') (checkpoint.f90:251-290) from host copies of the
    ! state: F, DF [nx,ny,6], gp, lam [nx,ny], IF [nodes of the frictional
    ! interfaces in order, 18]
    integer,intent(in) :: number
    real(dp),intent(in) :: F(:,:,:),DF(:,:,:),gp(:,:),lam(:,:),IF(:,:)
    integer :: u,l,i,s,k,n,off,side,nF,nU
    logical :: ok
    character(1),parameter :: sn(4) = (/ 'L','R','B','T' /)
    real(dp),allocatable :: fill(:),zero(:)

    nF = 6
    nU = 2
    if (pb_mode==3) then
       nF = 3
       nU = 1
    end if
    call open_ckpt(u,ckname('F',number),ok)
    if (.not.ok) return
    do l = 1,nF
       write(u) F(:,:,l)
    end do
    do l = 1,nF
       write(u) DF(:,:,l)
    end do
    if (block1_plastic()) then
       write(u) gp
       write(u) lam
    end if
    ! fields.f90:406-407
    if (any_pmlx()) write(u) ckWx
    if (any_pmly()) write(u) ckWy
    close(u)

    do i = 1,nblocks
       do s = 1,4
          n = size(HB(i)%side(s)%F0,1)
          allocate(fill(n),zero(n))
          fill = unset
          zero = 0.0_dp
          call open_ckpt(u,ckname('block'//itoa(i)//sn(s),number),ok)
          if (.not.ok) return
          do l = 1,nF
             write(u) fill
             write(u) fill
             write(u) HB(i)%side(s)%F0(:,l)
          end do
          do l = 1,nU
             write(u) zero
             write(u) fill
          end do
          write(u) zero
          write(u) fill
          close(u)
          deallocate(fill,zero)
       end do
    end do

    off = 0
    do k = 1,nifaces
       do side = 1,2
          call open_ckpt(u,ckname('iface'//itoa(k)//'side'//itoa(side),number),ok)
          if (.not.ok) return
          if (trim(HI(k)%coupling)=='friction') then
             n = HI(k)%n
             do l = 1,FDM_IF_NFIELDS
                write(u) IF(off+1:off+n,l)
             end do
          end if
          close(u)
       end do
       if (trim(HI(k)%coupling)=='friction') off = off+HI(k)%n
    end do
  end subroutine files_write


  subroutine files_read(number,F,DF,gp,lam,IF)
    ! checkpoint_op('read'): the field file and side 1 of each frictional
    ! interface
    integer,intent(in) :: number
    real(dp),intent(out) :: F(:,:,:),DF(:,:,:),gp(:,:),lam(:,:),IF(:,:)
    integer :: u,l,k,n,off,stat,nF
    character(len=:),allocatable :: name

    nF = 6
    if (pb_mode==3) nF = 3
    gp = 0.0_dp
    lam = 0.0_dp
    name = ckname('F',number)
    open(newunit=u,file=name,access='stream',form='unformatted',status='old', &
         action='read',iostat=stat)
    if (stat/=0) then
       call fail('Error opening '//name)
       return
    end if
    do l = 1,nF
       read(u,iostat=stat) F(:,:,l)
       if (stat/=0) exit
    end do
    do l = 1,nF
       if (stat/=0) exit
       read(u,iostat=stat) DF(:,:,l)
    end do
    if (stat==0.and.block1_plastic()) then
       read(u,iostat=stat) gp
       if (stat==0) read(u,iostat=stat) lam
    end if
    ! fields.f90:383-384
    call alloc_w()
    if (stat==0.and.any_pmlx()) read(u,iostat=stat) ckWx
    if (stat==0.and.any_pmly()) read(u,iostat=stat) ckWy
    close(u)
    if (stat/=0) then
       call fail('Error reading '//name)
       return
    end if

    off = 0
    do k = 1,nifaces
       if (trim(HI(k)%coupling)/='friction') cycle
       n = HI(k)%n
       name = ckname('iface'//itoa(k)//'side1',number)
       open(newunit=u,file=name,access='stream',form='unformatted',status='old', &
            action='read',iostat=stat)
       if (stat/=0) then
          call fail('Error opening '//name)
          return
       end if
       do l = 1,FDM_IF_NFIELDS
          read(u,iostat=stat) IF(off+1:off+n,l)
          if (stat/=0) exit
       end do
       close(u)
       if (stat/=0) then
          call fail('Error reading '//name)
          return
       end if
       off = off+n
    end do
  end subroutine files_read


  subroutine delete_file(name)
    character(*),intent(in) :: name
    integer :: u,stat
    open(newunit=u,file=name,status='old',iostat=stat)
    if (stat==0) close(u,status='delete')
  end subroutine delete_file


  subroutine files_delete(number)
    ! checkpoint_op('delete'): every file of checkpoint number, including
    ! the per-block E file the reference also removes
    integer,intent(in) :: number
    integer :: i,s,k
    character(1),parameter :: sn(4) = (/ 'L','R','B','T' /)
    call delete_file(ckname('F',number))
    do i = 1,nblocks
       do s = 1,4
          call delete_file(ckname('block'//itoa(i)//sn(s),number))
       end do
       call delete_file(ckname('Eblock'//itoa(i),number))
    end do
    do k = 1,nifaces
       call delete_file(ckname('iface'//itoa(k)//'side1',number))
       call delete_file(ckname('iface'//itoa(k)//'side2',number))
    end do
  end subroutine files_delete


  subroutine gather_state(ctx,F,DF,gp,lam,IF,ok)
    type(c_ptr),intent(in) :: ctx
    real(dp),intent(inout),target :: F(:,:,:),DF(:,:,:),gp(:,:),lam(:,:),IF(:,:)
    logical,intent(out) :: ok
    integer :: k,l,off,stat
    real(dp),allocatable,target :: buf(:)
    stat = fdm_get_fields(ctx,c_loc(F(1,1,1)),c_loc(DF(1,1,1)))
    if (stat==FDM_OK.and.block1_plastic()) stat = fdm_get_plastic(ctx,c_loc(gp(1,1)),c_loc(lam(1,1)))
    call alloc_w()
    if (stat==FDM_OK.and.any_pmlx()) stat = fdm_get_pml(ctx,c_loc(ckWx(1,1,1)),c_null_ptr)
    if (stat==FDM_OK.and.any_pmly()) stat = fdm_get_pml(ctx,c_null_ptr,c_loc(ckWy(1,1,1)))
    off = 0
    do k = 1,nifaces
       if (stat/=FDM_OK) exit
       if (trim(HI(k)%coupling)/='friction') cycle
       allocate(buf(HI(k)%n))
       do l = 1,FDM_IF_NFIELDS
          stat = fdm_get_iface(ctx,int(k-1,c_int32_t),int(l-1,c_int32_t),c_loc(buf(1)))
          if (stat/=FDM_OK) exit
          IF(off+1:off+HI(k)%n,l) = buf
       end do
       off = off+HI(k)%n
       deallocate(buf)
    end do
    ok = (stat==FDM_OK)
  end subroutine gather_state


  subroutine restore_state(ctx,F,DF,gp,lam,IF,ok)
    type(c_ptr),intent(in) :: ctx
    real(dp),intent(in),target :: F(:,:,:),DF(:,:,:),gp(:,:),lam(:,:),IF(:,:)
    logical,intent(out) :: ok
    integer :: k,l,off,stat
    real(dp),allocatable,target :: buf(:)
    stat = fdm_set_fields(ctx,c_loc(F(1,1,1)),c_loc(DF(1,1,1)))
    if (stat==FDM_OK.and.block1_plastic()) stat = fdm_set_plastic(ctx,c_loc(gp(1,1)),c_loc(lam(1,1)))
    if (stat==FDM_OK.and.any_pmlx()) stat = fdm_set_pml(ctx,c_loc(ckWx(1,1,1)),c_null_ptr)
    if (stat==FDM_OK.and.any_pmly()) stat = fdm_set_pml(ctx,c_null_ptr,c_loc(ckWy(1,1,1)))
    off = 0
    do k = 1,nifaces
       if (stat/=FDM_OK) exit
       if (trim(HI(k)%coupling)/='friction') cycle
       allocate(buf(HI(k)%n))
       do l = 1,FDM_IF_NFIELDS
          buf = IF(off+1:off+HI(k)%n,l)
          stat = fdm_set_iface(ctx,int(k-1,c_int32_t),int(l-1,c_int32_t),c_loc(buf(1)))
          if (stat/=FDM_OK) exit
       end do
       off = off+HI(k)%n
       deallocate(buf)
    end do
    ok = (stat==FDM_OK)
  end subroutine restore_state


  subroutine ckpt_read(ctx,ok)
    ! checkpoint_read (checkpoint.f90:153-178)
    type(c_ptr),intent(in) :: ctx
    logical,intent(out) :: ok
    real(dp),allocatable,target :: F(:,:,:),DF(:,:,:),gp(:,:),lam(:,:),IF(:,:)
    ok = .true.
    if (.not.ck%ckpt) return
    if (ck%begin==0) return
    write(6,'(a,i0)') '...reading  checkpoint at time step ',ck%begin
    allocate(F(gnx,gny,6),DF(gnx,gny,6),gp(gnx,gny),lam(gnx,gny), &
         IF(max(1,nfric_nodes()),FDM_IF_NFIELDS))
    call files_read(ck%begin,F,DF,gp,lam,IF)
    ok = (setup_error=='')
    if (ok) call restore_state(ctx,F,DF,gp,lam,IF,ok)
  end subroutine ckpt_read


  subroutine ckpt_write_delete(ctx,number,abort,ok)
    ! checkpoint_write_delete (checkpoint.f90:181-213): a checkpoint every
    ! interval steps, or now when aborting; then the previous one goes
    type(c_ptr),intent(in) :: ctx
    integer,intent(in) :: number
    logical,intent(in) :: abort
    logical,intent(out) :: ok
    real(dp),allocatable,target :: F(:,:,:),DF(:,:,:),gp(:,:),lam(:,:),IF(:,:)
    ok = .true.
    if (abort) then
       ck%ckpt = .true.
    else
       if (mod(number,ck%interval)/=0) return
    end if
    if (.not.ck%ckpt) return
    write(6,'(a,i0)') '...writing  checkpoint at time step ',number
    allocate(F(gnx,gny,6),DF(gnx,gny,6),gp(gnx,gny),lam(gnx,gny), &
         IF(max(1,nfric_nodes()),FDM_IF_NFIELDS))
    gp = 0.0_dp
    lam = 0.0_dp
    call gather_state(ctx,F,DF,gp,lam,IF,ok)
    if (.not.ok) return
    call files_write(number,F,DF,gp,lam,IF)
    if (setup_error/='') then
       ok = .false.
       return
    end if
    call write_number(number)
    call ckpt_delete(.false.)
    ck%previous = number
  end subroutine ckpt_write_delete


  subroutine ckpt_delete(final)
    ! checkpoint_delete (checkpoint.f90:216-248): remove the previous
    ! checkpoint; at the end of a finished run the number file reads -1
    logical,intent(in) :: final
    if (.not.ck%ckpt) return
    if (ck%previous==0) return
    write(6,'(a,i0)') '...deleting checkpoint at time step ',ck%previous
    call files_delete(ck%previous)
    if (final) call write_number(-1)
  end subroutine ckpt_delete


  ! ---------------------------------------------------------------
  ! C entry points over the files alone (no device), for tests
  ! ---------------------------------------------------------------

  function fdmap_host_ckpt_init(path) bind(C,name='fdmap_host_ckpt_init') result(begin)
    ! after fdmap_host_setup: parse checkpoint_list; returns the step the
    ! run starts after, or -1 with fdmap_host_error() set
    character(kind=c_char),intent(in) :: path(*)
    integer(c_int) :: begin
    character(len=:),allocatable :: f
    integer :: n
    n = 0
    do while (path(n+1)/=c_null_char)
       n = n+1
    end do
    allocate(character(len=n) :: f)
    f = transfer(path(1:n),f)
    setup_error = ''
    ck = ckpt_type()
    call init_checkpoint(f)
    begin = ck%begin
    if (setup_error/='') begin = -1
  end function fdmap_host_ckpt_init


  function fdmap_host_ckpt_files(op,number,F,DF,gp,lam,IF,Wx,Wy) bind(C,name='fdmap_host_ckpt_files') &
       result(rc)
    ! op 0: write, 1: read, 2: delete checkpoint number; F, DF [6][ny][nx],
    ! gp, lam [ny][nx], IF [18][frictional interface nodes], Wx, Wy
    ! [6][ny][nx] (used when a block sets pmlx / pmly)
    integer(c_int32_t),value :: op,number
    type(c_ptr),value :: F,DF,gp,lam,IF,Wx,Wy
    integer(c_int) :: rc
    real(dp),pointer :: pF(:,:,:),pDF(:,:,:),pgp(:,:),plam(:,:),pIF(:,:),pW(:,:,:)
    setup_error = ''
    if (op==2) then
       call files_delete(number)
       rc = 0
       return
    end if
    call c_f_pointer(F,pF,(/ gnx,gny,6 /))
    call c_f_pointer(DF,pDF,(/ gnx,gny,6 /))
    call c_f_pointer(gp,pgp,(/ gnx,gny /))
    call c_f_pointer(lam,plam,(/ gnx,gny /))
    call c_f_pointer(IF,pIF,(/ max(1,nfric_nodes()),FDM_IF_NFIELDS /))
    if (op==0) then
       call alloc_w()
       if (any_pmlx()) then
          call c_f_pointer(Wx,pW,(/ gnx,gny,6 /))
          ckWx = pW
       end if
       if (any_pmly()) then
          call c_f_pointer(Wy,pW,(/ gnx,gny,6 /))
          ckWy = pW
       end if
       call files_write(number,pF,pDF,pgp,plam,pIF)
    else
       call files_read(number,pF,pDF,pgp,plam,pIF)
       if (any_pmlx()) then
          call c_f_pointer(Wx,pW,(/ gnx,gny,6 /))
          pW = ckWx
       end if
       if (any_pmly()) then
          call c_f_pointer(Wy,pW,(/ gnx,gny,6 /))
          pW = ckWy
       end if
    end if
    rc = 0
    if (setup_error/='') rc = -1
  end function fdmap_host_ckpt_files

end module fdmap_checkpoint
