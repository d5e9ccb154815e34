! fdmap -- the Fortran driver executable: `fdmap problem.in [fp32-bulk]`.
! Same command line as the reference (main.f90:34-35, default 'default.in');
! an optional second argument selects the FP32-bulk precision mode.
program fdmap
  use iso_c_binding
  use fdmap_driver, only : run_problem
  use fdmap_capi, only : FDM_FP64, FDM_FP32_BULK
  implicit none
  character(256) :: input_file,prec
  integer :: stat,rc
  integer(c_int32_t) :: precision
  call get_command_argument(1,input_file,status=stat)
  if (stat/=0) input_file = 'default.in'
  precision = FDM_FP64
  call get_command_argument(2,prec,status=stat)
  if (stat==0) then
     if (trim(prec)=='fp32-bulk') precision = FDM_FP32_BULK
  end if
  rc = run_problem(trim(input_file),precision)
  if (rc/=0) stop 1
end program fdmap
