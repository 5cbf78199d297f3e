! afas_gpu_wrap.f90 -- the reference program with its solve on the MI355X
! engine, without a change to the reference sources.
!
! Linked into the reference program (its objects from lib/ and the problem
! module, plus bsam_gpu.o, afas_gpu.o and this file) with
!
!   -Wl,--wrap=_QMafasroutinesPmultigriditerations
!
! the linker sends every call of AFASRoutines' MultigridIterations
! (lib/afasroutines.f90:38-106) made from another object file -- the one in
! TakeTimeSteps, lib/bsamroutines.f90:714-718 -- to the routine below. It
! opens the engine on the first call (SetProb has run by then), hands it the
! hierarchy the regrid just built (the reference regrids every step,
! lib/bsamroutines.f90:649) and runs the solve on the GPU; the host arrays
! q afterwards hold what MultigridIterations would have left, so the rest
! of the time step (statistics, frames, the next regrid) runs unchanged.
!
! Environment: BSAM_DEVICE (HIP device, default 0), BSAM_PRECISION (0 FP64,
! 1 mixed; default 0), BSAM_TIMING (1: print the wall time of the hierarchy
! hand-over and of the solve with its copies, per call).
subroutine WrapMultigridIterations() bind(C, name='__wrap__QMafasroutinesPmultigriditerations')
  use iso_c_binding
  use bsam_gpu, only: bsam_tune_get
  use AFASGPU
  implicit none
  logical, save :: opened = .false.
  integer, save :: timing = 0
  integer :: dev, prec, n
  integer(8) :: t0, t1, t2, rate
  real(c_double) :: reused
  integer(c_int) :: rc
  if (.not. opened) then
    dev = env_int('BSAM_DEVICE', 0)
    prec = env_int('BSAM_PRECISION', 0)
    timing = env_int('BSAM_TIMING', 0)
    call GpuOpen(dev, prec)
    opened = .true.
  end if
  call system_clock(t0, rate)
  call GpuDescribeHierarchy()
  call system_clock(t1)
  call GpuMultigridIterations()
  call system_clock(t2)
  if (timing /= 0) then
    rc = bsam_tune_get(gpu_ctx, 'commits_reused'//c_null_char, reused)
    print '(a,f10.3,a,f10.3,a,i0,a,i0)', ' gpu describe_ms ', &
      1.0d3 * real(t1 - t0, 8) / real(rate, 8), ' solve_ms ', &
      1.0d3 * real(t2 - t1, 8) / real(rate, 8), ' cycles ', gpu_cycles, &
      ' reused ', int(reused)
  end if
  ! the per-cycle lines MultigridIterations prints (afasroutines.f90:96)
  do n = 1, min(gpu_cycles, size(gpu_residuals))
    print *, 'it =', n, ' residual =', gpu_residuals(n)
  end do
contains
  integer function env_int(name, default)
    character(len=*), intent(in) :: name
    integer, intent(in) :: default
    character(len=32) :: v
    integer :: st, ios
    env_int = default
    call get_environment_variable(name, v, status=st)
    if (st /= 0) return
    read(v, *, iostat=ios) env_int
    if (ios /= 0) env_int = default
  end function env_int
end subroutine WrapMultigridIterations
