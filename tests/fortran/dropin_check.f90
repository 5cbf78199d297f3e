! Drop-in check (test infrastructure): the reference BSAM host code -- its
! tree (TreeOps), its storage (AllocFields) and its solver modules -- drives
! the MI355X engine through the Fortran binding (bsam_amd/fortran).
!
!   dropin_check host SPEC INPUT   engine ghost programs (host-only context)
!                                  applied to the reference's own q arrays vs
!                                  the reference SetGhost, every level/pass
!   dropin_check gpu  SPEC INPUT   GpuMultigridIterations vs the reference
!                                  MultigridIterations, NSTEP time steps
!
! Prints "max_abs_diff <value>" (0 expected) and "cycles ...".
module DropinBuffers
  implicit none
  real, allocatable, save :: buf(:)
  integer, save :: pos = 0
  real, save :: worst = 0.0
  integer, save :: mode = 0      ! 0 save q, 1 compare q, 2 save qold, 3 restore q, 4 restore qold
end module DropinBuffers

module DropinCallbacks
contains
  integer function BufPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use DropinBuffers
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    integer :: n
    n = size(info%q)
    select case (mode)
    case (0)
      buf(pos+1:pos+n) = reshape(info%q, [n])
    case (1)
      worst = max(worst, maxval(abs(buf(pos+1:pos+n) - reshape(info%q, [n]))))
    case (2)
      buf(pos+1:pos+n) = reshape(info%qold, [n])
    case (3)
      info%q = reshape(buf(pos+1:pos+n), shape(info%q))
    case (4)
      info%qold = reshape(buf(pos+1:pos+n), shape(info%qold))
    end select
    pos = pos + n
    BufPatch = err_ok
  end function BufPatch

  integer function SizePatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use DropinBuffers
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    pos = pos + size(info%q)
    SizePatch = err_ok
  end function SizePatch
end module DropinCallbacks

program DropinCheck
  use iso_c_binding
  use NodeInfoDef
  use TreeOps
  use Boundary, only: SetGhost
  use AFASRoutines, only: FillDown, MultigridIterations
  use HarnessState
  use HarnessCallbacks
  use HarnessTree
  use AFASGPU
  use DropinBuffers
  use DropinCallbacks
  implicit none
  character(len=512) :: what, specname, inname, arg
  type(funcparam) :: dummy
  integer :: maxv_user, lev, ip, nstep, s, total

  call get_command_argument(1, what)
  call get_command_argument(2, specname)
  call get_command_argument(3, inname)
  nstep = 2
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) nstep
  end if
  open(10, file=trim(specname), status='old', action='read')
  call build_tree(10, maxv_user)
  close(10)
  errortype = 1
  updateauxfreq = 1
  maxvcycles = maxv_user

  pos = 0
  do lev = minlevel, finestlevel
    call ApplyOnLevel(lev, SizePatch, dummy)
  end do
  total = pos
  allocate(buf(total))

  if (trim(what) == 'host') then
    call GpuOpen(-1, 0)
    call GpuDescribeHierarchy()
    do lev = minlevel, finestlevel
      do ip = 0, 2
        call load_inputs()
        call SetGhost(lev, ip)
        call walk_buf(0)
        call load_inputs()
        call GpuApplyGhostHost(lev, ip)
        call walk_buf(1)
      end do
    end do
    print '(a,es12.4)', 'max_abs_diff ', worst
    print '(a,i0)', 'patches ', gpu_npatch
  else
    call GpuOpen(0, 0)
    call GpuDescribeHierarchy()
    call load_inputs()
    do s = 1, nstep
      ! one time step as TakeTimeSteps prepares it (lib/bsamroutines.f90:648-661)
      do lev = rootlevel, finestlevel
        call SetGhost(lev, 0)
      end do
      call FillDown(solutionfield)
      call copy_qold()
      call walk_buf(0)          ! state before the solve
      call MultigridIterations  ! reference, on the host
      call swap_with_buf()      ! buf <- reference result, q <- state before
      call GpuMultigridIterations()
      call walk_buf(1)
      print '(a,i0,a,i0,a,i0)', 'step ', s, ' gpu_cycles ', gpu_cycles, ' converged ', &
        gpu_converged
    end do
    print '(a,es12.4)', 'max_abs_diff ', worst
  end if
  call GpuClose()

contains

  subroutine load_inputs()
    integer :: l
    h_unit = 11
    open(h_unit, file=trim(inname), access='stream', form='unformatted', &
         status='old', action='read')
    do l = minlevel, finestlevel
      call ApplyOnLevel(l, LoadPatch, dummy)
    end do
    close(h_unit)
  end subroutine load_inputs

  subroutine walk_buf(m)
    integer, intent(in) :: m
    integer :: l
    mode = m
    pos = 0
    do l = minlevel, finestlevel
      call ApplyOnLevel(l, BufPatch, dummy)
    end do
  end subroutine walk_buf

  subroutine copy_qold()
    integer :: l
    do l = minlevel, finestlevel
      call ApplyOnLevel(l, CopyToOld, dummy)
    end do
  end subroutine copy_qold

  subroutine swap_with_buf()
    real, allocatable :: tmp(:)
    allocate(tmp(total))
    tmp = buf
    call walk_buf(0)         ! buf <- reference result
    call swap_restore(tmp)
  end subroutine swap_with_buf

  subroutine swap_restore(tmp)
    real, intent(in) :: tmp(:)
    real, allocatable :: keep(:)
    allocate(keep(total))
    keep = buf
    buf = tmp
    call walk_buf(3)         ! q <- state before the solve
    buf = keep
  end subroutine swap_restore

end program DropinCheck
