! Drop-in check (test infrastructure): the reference BSAM host code -- its
! tree (TreeOps), its storage (AllocFields) and its solver modules -- drives
! the MI355X engine through the Fortran binding (bsam_amd/fortran).
!
!   dropin_check host SPEC INPUT   engine ghost programs (host-only context)
!                                  applied to the reference's own q arrays vs
!                                  the reference SetGhost, every level/pass
!   dropin_check gpu  SPEC INPUT   GpuMultigridIterations vs the reference
!                                  MultigridIterations, NSTEP time steps
!   dropin_check trunc SPEC INPUT  GpuRelativeTruncationError vs the
!                                  reference calls of EstimateLevelErrors
!                                  (q, qold, qc, qrte), GpuSetErrFlags'
!                                  flags vs the SetErrFlags rule on that
!                                  qrte, and GpuSetErrFlagsUser vs the
!                                  reference SetErrFlagsUser2D/3D, every
!                                  level from the root
!
! Prints "max_abs_diff <value>" (0 expected) and "cycles ...".
module DropinBuffers
  implicit none
  real, allocatable, save :: buf(:)
  integer, save :: pos = 0
  real, save :: worst = 0.0
  integer, save :: mode = 0      ! 0 save q, 1 compare q, 2 save qold, 3 restore q, 4 restore qold
                                 ! 5 save, 6 compare: q, qold (+ qc, qrte on fulllevel);
                                 ! 7 poisons the same arrays but q (the input)
  integer, save :: fulllevel = -1
  real, save :: tol = 0.0, scale = 0.0, hlev = 0.0
  integer, save :: nflagged = 0, nbadflags = 0
  integer, allocatable, save :: ibuf(:)
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
    BufPatch = err_ok
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
    case (5, 6, 7)
      if (mode /= 7) call arr(info%q)
      call arr(info%qold)
      if (info%level == fulllevel) then
        call arr(info%qc)
        call arr(info%qrte)
      end if
      return
    end select
    pos = pos + n
  contains
    subroutine arr(a)
      real, intent(inout) :: a(:,:,:,:)
      integer :: m
      m = size(a)
      select case (mode)
      case (5)
        buf(pos+1:pos+m) = reshape(a, [m])
      case (6)
        worst = max(worst, maxval(abs(buf(pos+1:pos+m) - reshape(a, [m]))))
      case (7)
        a = 1.0e300
      end select
      pos = pos + m
    end subroutine arr
  end function BufPatch

  ! Largest |qrte| of the patches of fulllevel (mode 0), the check of their
  ! error flags against SetErrFlags2D/3D (lib/bsamroutines.f90:1093-1157) on
  ! their qrte at tolerance tol (mode 1), or their flags saved to (mode 2)
  ! and compared with (mode 3) ibuf.
  integer function FlagPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use DropinBuffers
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    integer :: i, j, k, c(3), want, nd
    FlagPatch = err_ok
    if (info%level /= fulllevel) return
    nd = ndims
    c = 1
    c(1:nd) = info%mx(1:nd) / 2
    if (mode == 0) then
      scale = max(scale, maxval(abs(info%qrte(1:c(1), 1:c(2), 1:c(3), :))))
      hlev = info%dx(1)
      return
    end if
    if (mode == 2 .or. mode == 3) then
      ! save (2) or compare (3) the patch's flags
      i = size(info%errorflags)
      if (mode == 2) then
        ibuf(pos+1:pos+i) = reshape(info%errorflags, [i])
      else
        nbadflags = nbadflags + count(ibuf(pos+1:pos+i) /= reshape(info%errorflags, [i]))
        nflagged = nflagged + sum(info%errorflags)
      end if
      pos = pos + i
      return
    end if
    do k = 1, info%mx(3)
      do j = 1, info%mx(2)
        do i = 1, info%mx(1)
          want = 0
          if ((i + 1) / 2 <= c(1) .and. (j + 1) / 2 <= c(2) .and. &
              (nd == 2 .or. (k + 1) / 2 <= c(3))) then
            if (maxval(abs(info%qrte((i + 1) / 2, (j + 1) / 2, &
                                     merge(1, (k + 1) / 2, nd == 2), :))) > tol) want = 1
          end if
          if (info%errorflags(i, j, k) /= want) nbadflags = nbadflags + 1
          nflagged = nflagged + info%errorflags(i, j, k)
        end do
      end do
    end do
  end function FlagPatch

  integer function ZeroFlags(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    info%errorflags = 7
    ZeroFlags = err_ok
  end function ZeroFlags

  integer function SizePatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use DropinBuffers
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    pos = pos + size(info%q) + size(info%qold) + size(info%qc) + size(info%qrte)
    SizePatch = err_ok
  end function SizePatch
end module DropinCallbacks

program DropinCheck
  use iso_c_binding
  use NodeInfoDef
  use TreeOps
  use Boundary, only: SetGhost, GetCoarseGhostPoints
  use AFASRoutines, only: FillDown, MultigridIterations, RestrictSolution, &
                          RelativeTruncationError
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
  allocate(buf(total), ibuf(total))

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
  else if (trim(what) == 'trunc') then
    call GpuOpen(0, 0)
    call GpuDescribeHierarchy()
    do lev = max(minlevel + 1, 0), finestlevel
      fulllevel = lev
      ! the truncation-error calls of EstimateLevelErrors
      ! (lib/bsamroutines.f90:971-981), on the host
      call load_inputs()
      call ApplyOnLevel(lev, RestrictSolution, dummy)
      call SetGhost(lev - 1, 0)
      call ApplyOnLevel(lev, GetCoarseGhostPoints, dummy)
      call ApplyOnLevel(lev, CopyToOld, dummy)
      call ApplyOnLevel(lev - 1, CopyToOld, dummy)
      call ApplyOnLevel(lev, RelativeTruncationError, dummy)
      call walk_levels(5, lev - 1, lev)   ! buf <- reference arrays
      call load_inputs()
      call walk_levels(7, lev - 1, lev)   ! poison what the engine must write
      call GpuRelativeTruncationError(lev)
      call walk_levels(6, lev - 1, lev)
      ! default error tagging at a tolerance inside the data's range
      mode = 0
      scale = 0.0
      call ApplyOnLevel(lev, FlagPatch, dummy)
      qtolerance(lev) = 0.3 * scale * hlev**ndims
      tol = qtolerance(lev) / hlev / hlev
      if (ndims == 3) tol = tol / hlev
      call GpuSetErrFlags(lev)
      mode = 1
      call ApplyOnLevel(lev, FlagPatch, dummy)
      ! the problem's user tagging after SetGhost(lev, 0), as AMR runs it:
      ! reference SetErrFlagsUser2D/3D on the host vs GpuSetErrFlagsUser
      call load_inputs()
      call SetGhost(lev, 0)
      qtolerance(lev) = 0.05
      call ApplyOnLevel(lev, UserFlagPatch, dummy)
      mode = 2
      pos = 0
      call ApplyOnLevel(lev, FlagPatch, dummy)
      call ApplyOnLevel(lev, ZeroFlags, dummy)
      call GpuSetErrFlagsUser(lev)
      mode = 3
      pos = 0
      call ApplyOnLevel(lev, FlagPatch, dummy)
    end do
    if (nbadflags > 0) worst = max(worst, 1.0)
    print '(a,i0,a,i0)', 'flagged ', nflagged, ' bad_flags ', nbadflags
    print '(a,es12.4)', 'max_abs_diff ', worst
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

  subroutine walk_levels(m, l0, l1)
    integer, intent(in) :: m, l0, l1
    integer :: l
    mode = m
    pos = 0
    do l = l0, l1
      call ApplyOnLevel(l, BufPatch, dummy)
    end do
  end subroutine walk_levels

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
