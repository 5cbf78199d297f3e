! afas_gpu.f90 -- drop-in glue between the reference BSAM host code and the
! MI355X FAS engine.
!
! Compiled together with the reference modules NodeInfoDef, TreeOps and
! ProblemDef, it replaces the solve of AFASRoutines (lib/afasroutines.f90):
!
!   call GpuOpen(device, precision)      ! once, after SetProb
!   call GpuDescribeHierarchy()          ! after every regrid (AMR)
!   call GpuMultigridIterations()        ! in place of MultigridIterations
!
! and, in EstimateLevelErrors (lib/bsamroutines.f90:963-1002), the relative
! truncation error and the default error tagging of a refined level:
!
!   call GpuRelativeTruncationError(level) ! RestrictSolution .. RelativeTruncationError
!   call GpuSetErrFlags(level)             ! SetErrFlags2D/3D on every patch
!   call GpuSetErrFlagsUser(level)         ! or the problem's SetErrFlagsUser2D/3D
!
! and TakeTimeSteps' after-step statistics (mass, 2D energy):
!
!   call GpuAfterStepStatistics()          ! integralresult(1:2)
!
! The tree stays on the host. GpuDescribeHierarchy walks it with
! ApplyOnLevel (levels minlevel..finestlevel, youngest patch first) and hands
! every nodeinfo's geometry to the engine; GpuMultigridIterations uploads the
! patch arrays q and qold (the reference's own allocations, passed by
! address), runs the solve on the GPU and copies q back, so the host state
! afterwards is what MultigridIterations would have left in q.
module AFASGPU
  use iso_c_binding
  use bsam_gpu
  implicit none
  private
  public :: GpuOpen, GpuClose, GpuDescribeHierarchy, GpuMultigridIterations, &
            GpuSetGhost, GpuUploadAll, GpuDownloadAll, GpuApplyGhostHost, &
            GpuRelativeTruncationError, GpuSetErrFlags, GpuSetErrFlagsUser, &
            GpuAfterStepStatistics, gpu_ctx, &
            gpu_npatch, gpu_cycles, gpu_converged, gpu_residuals

  type(c_ptr), save :: gpu_ctx = c_null_ptr
  integer, save :: gpu_npatch = 0, gpu_cycles = 0, gpu_converged = 0
  real(c_double), save :: gpu_residuals(200)
  ! q-array address of each patch in engine order, to find parent indices
  integer(c_intptr_t), allocatable, save :: qaddr(:)
  type(c_ptr), allocatable, save :: qptr(:)
  integer, save :: walk_field = 0, walk_dir = 0, walk_level = -1

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      print *, 'AFASGPU: ', what, ' failed: ', trim(bsam_error_message(gpu_ctx))
      stop 1
    end if
  end subroutine check

  ! Creates the engine context from the run globals (NodeInfoDef) and the
  ! problem constants (ProblemDef, problem.f90:781-813: Dfc = alpha (c-1/2)^3,
  ! Dfe = alpha (c-1/2)/4).
  subroutine GpuOpen(device, precision)
    use NodeInfoDef
    use ProblemDef
    integer, intent(in) :: device, precision
    type(bsam_cfg) :: cfg
    real(c_double), dimension(2) :: qpo2
    cfg%ndims = ndims
    cfg%nrvars = 2
    cfg%mbc = 1
    cfg%minlevel = minlevel
    cfg%dt = dt
    cfg%eps2 = eps2
    cfg%fc_alpha = alpha
    cfg%fc_shift = 0.5_c_double
    cfg%fe_num = alpha
    cfg%fe_den = 4.0_c_double
    cfg%nsmoothingpasses = nsmoothingpasses
    cfg%maxvcycles = maxvcycles
    cfg%qerrortol = qerrortol
    cfg%precision = precision
    cfg%device = device
    cfg%errortype = errortype
    gpu_ctx = bsam_ctx_create(cfg)
    if (.not. c_associated(gpu_ctx)) then
      print *, 'AFASGPU: cannot create the engine context: ', &
               trim(bsam_error_message(c_null_ptr))
      stop 1
    end if
    ! PeriodicBC's and PeriodicGhostOverlap's quasi-periodic offsets
    ! (boundary.f90:432-461, :578-591)
    qpo2 = real(qpo(1:2), c_double)
    call check(bsam_set_qpo(gpu_ctx, qpo2, 2_c_int), 'bsam_set_qpo')
  end subroutine GpuOpen

  ! The module's dt as it is now: TakeTimeSteps sets it to 0 for the first
  ! pass of a syncelliptic run and back after (lib/bsamroutines.f90:620-622,
  ! :748-750). The engine returns at once when it is unchanged.
  subroutine GpuSetDt()
    use NodeInfoDef, only: dt
    call check(bsam_set_dt(gpu_ctx, real(dt, c_double)), 'bsam_set_dt')
  end subroutine GpuSetDt

  subroutine GpuClose()
    integer(c_int) :: rc
    if (c_associated(gpu_ctx)) rc = bsam_ctx_destroy(gpu_ctx)
    gpu_ctx = c_null_ptr
  end subroutine GpuClose

  integer function CountPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    type(nodeinfo) :: info
    type(funcparam) :: dummy
    gpu_npatch = gpu_npatch + 1
    CountPatch = err_ok
  end function CountPatch

  integer function AddPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok, GetParentInfo
    type(nodeinfo) :: info
    type(funcparam) :: dummy
    type(nodeinfo), pointer :: parent
    integer :: ierror, n, pidx
    integer(c_int) :: mx(3), mb(6), mg(6), bc(6)
    real(c_double) :: xl(3)
    integer(c_intptr_t) :: pa
    gpu_npatch = gpu_npatch + 1
    qptr(gpu_npatch) = c_loc(info%q)
    qaddr(gpu_npatch) = transfer(qptr(gpu_npatch), 0_c_intptr_t)
    pidx = -1
    if (info%level > minlevel) then
      ierror = GetParentInfo(parent)
      pa = transfer(c_loc(parent%q), 0_c_intptr_t)
      do n = 1, gpu_npatch - 1
        if (qaddr(n) == pa) then
          pidx = n - 1
          exit
        end if
      end do
    end if
    do n = 1, 3
      mx(n) = info%mx(n)
      mb(2*n-1) = info%mbounds(n,1)
      mb(2*n) = info%mbounds(n,2)
      mg(2*n-1) = info%mglobal(n,1)
      mg(2*n) = info%mglobal(n,2)
      xl(n) = info%xlower(n)
    end do
    bc = info%mthbc
    call check(bsam_add_patch(gpu_ctx, info%ngrid, info%level, pidx, mx, mb, mg, bc, &
                              info%dx(1), xl), 'bsam_add_patch')
    AddPatch = err_ok
  end function AddPatch

  ! Hands the current tree to the engine (after InitForest/regrid).
  subroutine GpuDescribeHierarchy()
    use NodeInfoDef
    use TreeOps, only: ApplyOnLevel
    type(funcparam) :: dummy
    integer :: level
    gpu_npatch = 0
    do level = minlevel, finestlevel
      call ApplyOnLevel(level, CountPatch, dummy)
    end do
    if (allocated(qaddr)) deallocate(qaddr, qptr)
    allocate(qaddr(gpu_npatch), qptr(gpu_npatch))
    call check(bsam_hierarchy_begin(gpu_ctx), 'bsam_hierarchy_begin')
    gpu_npatch = 0
    do level = minlevel, finestlevel
      call ApplyOnLevel(level, AddPatch, dummy)
    end do
    call check(bsam_hierarchy_commit(gpu_ctx), 'bsam_hierarchy_commit')
  end subroutine GpuDescribeHierarchy

  ! Copies one patch array between the host nodeinfo and the engine:
  ! walk_dir 1 uploads, 0 downloads; walk_field names the array (a field id,
  ! or 0 for the error flags); walk_level >= 0 restricts the copy to one
  ! level (the walk still numbers every patch).
  integer function MovePatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    type(nodeinfo) :: info
    type(funcparam) :: dummy
    integer(c_int) :: f, p
    type(c_ptr) :: a
    integer(c_long) :: n
    gpu_npatch = gpu_npatch + 1
    MovePatch = err_ok
    if (walk_level >= 0 .and. info%level /= walk_level) return
    f = walk_field
    p = gpu_npatch - 1
    select case (f)
    case (BSAM_FIELD_Q)
      a = c_loc(info%q)
      n = size(info%q)
    case (BSAM_FIELD_QOLD)
      a = c_loc(info%qold)
      n = size(info%qold)
    case (BSAM_FIELD_QC)
      a = c_loc(info%qc)
      n = size(info%qc)
    case (BSAM_FIELD_QRTE)
      a = c_loc(info%qrte)
      n = size(info%qrte)
    case default
      call check(bsam_download_flags(gpu_ctx, p, c_loc(info%errorflags), &
                                     int(size(info%errorflags), c_long)), 'bsam_download_flags')
      return
    end select
    if (walk_dir == 1) then
      call check(bsam_upload(gpu_ctx, p, f, a, n), 'bsam_upload')
    else
      call check(bsam_download(gpu_ctx, p, f, a, n), 'bsam_download')
    end if
  end function MovePatch

  subroutine walk(field, dir, onlevel)
    use NodeInfoDef
    use TreeOps, only: ApplyOnLevel
    integer, intent(in) :: field, dir
    integer, intent(in), optional :: onlevel
    type(funcparam) :: dummy
    integer :: level
    walk_field = field
    walk_dir = dir
    walk_level = -1
    if (present(onlevel)) walk_level = onlevel
    gpu_npatch = 0
    do level = minlevel, finestlevel
      call ApplyOnLevel(level, MovePatch, dummy)
    end do
    walk_level = -1
  end subroutine walk

  subroutine GpuUploadAll()
    call walk(BSAM_FIELD_Q, 1)
    call walk(BSAM_FIELD_QOLD, 1)
  end subroutine GpuUploadAll

  subroutine GpuDownloadAll()
    call walk(BSAM_FIELD_Q, 0)
  end subroutine GpuDownloadAll

  ! MultigridIterations (lib/afasroutines.f90:38-106) on the GPU.
  subroutine GpuMultigridIterations()
    use NodeInfoDef
    integer(c_int) :: n, conv
    call GpuSetDt()
    call GpuUploadAll()
    call check(bsam_multigrid_iterations(gpu_ctx, n, gpu_residuals, &
                                         int(size(gpu_residuals), c_int), conv), &
               'bsam_multigrid_iterations')
    gpu_cycles = n
    gpu_converged = conv
    lconverged = conv /= 0
    call GpuDownloadAll()
  end subroutine GpuMultigridIterations

  ! The truncation-error part of EstimateLevelErrors (lib/bsamroutines.f90:
  ! 971-981) on the GPU: uploads q of every patch, restricts level onto
  ! level - 1, fills the coarse ghost layer, copies q to qold on both levels
  ! and evaluates RelativeTruncationError (lib/afasroutines.f90:716-797) on
  ! the level's patches. Downloads q and qold of levels level - 1 and level,
  ! qc and qrte of level: the host arrays are then what the reference calls
  ! leave.
  subroutine GpuRelativeTruncationError(level)
    integer, intent(in) :: level
    call GpuSetDt()
    call walk(BSAM_FIELD_Q, 1)
    call check(bsam_truncation_error(gpu_ctx, level), 'bsam_truncation_error')
    call walk(BSAM_FIELD_Q, 0, level - 1)
    call walk(BSAM_FIELD_QOLD, 0, level - 1)
    call walk(BSAM_FIELD_Q, 0, level)
    call walk(BSAM_FIELD_QOLD, 0, level)
    call walk(BSAM_FIELD_QC, 0, level)
    call walk(BSAM_FIELD_QRTE, 0, level)
  end subroutine GpuRelativeTruncationError

  ! SetErrFlags2D/3D (lib/bsamroutines.f90:1093-1157), the default tagging
  ! of ErrFlag, on every patch of a level from the device qrte that
  ! GpuRelativeTruncationError left, with qtolerance(level); the flags land
  ! in each nodeinfo's errorflags.
  subroutine GpuSetErrFlags(level)
    use NodeInfoDef, only: qtolerance
    integer, intent(in) :: level
    call check(bsam_error_flags(gpu_ctx, level, real(qtolerance(level), c_double)), &
               'bsam_error_flags')
    call walk(0, 0, level)
  end subroutine GpuSetErrFlags

  ! ErrFlag's user tagging (lib/bsamroutines.f90:1061-1075: the problem
  ! module's SetErrFlagsUser2D/3D, problem.f90:135-163, :425-452) on every
  ! patch of a level: uploads the level's q (ghosts as AMR's SetGhost(level,
  ! 0) left them), tags on the GPU with qtolerance(level) and downloads each
  ! nodeinfo's errorflags.
  subroutine GpuSetErrFlagsUser(level)
    use NodeInfoDef, only: qtolerance
    integer, intent(in) :: level
    call walk(BSAM_FIELD_Q, 1, level)
    call check(bsam_error_flags_user(gpu_ctx, level, real(qtolerance(level), c_double)), &
               'bsam_error_flags_user')
    call walk(0, 0, level)
  end subroutine GpuSetErrFlagsUser

  ! TakeTimeSteps' after-step statistics (lib/bsamroutines.f90:690-697) from
  ! the device q the last solve left: integralresult(1) the mass of
  ! variable 1 on the composite grid, integralresult(2) the 2D energy.
  subroutine GpuAfterStepStatistics()
    use NodeInfoDef, only: integralresult
    real(c_double) :: m, e
    call check(bsam_after_step_stats(gpu_ctx, m, e), 'bsam_after_step_stats')
    integralresult(1) = m
    integralresult(2) = e
  end subroutine GpuAfterStepStatistics

  ! SetGhost(level, ipass) (lib/boundary.f90:58-89) on the device copy.
  subroutine GpuSetGhost(level, ipass)
    integer, intent(in) :: level, ipass
    call check(bsam_set_ghost(gpu_ctx, level, ipass), 'bsam_set_ghost')
  end subroutine GpuSetGhost

  ! The engine's SetGhost program applied on the host to the reference's own
  ! q arrays (valid on a host-only context, device = -1).
  subroutine GpuApplyGhostHost(level, ipass)
    integer, intent(in) :: level, ipass
    call check(bsam_ghost_program_apply_host(gpu_ctx, level, ipass, qptr), &
               'bsam_ghost_program_apply_host')
  end subroutine GpuApplyGhostHost

end module AFASGPU
