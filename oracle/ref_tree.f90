! Tree-building helpers for programs linked against the reference BSAM
! modules (test infrastructure, see oracle/ref_harness.f90): create patches
! with the reference's own TreeOps/AllocFields following the geometry rules
! of InitSeed and MakeNewGrid, and stream patch arrays in and out.
module HarnessState
  use NodeInfoDef
  implicit none
  integer :: target_parent, new_id, new_level, h_unit
  integer, dimension(1:3,1:2) :: new_mb
  integer :: ghost_mbc
end module HarnessState

module HarnessCallbacks
contains
  integer function SeedFromChild(seedinfo, dummy)
    ! Below-root single-patch levels: half the child in every dimension,
    ! following the geometry of InitSeed (lib/bsamroutines.f90:426-494).
    use NodeInfoDef
    use TreeOps, only: err_ok, GetChildInfo
    use BSAMStorage, only: AllocFields
    implicit none
    type(nodeinfo)          :: seedinfo
    type(funcparam)         :: dummy
    type(nodeinfo), pointer :: child
    integer                 :: ierror
    ierror = GetChildInfo(child)
    seedinfo%nrvars = child%nrvars
    seedinfo%mbc = child%mbc
    seedinfo%maxlevel = maxlevel
    seedinfo%level = child%level-1
    seedinfo%mx = 1
    seedinfo%mx(1:ndims) = child%mx(1:ndims)/2
    seedinfo%dx = 0.0
    seedinfo%dx(1:ndims) = child%dx(1:ndims)*2.0
    seedinfo%mglobal = 1
    seedinfo%mglobal(1:ndims,2) = child%mglobal(1:ndims,2)/2
    child%mbounds = 1
    child%mbounds(1:ndims,2) = seedinfo%mx(1:ndims)
    seedinfo%mbounds = 1
    seedinfo%mbounds(1:ndims,2) = seedinfo%mx(1:ndims)
    seedinfo%ngrid = seedinfo%level
    seedinfo%tobedeleted = .false.
    seedinfo%time = child%time
    seedinfo%xlower = child%xlower
    seedinfo%xupper = child%xupper
    seedinfo%mthbc = child%mthbc
    seedinfo%maux = child%maux
    seedinfo%nroutvars = seedinfo%nrvars
    call AllocFields(seedinfo)
    SeedFromChild = err_ok
  end function SeedFromChild

  integer function SpawnChild(parent, dummy)
    ! Creates one child of the patch whose id is target_parent, using the
    ! child geometry rules of MakeNewGrid (lib/bsamroutines.f90:1581-1683).
    use NodeInfoDef
    use TreeOps, only: err_ok, CreateChild, GetChildInfo
    use BSAMStorage, only: AllocFields
    use HarnessState
    implicit none
    type(nodeinfo)          :: parent
    type(funcparam)         :: dummy
    type(nodeinfo), pointer :: child
    integer                 :: ierror, n, nb
    integer, dimension(1:3,1:2) :: gb
    SpawnChild = err_ok
    if (parent%ngrid /= target_parent) return
    call CreateChild
    ierror = GetChildInfo(child)
    child%tobedeleted = .false.
    child%newgrid = .false.
    child%fieldsallocated = .false.
    child%initialgrid = .true.
    child%maxlevel = parent%maxlevel
    child%nsubgrids = 0
    child%level = parent%level+1
    child%nrvars = parent%nrvars
    child%mbc = parent%mbc
    child%maux = parent%maux
    child%nroutvars = parent%nroutvars
    child%mx = 1
    child%mx(1:ndims) = (new_mb(1:ndims,2)-new_mb(1:ndims,1)+1)*2
    child%mbounds = 1
    child%mbounds(1:ndims,:) = new_mb(1:ndims,:)
    child%mglobal = 1
    gb(1:ndims,1) = parent%mglobal(1:ndims,1)+new_mb(1:ndims,1)-1
    gb(1:ndims,2) = gb(1:ndims,1)+new_mb(1:ndims,2)-new_mb(1:ndims,1)
    child%mglobal(1:ndims,1) = (gb(1:ndims,1)-1)*2+1
    child%mglobal(1:ndims,2) = gb(1:ndims,2)*2
    child%mthbc = internalbc
    do n = 1, ndims
      nb = 2*n-1
      if (parent%mthbc(nb)<internalbc .and. child%mbounds(n,1)==1) &
        child%mthbc(nb) = parent%mthbc(nb)
      nb = nb+1
      if (parent%mthbc(nb)<internalbc .and. child%mbounds(n,2)==parent%mx(n)) &
        child%mthbc(nb) = parent%mthbc(nb)
    end do
    child%ngrid = new_id
    child%time = parent%time
    child%xlower = 0.0
    child%xupper = 0.0
    child%xlower(1:ndims) = real(new_mb(1:ndims,1)-1)*parent%dx(1:ndims) &
                            + parent%xlower(1:ndims)
    child%xupper(1:ndims) = real(new_mb(1:ndims,2))*parent%dx(1:ndims) &
                            + parent%xlower(1:ndims)
    child%dx = 0.0
    child%dx(1:ndims) = parent%dx(1:ndims)/2.0
    call AllocFields(child, parent)
    if (child%level>finestlevel) finestlevel = child%level
  end function SpawnChild

  integer function LoadPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use HarnessState
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    read(h_unit) info%q
    read(h_unit) info%f
    read(h_unit) info%qc
    info%rf = 0.0
    LoadPatch = err_ok
  end function LoadPatch

  integer function DumpPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use HarnessState
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    write(h_unit) info%q
    write(h_unit) info%f
    write(h_unit) info%qc
    write(h_unit) info%rf
    write(h_unit) info%qold
    write(h_unit) info%qrte
    write(h_unit) real(info%errorflags)
    DumpPatch = err_ok
  end function DumpPatch

  ! ErrFlag's user tagging (lib/bsamroutines.f90:1061-1075) on one patch,
  ! with the same array sections: Problem's SetErrFlagsUser2D/3D.
  integer function UserFlagPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use Problem, only: SetErrFlagsUser2D, SetErrFlagsUser3D
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    integer :: nrvars
    integer, dimension(1:maxdims) :: mx, cmx
    real :: h
    real, dimension(1:maxdims) :: xlower
    nrvars = info%nrvars
    h = info%dx(1)
    xlower = 0.0
    xlower(1:ndims) = info%xlower(1:ndims)
    mx = 1
    mx(1:ndims) = info%mx(1:ndims)
    cmx = 1
    cmx(1:ndims) = mx(1:ndims)/2
    select case(ndims)
    case(2)
      call SetErrFlagsUser2D(info%qrte(1:cmx(1),1:cmx(2),1,1:nrvars), &
                             info%q(0:mx(1)+1,0:mx(2)+1,1,1:nrvars), &
                             info%errorflags(1:mx(1),1:mx(2),1), &
                             mx(1:2),cmx(1:2),nrvars,h,xlower(1:2),info%level)
    case(3)
      call SetErrFlagsUser3D(info%qrte(1:cmx(1),1:cmx(2),1:cmx(3),1:nrvars), &
                             info%q(0:mx(1)+1,0:mx(2)+1,0:mx(3)+1,1:nrvars), &
                             info%errorflags(1:mx(1),1:mx(2),1:mx(3)), &
                             mx(1:3),cmx(1:3),nrvars,h,xlower(1:3),info%level)
    end select
    UserFlagPatch = err_ok
  end function UserFlagPatch

  ! AfterStepStatistics (lib/bsamroutines.f90:896-939) on one patch, with
  ! the same array sections: Problem's AfterStep2D/3D accumulate mass and
  ! (2D) energy into integralresult.
  integer function StatsPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok, GetParentInfo
    use Problem, only: AfterStep2D, AfterStep3D
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    type(nodeinfo), pointer :: parent
    integer :: ierror, nrvars
    integer, dimension(1:maxdims) :: mx
    integer, dimension(1:maxdims,1:2) :: mb
    real :: h
    real, dimension(1:maxdims) :: xlower
    mx = 1
    mx(1:ndims) = info%mx(1:ndims)
    nrvars = info%nrvars
    h = info%dx(1)
    xlower = 0.0
    xlower(1:ndims) = info%xlower(1:ndims)
    mb = 1
    mb(1:ndims,1:2) = info%mbounds(1:ndims,1:2)
    ierror = GetParentInfo(parent)
    select case(ndims)
    case(2)
      call AfterStep2D(info%q(0:mx(1)+1,0:mx(2)+1,1,1:nrvars), &
                       parent%q(mb(1,1)-1:mb(1,2)+1,mb(2,1)-1:mb(2,2)+1,1,1:nrvars), &
                       mx(1:2),nrvars,h,xlower(1:2),info%level)
    case(3)
      call AfterStep3D(info%q(0:mx(1)+1,0:mx(2)+1,0:mx(3)+1,1:nrvars), &
                       parent%q(mb(1,1)-1:mb(1,2)+1,mb(2,1)-1:mb(2,2)+1, &
                                mb(3,1)-1:mb(3,2)+1,1:nrvars), &
                       mx(1:3),nrvars,h,xlower(1:3),info%level)
    end select
    StatsPatch = err_ok
  end function StatsPatch

  integer function ListPatch(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    use HarnessState
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    write(h_unit,'(i6,1x,i4,1x,3(i6,1x),6(i7,1x),6(i7,1x),6(i5,1x))') &
      info%ngrid, info%level, info%mx, info%mbounds, info%mglobal, info%mthbc
    ListPatch = err_ok
  end function ListPatch

  integer function CopyToOld(info, dummy)
    use NodeInfoDef
    use TreeOps, only: err_ok
    implicit none
    type(nodeinfo)  :: info
    type(funcparam) :: dummy
    info%qold = info%q
    CopyToOld = err_ok
  end function CopyToOld
end module HarnessCallbacks


module HarnessTree
contains
  ! Reads the spec header (see ref_harness.f90) from `unit` and builds the
  ! hierarchy through the reference's TreeOps: root as RootInit sets it up,
  ! the below-root chain, then the refined patches in creation order.
  subroutine build_tree(unit, maxv_user)
    use NodeInfoDef
    use ProblemDef
    use TreeOps
    use BSAMStorage, only: AllocFields
    use Boundary, only: PeriodicSetup
    use HarnessState
    use HarnessCallbacks
    implicit none
    integer, intent(in)  :: unit
    integer, intent(out) :: maxv_user
    type(funcparam)         :: dummy
    type(nodeinfo), pointer :: rootinfo
    integer :: ierror, mbc_in, nrv_in, finest_in, nref, i, lev
    integer, dimension(1:3) :: rmx
    integer, dimension(1:6) :: rmthbc
    real, dimension(1:3)    :: rxl
    real    :: rdx
    integer, allocatable :: cid(:), clev(:), cpar(:), cmb(:,:,:)
    character(len=128) :: qstr
    integer :: qlen, qstat
  read(unit,*) ndims, nrv_in, mbc_in
  read(unit,*) dt, alpha, eps
  eps2 = eps*eps
  read(unit,*) nsmoothingpasses, maxv_user, qerrortol
  read(unit,*) minlevel, finest_in
  read(unit,*) rmx, rxl, rdx, rmthbc
  read(unit,*) nref
  allocate(cid(nref), clev(nref), cpar(nref), cmb(3,2,nref))
  do i = 1, nref
    read(unit,*) cid(i), clev(i), cpar(i), cmb(1,1,i), cmb(1,2,i), &
               cmb(2,1,i), cmb(2,2,i), cmb(3,1,i), cmb(3,2,i)
  end do

  errortype = 1
  updateauxfreq = 1
  eqtag = 1
  nrootgrids = 1
  maxlevel = finest_in
  finestlevel = 0
  ! qpo(1:nrvars), the griddata namelist's quasi-periodic offsets
  ! (bsamroutines.f90:221): from BSAM_HARNESS_QPO when set, else 0 as
  ! RootInit leaves them
  qpo = 0.0
  call get_environment_variable('BSAM_HARNESS_QPO', qstr, qlen, qstat)
  if (qstat == 0 .and. qlen > 0) read(qstr,*) qpo(1:nrv_in)
  lconverged = .false.

  call InitForest
  ierror = GetRootInfo(rootinfo)
  rootinfo%ngrid = 0
  rootinfo%level = rootlevel
  rootinfo%maxlevel = maxlevel
  rootinfo%tobedeleted = .false.
  rootinfo%initialgrid = .true.
  rootinfo%mx = 1
  rootinfo%mx(1:ndims) = rmx(1:ndims)
  rootinfo%mglobal = 1
  rootinfo%mglobal(1:ndims,2) = rmx(1:ndims)
  rootinfo%time = 0.0
  rootinfo%xlower = 0.0
  rootinfo%xlower(1:ndims) = rxl(1:ndims)
  rootinfo%dx = 0.0
  rootinfo%dx(1:ndims) = rdx
  rootinfo%xupper = 0.0
  rootinfo%xupper(1:ndims) = rxl(1:ndims) + rdx*real(rmx(1:ndims))
  rootinfo%mbc = mbc_in
  rootinfo%mthbc = 1
  rootinfo%mthbc(1:2*ndims) = rmthbc(1:2*ndims)
  rootinfo%nrvars = nrv_in
  rootinfo%maux = 0
  rootinfo%nroutvars = nrv_in
  rootinfo%mbounds = 1
  rootinfo%mbounds(1:ndims,2) = rootinfo%mx(1:ndims)
  mxmax = 1
  mxmax(0,1:ndims) = rmx(1:ndims)
  do i = 1, maxdepth
    mxmax(i,1:ndims) = 2*mxmax(i-1,1:ndims)
  end do
  call PeriodicSetup(rootinfo)
  call AllocFields(rootinfo)

  call CreateBelowSeedLevels(minlevel)
  do lev = rootlevel-1, minlevel, -1
    call ApplyOnLevel(lev, SeedFromChild, dummy)
  end do

  do i = 1, nref
    target_parent = cpar(i)
    new_id = cid(i)
    new_level = clev(i)
    new_mb = cmb(:,:,i)
    call ApplyOnLevel(clev(i)-1, SpawnChild, dummy)
  end do
  if (finestlevel /= finest_in) then
    print *, 'harness: finest level mismatch', finestlevel, finest_in
    stop 2
  end if

  end subroutine build_tree
end module HarnessTree
