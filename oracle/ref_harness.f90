! Fixture driver for the reference BSAM Fortran modules.
!
! This program is test infrastructure written for this repository. It is
! linked against the reference's own modules (lib/*.f90 and the
! simple_cahn-hilliard problem module, compiled unmodified from the
! reference tree by oracle/Makefile) and never ships with the product.
!
! It builds a patch hierarchy from a text spec, loads FP64 field arrays from
! a stream file, runs a sequence of reference operations, and writes every
! patch's arrays back out as raw FP64 (column-major, as allocated by
! AllocFields, lib/bsamstorage.f90:46-123).
!
! Spec file (list-directed):
!   ndims nrvars mbc
!   dt alpha eps
!   nsmoothingpasses maxvcycles qerrortol
!   minlevel finestlevel
!   mx(1:3) xlower(1:3) dx mthbc(1:6)          -- root patch, id 0
!   nrefined
!   id level parentid mb(1,1) mb(1,2) mb(2,1) mb(2,2) mb(3,1) mb(3,2)   (x nrefined, creation order)
!   nops
!   opname arg1 arg2                            (x nops)
!
! Ops: setghost L ipass | relax L color | restrict L | coarseghost L |
!      coarseload L | correct L | filldown field | source 0 0 | vcycle L 0 |
!      residual L 0 | mgiter 0 0 | step nsteps 0 | copyqold 0 0
!
! Input stream: for every level minlevel..finestlevel in level-list order
! (ApplyOnLevel order), the arrays q, f, qc of each patch.
! Output stream: the same order, arrays q, f, qc, rf.
! Residual history (mgiter/step ops): one line per V-cycle in res.txt.
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
    DumpPatch = err_ok
  end function DumpPatch

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

program RefHarness
  use NodeInfoDef
  use ProblemDef
  use TreeOps
  use BSAMStorage, only: AllocFields
  use Boundary, only: SetGhost, PeriodicSetup, GetCoarseGhostPoints
  use AFASRoutines
  use HarnessState
  use HarnessCallbacks
  implicit none
  character(len=512)     :: specname, inname, outname, resname
  character(len=32)      :: opname
  type(funcparam)        :: dummy
  type(nodeinfo), pointer :: rootinfo
  integer :: ierror, mbc_in, nrv_in, finest_in, nref, i, lev, nops, iop
  integer :: a1, a2, it, istep, maxv_user
  integer, dimension(1:3) :: rmx
  integer, dimension(1:6) :: rmthbc
  real, dimension(1:3)    :: rxl
  real    :: rdx, r
  integer, allocatable :: cid(:), clev(:), cpar(:), cmb(:,:,:)

  call get_command_argument(1, specname)
  call get_command_argument(2, inname)
  call get_command_argument(3, outname)
  call get_command_argument(4, resname)

  open(10, file=trim(specname), status='old', action='read')
  read(10,*) ndims, nrv_in, mbc_in
  read(10,*) dt, alpha, eps
  eps2 = eps*eps
  read(10,*) nsmoothingpasses, maxv_user, qerrortol
  read(10,*) minlevel, finest_in
  read(10,*) rmx, rxl, rdx, rmthbc
  read(10,*) nref
  allocate(cid(nref), clev(nref), cpar(nref), cmb(3,2,nref))
  do i = 1, nref
    read(10,*) cid(i), clev(i), cpar(i), cmb(1,1,i), cmb(1,2,i), &
               cmb(2,1,i), cmb(2,2,i), cmb(3,1,i), cmb(3,2,i)
  end do

  errortype = 1
  updateauxfreq = 1
  eqtag = 1
  nrootgrids = 1
  maxlevel = finest_in
  finestlevel = 0
  qpo = 0.0
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

  ! Patch list in level-list order, for the caller to check its own order.
  h_unit = 13
  open(h_unit, file=trim(outname)//'.patches', status='replace', action='write')
  do lev = minlevel, finestlevel
    call ApplyOnLevel(lev, ListPatch, dummy)
  end do
  close(h_unit)

  h_unit = 11
  open(h_unit, file=trim(inname), access='stream', form='unformatted', &
       status='old', action='read')
  do lev = minlevel, finestlevel
    call ApplyOnLevel(lev, LoadPatch, dummy)
  end do
  close(h_unit)

  open(14, file=trim(resname), status='replace', action='write')
  read(10,*) nops
  do iop = 1, nops
    read(10,*) opname, a1, a2
    select case(trim(opname))
    case('setghost')
      call SetGhost(a1, a2)
    case('relax')
      dummy%iswitch = a2
      call ApplyOnLevel(a1, RelaxPatch, dummy)
    case('restrict')
      call ApplyOnLevel(a1, RestrictSolution, dummy)
    case('coarseghost')
      call ApplyOnLevel(a1, GetCoarseGhostPoints, dummy)
    case('coarseload')
      call ApplyOnLevel(a1, CoarseLoadingFunction, dummy)
    case('correct')
      call ApplyOnLevel(a1, CorrectFine, dummy)
    case('filldown')
      call FillDown(a1)
    case('source')
      do lev = finestlevel, rootlevel, -1
        call ApplyOnLevel(lev, GetSourceFunction, dummy)
      end do
      do lev = finestlevel, rootlevel, -1
        call ApplyOnLevel(lev, UpdateSourceFunction, dummy)
      end do
    case('vcycle')
      call AFASVcycle(a1)
    case('residual')
      r = ErrorAFAS(a1)
      write(14,'(i6,1x,i6,1x,es25.17)') 0, 0, r
    case('copyqold')
      do lev = minlevel, finestlevel
        call ApplyOnLevel(lev, CopyToOld, dummy)
      end do
    case('mgiter')
      ! MultigridIterations (lib/afasroutines.f90:38-106) called once per
      ! V-cycle so that every cycle's residual is recorded; repeating the
      ! prologue between cycles is idempotent on a fixed hierarchy.
      maxvcycles = 1
      do it = 1, maxv_user
        call MultigridIterations
        r = ErrorAFAS(finestlevel)
        write(14,'(i6,1x,i6,1x,es25.17)') 1, it, r
        if (r < qerrortol) exit
      end do
    case('step')
      ! a1 time steps on a fixed hierarchy, following TakeTimeSteps
      ! (lib/bsamroutines.f90:648-723) without regridding.
      maxvcycles = 1
      do istep = 1, a1
        do lev = rootlevel, finestlevel
          call SetGhost(lev, 0)
        end do
        call FillDown(solutionfield)
        do lev = minlevel, finestlevel
          call ApplyOnLevel(lev, CopyToOld, dummy)
        end do
        do it = 1, maxv_user
          call MultigridIterations
          r = ErrorAFAS(finestlevel)
          write(14,'(i6,1x,i6,1x,es25.17)') istep, it, r
          if (r < qerrortol) exit
        end do
      end do
    case default
      print *, 'harness: unknown op ', trim(opname)
      stop 3
    end select
  end do
  close(14)
  close(10)

  h_unit = 12
  open(h_unit, file=trim(outname), access='stream', form='unformatted', &
       status='replace', action='write')
  do lev = minlevel, finestlevel
    call ApplyOnLevel(lev, DumpPatch, dummy)
  end do
  close(h_unit)
end program RefHarness
