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
!      residual L 0 | mgiter 0 0 | step nsteps 0 | copyqold 0 0 | truncerr L 0 |
!      userflags L N (qtolerance(L) = N * 1e-3) | stats 0 0 |
!      uniform R 0 (WriteUniformMeshQ at time 0 with restartframe R: it
!      writes ./out/uRRRRR.dat under the working directory) |
!      errortype E 0 (ErrorAFAS's errortype for the ops that follow)
!
! Input stream: for every level minlevel..finestlevel in level-list order
! (ApplyOnLevel order), the arrays q, f, qc of each patch.
! Output stream: the same order, arrays q, f, qc, rf, qold, qrte and the
! error flags (as reals).
! Residual history (mgiter/step ops): MultigridIterations prints one line per
! V-cycle to stdout, after a 'harness step N' line per time step.
program RefHarness
  use NodeInfoDef
  use ProblemDef
  use TreeOps
  use BSAMStorage, only: AllocFields
  use Boundary, only: SetGhost, PeriodicSetup, GetCoarseGhostPoints
  use AFASRoutines
  use BSAMInputOutput, only: WriteUniformMeshQ
  use HarnessState
  use HarnessCallbacks
  use HarnessTree
  implicit none
  character(len=512)     :: specname, inname, outname, resname
  character(len=32)      :: opname
  type(funcparam)        :: dummy
  integer :: lev, nops, iop
  integer :: a1, a2, it, istep, maxv_user
  real    :: r

  call get_command_argument(1, specname)
  call get_command_argument(2, inname)
  call get_command_argument(3, outname)
  call get_command_argument(4, resname)

  open(10, file=trim(specname), status='old', action='read')
  call build_tree(10, maxv_user)

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
    case('truncerr')
      ! The truncation-error part of EstimateLevelErrors
      ! (lib/bsamroutines.f90:971-981), whose module keeps it private: the
      ! same public calls in the same order, CopyQToQold as qold = q.
      call ApplyOnLevel(a1, RestrictSolution, dummy)
      call SetGhost(a1 - 1, 0)
      call ApplyOnLevel(a1, GetCoarseGhostPoints, dummy)
      call ApplyOnLevel(a1, CopyToOld, dummy)
      call ApplyOnLevel(a1 - 1, CopyToOld, dummy)
      call ApplyOnLevel(a1, RelativeTruncationError, dummy)
    case('userflags')
      ! ErrFlag's errflaguser case on every patch of the level
      qtolerance(a1) = real(a2) * 1.0e-3
      call ApplyOnLevel(a1, UserFlagPatch, dummy)
    case('stats')
      ! TakeTimeSteps' after-step statistics (lib/bsamroutines.f90:690-697):
      ! levels finest to root, integralresult(1:2) = mass, energy
      integralresult = 0.0
      do lev = finestlevel, rootlevel, -1
        call ApplyOnLevel(lev, StatsPatch, dummy)
      end do
      write(14,'(i6,1x,i6,1x,es25.17)') -1, 1, integralresult(1)
      write(14,'(i6,1x,i6,1x,es25.17)') -1, 2, integralresult(2)
    case('errortype')
      errortype = a1
    case('uniform')
      restartframe = a1
      call WriteUniformMeshQ(0, 0.0)
    case('copyqold')
      do lev = minlevel, finestlevel
        call ApplyOnLevel(lev, CopyToOld, dummy)
      end do
    case('mgiter')
      ! MultigridIterations (lib/afasroutines.f90:38-106) itself, with
      ! maxvcycles from the spec; it prints every cycle's residual to stdout,
      ! which tests/golden/refrun.py records.
      write(*,'(a,i6)') 'harness step ', 1
      maxvcycles = maxv_user
      call MultigridIterations
    case('step')
      ! a1 time steps on a fixed hierarchy, following TakeTimeSteps
      ! (lib/bsamroutines.f90:648-723) without regridding.
      maxvcycles = maxv_user
      do istep = 1, a1
        do lev = rootlevel, finestlevel
          call SetGhost(lev, 0)
        end do
        call FillDown(solutionfield)
        do lev = minlevel, finestlevel
          call ApplyOnLevel(lev, CopyToOld, dummy)
        end do
        write(*,'(a,i6)') 'harness step ', istep
        call MultigridIterations
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
