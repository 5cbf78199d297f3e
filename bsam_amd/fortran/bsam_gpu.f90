! bsam_gpu.f90 -- ISO_C_BINDING interface to the MI355X FAS engine
! (include/bsam_gpu.h). Every procedure returns 0 on success, <0 on error;
! bsam_last_error returns the message as a C string.
module bsam_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: BSAM_FIELD_F = 1, BSAM_FIELD_Q = 2, BSAM_FIELD_QOLD = 3, &
                               BSAM_FIELD_QC = 4, BSAM_FIELD_FTMP = 5, BSAM_FIELD_RF = 6, &
                               BSAM_FIELD_QRTE = 7
  integer(c_int), parameter :: BSAM_PREC_FP64 = 0, BSAM_PREC_MIXED = 1

  type, bind(C) :: bsam_cfg
    integer(c_int) :: ndims, nrvars, mbc, minlevel
    real(c_double) :: dt, eps2
    real(c_double) :: fc_alpha, fc_shift, fe_num, fe_den
    integer(c_int) :: nsmoothingpasses, maxvcycles
    real(c_double) :: qerrortol
    integer(c_int) :: precision, device
    integer(c_int) :: errortype
  end type bsam_cfg

  interface
    function bsam_ctx_create(cfg) bind(C, name='bsam_ctx_create') result(ctx)
      import :: c_ptr, bsam_cfg
      type(bsam_cfg), intent(in) :: cfg
      type(c_ptr) :: ctx
    end function
    function bsam_ctx_destroy(ctx) bind(C, name='bsam_ctx_destroy') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function bsam_last_error(ctx) bind(C, name='bsam_last_error') result(msg)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: msg
    end function
    function bsam_hierarchy_begin(ctx) bind(C, name='bsam_hierarchy_begin') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function bsam_add_patch(ctx, id, level, parent, mx, mbounds, mglobal, mthbc, dx, &
                            xlower) bind(C, name='bsam_add_patch') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id, level, parent
      integer(c_int), intent(in) :: mx(3), mbounds(6), mglobal(6), mthbc(6)
      real(c_double), value :: dx
      real(c_double), intent(in) :: xlower(3)
      integer(c_int) :: rc
    end function
    function bsam_hierarchy_commit(ctx) bind(C, name='bsam_hierarchy_commit') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function bsam_upload(ctx, patch, field, host, n) bind(C, name='bsam_upload') result(rc)
      import :: c_ptr, c_int, c_long
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: patch, field
      integer(c_long), value :: n
      integer(c_int) :: rc
    end function
    function bsam_download(ctx, patch, field, host, n) bind(C, name='bsam_download') &
        result(rc)
      import :: c_ptr, c_int, c_long
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: patch, field
      integer(c_long), value :: n
      integer(c_int) :: rc
    end function
    function bsam_set_ghost(ctx, level, ipass) bind(C, name='bsam_set_ghost') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, ipass
      integer(c_int) :: rc
    end function
    function bsam_fill_down(ctx, field) bind(C, name='bsam_fill_down') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      integer(c_int) :: rc
    end function
    function bsam_vcycle(ctx, level) bind(C, name='bsam_vcycle') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      integer(c_int) :: rc
    end function
    function bsam_error_afas(ctx, level, residual) bind(C, name='bsam_error_afas') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), intent(out) :: residual
      integer(c_int) :: rc
    end function
    function bsam_multigrid_iterations(ctx, ncycles, hist, maxhist, converged) &
        bind(C, name='bsam_multigrid_iterations') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: ncycles, converged
      real(c_double), intent(out) :: hist(*)
      integer(c_int), value :: maxhist
      integer(c_int) :: rc
    end function
    function bsam_truncation_error(ctx, level) bind(C, name='bsam_truncation_error') &
        result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      integer(c_int) :: rc
    end function
    function bsam_error_flags(ctx, level, qtolerance) bind(C, name='bsam_error_flags') &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), value :: qtolerance
      integer(c_int) :: rc
    end function
    function bsam_error_flags_user(ctx, level, qtolerance) &
        bind(C, name='bsam_error_flags_user') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), value :: qtolerance
      integer(c_int) :: rc
    end function
    function bsam_after_step_stats(ctx, mass, energy) bind(C, name='bsam_after_step_stats') &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mass, energy
      integer(c_int) :: rc
    end function
    function bsam_download_flags(ctx, patch, host, n) bind(C, name='bsam_download_flags') &
        result(rc)
      import :: c_ptr, c_int, c_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: patch
      type(c_ptr), value :: host
      integer(c_long), value :: n
      integer(c_int) :: rc
    end function
    function bsam_tune_get(ctx, key, value) bind(C, name='bsam_tune_get') result(rc)
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: key(*)
      real(c_double), intent(out) :: value
      integer(c_int) :: rc
    end function
    function bsam_set_dt(ctx, dt) bind(C, name='bsam_set_dt') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int) :: rc
    end function
    function bsam_set_qpo(ctx, qpo, n) bind(C, name='bsam_set_qpo') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: qpo
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function
    function bsam_sync(ctx) bind(C, name='bsam_sync') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function bsam_ghost_program_apply_host(ctx, level, ipass, q_patches) &
        bind(C, name='bsam_ghost_program_apply_host') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, ipass
      type(c_ptr), intent(in) :: q_patches(*)
      integer(c_int) :: rc
    end function
  end interface

contains

  ! The message of the last failed call as a Fortran string.
  function bsam_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    msg = ''
    call c_f_pointer(bsam_last_error(ctx), s, [256])
    do n = 1, 256
      if (s(n) == c_null_char) exit
      msg(n:n) = s(n)
    end do
  end function bsam_error_message

end module bsam_gpu
