!*****
!***** ISO_C_BINDING INTERFACE TO libcnd (include/cnd.h)
!*****
! Fortran declarations of the C ABI that replaces the reference's per-block
! step (src/mo_canard_driver.f90:226-431).  The derived types mirror
! cnd_params and cnd_block field by field.

module mo_cnd
   use iso_c_binding
   implicit none
   public

   integer(c_int), parameter :: CND_OK = 0, CND_FP64 = 0, CND_MIXED = 1
   integer(c_int), parameter :: CND_NSEC = 8
   integer(c_int), parameter :: CND_KIND_FILTER = 0, CND_KIND_GRAD = 1, CND_KIND_FLUX = 2, CND_KIND_FIELD = 3

   type, bind(C) :: cnd_params
      real(c_double)   :: fltk, fltrbc
      real(c_double)   :: reoo, tempoo, amach1, amach2, amach3, timf
      integer(c_int)   :: nsmf
      real(c_double)   :: cfl
      integer(c_int)   :: viscous
      integer(c_int)   :: sponge
      real(c_double)   :: szco, szth0, szth1, doml0, doml1
   end type cnd_params

   type, bind(C) :: cnd_block
      integer(c_int) :: lxi, let, lze
      integer(c_int) :: nbc(2,3)      ! C nbc[3][2]: (side, direction)
      integer(c_int) :: mcd(2,3)
      integer(c_int) :: inst
   end type cnd_block

   interface
      integer(c_int) function cnd_create(params, ninst, blocks, nblocks, precision, ctx) bind(C, name="cnd_create")
         import :: c_int, c_ptr, cnd_params, cnd_block
         type(cnd_params), intent(in) :: params(*)
         integer(c_int), value :: ninst, nblocks, precision
         type(cnd_block), intent(in) :: blocks(*)
         type(c_ptr), intent(out) :: ctx
      end function
      subroutine cnd_destroy(ctx) bind(C, name="cnd_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine
      type(c_ptr) function cnd_last_error(ctx) bind(C, name="cnd_last_error")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_set_grid(ctx, b, xyz) bind(C, name="cnd_set_grid")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: b
         real(c_double), intent(in) :: xyz(*)
      end function
      integer(c_int) function cnd_init_metrics(ctx) bind(C, name="cnd_init_metrics")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_set_state(ctx, b, q) bind(C, name="cnd_set_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: b
         real(c_double), intent(in) :: q(*)
      end function
      integer(c_int) function cnd_get_state(ctx, b, q) bind(C, name="cnd_get_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: b
         real(c_double), intent(out) :: q(*)
      end function
      integer(c_int) function cnd_set_clock(ctx, inst, n, ndt, dt, dts, dte, timo) bind(C, name="cnd_set_clock")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: inst, n, ndt
         real(c_double), value :: dt, dts, dte, timo
      end function
      integer(c_int) function cnd_get_clock(ctx, inst, n, ndt, dt, dts, dte, timo) bind(C, name="cnd_get_clock")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: inst
         integer(c_int), intent(out) :: n, ndt
         real(c_double), intent(out) :: dt, dts, dte, timo
      end function
      integer(c_int) function cnd_step(ctx, nsteps) bind(C, name="cnd_step")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsteps
      end function
      integer(c_int) function cnd_filter(ctx) bind(C, name="cnd_filter")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_rk_stage(ctx, nk) bind(C, name="cnd_rk_stage")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nk
      end function
      integer(c_int) function cnd_end_step(ctx) bind(C, name="cnd_end_step")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_record(ctx, on) bind(C, name="cnd_record")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
      end function
      integer(c_int) function cnd_record_read(ctx, b, out) bind(C, name="cnd_record_read")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: ctx
         integer(c_int), value :: b
         real(c_float), intent(out) :: out(*)
      end function
      integer(c_int) function cnd_record_reset(ctx) bind(C, name="cnd_record_reset")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_make_grid(ctx, b, kind, par, npar) bind(C, name="cnd_make_grid")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: b, kind, npar
         real(c_double), intent(in) :: par(*)
      end function
      integer(c_int) function cnd_output_schedule(ctx, inst, tmax, ndata) bind(C, name="cnd_output_schedule")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: inst, ndata
         real(c_double), value :: tmax
      end function
      integer(c_int) function cnd_output_due(ctx, inst, due, ndati, time) bind(C, name="cnd_output_due")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: inst
         integer(c_int), intent(out) :: due, ndati
         real(c_double), intent(out) :: time
      end function
      integer(c_int) function cnd_output_write(ctx, inst, path, async) bind(C, name="cnd_output_write")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: ctx
         integer(c_int), value :: inst, async
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function cnd_output_grid(ctx, inst, path) bind(C, name="cnd_output_grid")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: ctx
         integer(c_int), value :: inst
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function cnd_output_wait(ctx) bind(C, name="cnd_output_wait")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_use_graph(ctx, enable) bind(C, name="cnd_use_graph")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable
      end function
      integer(c_int) function cnd_profile(ctx, enable) bind(C, name="cnd_profile")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable
      end function
      integer(c_int) function cnd_profile_read(ctx, section, ms, launches) bind(C, name="cnd_profile_read")
         import :: c_int, c_ptr, c_double, c_long_long
         type(c_ptr), value :: ctx
         integer(c_int), value :: section
         real(c_double), intent(out) :: ms
         integer(c_long_long), intent(out) :: launches
      end function
      integer(c_int) function cnd_tune(ctx) bind(C, name="cnd_tune")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function cnd_set_shape(ctx, kind, dir, shape) bind(C, name="cnd_set_shape")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind, dir, shape
      end function
      integer(c_int) function cnd_tuning(ctx, kind, dir, shape, seg_len, pencils, segments, ms) &
            bind(C, name="cnd_tuning")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind, dir
         integer(c_int), intent(out) :: shape, seg_len, pencils, segments
         real(c_double), intent(out) :: ms
      end function
   end interface

contains

   ! Message of the last error of ctx (or of cnd_create when ctx is null).
   function cnd_error_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=256) :: msg
      character(kind=c_char), pointer :: s(:)
      type(c_ptr) :: p
      integer :: i
      msg = ' '
      p = cnd_last_error(ctx)
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [256])
      do i = 1, 256
         if (s(i) == c_null_char) exit
         msg(i:i) = s(i)
      end do
   end function cnd_error_message

   ! Stop with the library's message when a call did not succeed.
   subroutine cnd_check(rc, ctx, what)
      integer(c_int), intent(in) :: rc
      type(c_ptr), intent(in) :: ctx
      character(len=*), intent(in) :: what
      if (rc /= CND_OK) then
         write(*,'(a,a,a,i0,a,a)') 'Canard: ', what, ' failed (', rc, '): ', trim(cnd_error_message(ctx))
         stop 1
      end if
   end subroutine cnd_check

end module mo_cnd
