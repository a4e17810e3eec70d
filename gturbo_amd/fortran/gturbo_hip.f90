! gturbo_hip.f90 -- ISO_C_BINDING interface to libgturbo_hip.so (the C ABI of
! include/gturbo.h).  Array arguments are assumed-size, so Fortran passes the
! address of the caller's array, which is the key of the library's present
! table (the role of OpenACC's present table in the reference,
! src/main.f90:67-83).
module gturbo_hip
  use iso_c_binding
  implicit none
  private
  public :: gtctx, gtcheck, gt_last_error_string
  public :: gt_create, gt_destroy, gt_info, gt_plan_create, gt_mem_info
  public :: gt_enter_data, gt_exit_data, gt_update_host, gt_update_device, gt_copy
  public :: gt_synchronize
  public :: gt_fft_dir, gt_fft_inv, gt_produ, gt_trunc, gt_nlt, gt_step, gt_step4
  public :: gt_forcing, gt_spectrum, gt_fluxes, gt_glob
  public :: gt_adj_forward, gt_adj_backward
  public :: GT_PLAN_D2Z, GT_PLAN_Z2D

  integer(c_int32_t), parameter :: GT_PLAN_D2Z = 1, GT_PLAN_Z2D = 2
  type(c_ptr), save :: gtctx = c_null_ptr

  interface
    integer(c_int) function gt_create(n, params, ctx) bind(C, name="gt_create")
      import :: c_int, c_int32_t, c_double, c_ptr
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: params(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function gt_destroy(ctx) bind(C, name="gt_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function gt_last_error() bind(C, name="gt_last_error")
      import :: c_ptr
    end function
    integer(c_int) function gt_info(ctx, iinfo, dinfo) bind(C, name="gt_info")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: iinfo(*)
      real(c_double), intent(out) :: dinfo(*)
    end function
    integer(c_int) function gt_plan_create(ctx, kind, plan) bind(C, name="gt_plan_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind
      integer(c_int32_t), intent(out) :: plan
    end function
    integer(c_int) function gt_mem_info(free_b, total_b) bind(C, name="gt_mem_info")
      import :: c_int, c_size_t
      integer(c_size_t), intent(out) :: free_b, total_b
    end function
    integer(c_int) function gt_enter_data(ctx, host, bytes, copyin) bind(C, name="gt_enter_data")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: host(*)
      integer(c_int64_t), value :: bytes
      integer(c_int32_t), value :: copyin
    end function
    integer(c_int) function gt_exit_data(ctx, host, copyout) bind(C, name="gt_exit_data")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: host(*)
      integer(c_int32_t), value :: copyout
    end function
    integer(c_int) function gt_update_host(ctx, host) bind(C, name="gt_update_host")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: host(*)
    end function
    integer(c_int) function gt_update_device(ctx, host) bind(C, name="gt_update_device")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: host(*)
    end function
    integer(c_int) function gt_copy(ctx, dst, src) bind(C, name="gt_copy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: dst(*), src(*)
    end function
    integer(c_int) function gt_synchronize(ctx) bind(C, name="gt_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gt_fft_dir(ctx, z, plan) bind(C, name="gt_fft_dir")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*)
      integer(c_int32_t), value :: plan
    end function
    integer(c_int) function gt_fft_inv(ctx, z, plan) bind(C, name="gt_fft_inv")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*)
      integer(c_int32_t), value :: plan
    end function
    integer(c_int) function gt_produ(ctx, z, u, v) bind(C, name="gt_produ")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*), u(*), v(*)
    end function
    integer(c_int) function gt_trunc(ctx, z) bind(C, name="gt_trunc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*)
    end function
    integer(c_int) function gt_nlt(ctx, z, u, v) bind(C, name="gt_nlt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*), u(*), v(*)
    end function
    integer(c_int) function gt_step(ctx, z, znl, u, v) bind(C, name="gt_step")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*), znl(*), u(*), v(*)
    end function
    integer(c_int) function gt_step4(ctx, z, znl, znew, u, v) bind(C, name="gt_step4")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*), znl(*), znew(*), u(*), v(*)
    end function
    integer(c_int) function gt_forcing(ctx, z, seed) bind(C, name="gt_forcing")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*)
      integer(c_int32_t), intent(inout) :: seed
    end function
    integer(c_int) function gt_spectrum(ctx, z, sp, sz) bind(C, name="gt_spectrum")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*)
      real(c_double), intent(out) :: sp(*), sz(*)
    end function
    integer(c_int) function gt_fluxes(ctx, z, znl, fle, flz) bind(C, name="gt_fluxes")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*), znl(*)
      real(c_double), intent(out) :: fle(*), flz(*)
    end function
    integer(c_int) function gt_glob(ctx, z, out) bind(C, name="gt_glob")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: z(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function gt_adj_forward(ctx, w0, wt, nsteps, kchk, seed) &
        bind(C, name="gt_adj_forward")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: w0(*), wt(*)
      integer(c_int32_t), value :: nsteps, kchk
      integer(c_int32_t), intent(inout) :: seed
    end function
    integer(c_int) function gt_adj_backward(ctx, gt, g0) bind(C, name="gt_adj_backward")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: gt(*), g0(*)
    end function
  end interface

contains

  function gt_last_error_string() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ''
    cp = gt_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function

  ! Error convention: every entry returns 0 or a GT_ERR_* code; the Fortran
  ! side stops with the library's message.
  subroutine gtcheck(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= 0) then
      write(0, '(a,a,a,i0,a,a)') 'gturbo: ', what, ' failed (', rc, '): ', &
        trim(gt_last_error_string())
      error stop 1
    end if
  end subroutine

end module gturbo_hip
