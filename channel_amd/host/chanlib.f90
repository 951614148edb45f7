!
! chanlib.f90 -- ISO_C_BINDING interface to the MI355X time-step library
! (include/channel_hip.h).  The derived types mirror the C structs field by
! field; every function returns 0 on success, a HIP error code (> 0) or a
! CHAN_E_* code (< 0) whose message chan_error_message() returns.
!
MODULE chanlib
  USE, intrinsic :: iso_c_binding
  IMPLICIT NONE

  integer(C_INT), parameter :: CHAN_FP64 = 0, CHAN_SPLIT = 1

  TYPE, bind(C) :: chan_params
    integer(C_INT32_T) :: nx, ny, nz, precision
    real(C_DOUBLE) :: alfa0, beta0, ni, a, ymin, ymax
    integer(C_INT32_T) :: cpi, cpi_type
    real(C_DOUBLE) :: gamma, meanpx, meanpz, meanflowx, meanflowz, u0, uN, deltat, cflmax, time
  END TYPE chan_params

  TYPE, bind(C) :: chan_scalars
    real(C_DOUBLE) :: time, deltat, meanpx, meanpz, corrpx, corrpz, fr(3), cfl
  END TYPE chan_scalars

  INTERFACE
    integer(C_INT) FUNCTION chan_abi_version() bind(C, name="chan_abi_version")
      import :: C_INT
    END FUNCTION
    type(C_PTR) FUNCTION chan_last_error() bind(C, name="chan_last_error")
      import :: C_PTR
    END FUNCTION
    integer(C_INT) FUNCTION chan_create(p, ctx) bind(C, name="chan_create")
      import :: C_INT, C_PTR, chan_params
      type(chan_params), intent(in) :: p
      type(C_PTR), intent(out) :: ctx
    END FUNCTION
    integer(C_INT) FUNCTION chan_destroy(ctx) bind(C, name="chan_destroy")
      import :: C_INT, C_PTR
      type(C_PTR), value :: ctx
    END FUNCTION
    integer(C_INT) FUNCTION chan_sizes(ctx, sizes) bind(C, name="chan_sizes")
      import :: C_INT, C_INT32_T, C_PTR
      type(C_PTR), value :: ctx
      integer(C_INT32_T), intent(out) :: sizes(5)
    END FUNCTION
    integer(C_INT) FUNCTION chan_setup_host(p, sizes, y, der, bnd, bc, d0fac) bind(C, name="chan_setup_host")
      import :: C_INT, C_INT32_T, C_PTR, chan_params
      type(chan_params), intent(in) :: p
      integer(C_INT32_T), intent(out) :: sizes(5)
      type(C_PTR), value :: y, der, bnd, bc, d0fac
    END FUNCTION
    integer(C_INT) FUNCTION chan_set_field(ctx, V) bind(C, name="chan_set_field")
      import :: C_INT, C_PTR, C_DOUBLE_COMPLEX
      type(C_PTR), value :: ctx
      complex(C_DOUBLE_COMPLEX), intent(in) :: V(*)
    END FUNCTION
    integer(C_INT) FUNCTION chan_get_field(ctx, V) bind(C, name="chan_get_field")
      import :: C_INT, C_PTR, C_DOUBLE_COMPLEX
      type(C_PTR), value :: ctx
      complex(C_DOUBLE_COMPLEX), intent(out) :: V(*)
    END FUNCTION
    integer(C_INT) FUNCTION chan_get_scalars(ctx, s) bind(C, name="chan_get_scalars")
      import :: C_INT, C_PTR, chan_scalars
      type(C_PTR), value :: ctx
      type(chan_scalars), intent(out) :: s
    END FUNCTION
    integer(C_INT) FUNCTION chan_set_scalars(ctx, s) bind(C, name="chan_set_scalars")
      import :: C_INT, C_PTR, chan_scalars
      type(C_PTR), value :: ctx
      type(chan_scalars), intent(in) :: s
    END FUNCTION
    integer(C_INT) FUNCTION chan_initial(ctx) bind(C, name="chan_initial")
      import :: C_INT, C_PTR
      type(C_PTR), value :: ctx
    END FUNCTION
    integer(C_INT) FUNCTION chan_buildrhs(ctx, ode, compute_cfl) bind(C, name="chan_buildrhs")
      import :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
      type(C_PTR), value :: ctx
      real(C_DOUBLE), intent(in) :: ode(3)
      integer(C_INT32_T), value :: compute_cfl
    END FUNCTION
    integer(C_INT) FUNCTION chan_linsolve(ctx, lambda) bind(C, name="chan_linsolve")
      import :: C_INT, C_PTR, C_DOUBLE
      type(C_PTR), value :: ctx
      real(C_DOUBLE), value :: lambda
    END FUNCTION
    integer(C_INT) FUNCTION chan_outstats(ctx, row) bind(C, name="chan_outstats")
      import :: C_INT, C_PTR, C_DOUBLE
      type(C_PTR), value :: ctx
      real(C_DOUBLE), intent(out) :: row(11)
    END FUNCTION
    integer(C_INT) FUNCTION chan_step(ctx, row) bind(C, name="chan_step")
      import :: C_INT, C_PTR, C_DOUBLE
      type(C_PTR), value :: ctx
      real(C_DOUBLE), intent(out) :: row(11)
    END FUNCTION
    integer(C_INT) FUNCTION chan_set_body_force_feedback(ctx, src, gain) bind(C, name="chan_set_body_force_feedback")
      import :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
      type(C_PTR), value :: ctx
      integer(C_INT32_T), intent(in) :: src(3)
      real(C_DOUBLE), intent(in) :: gain(*)
    END FUNCTION
    integer(C_INT) FUNCTION chan_keep_body_force(ctx, on) bind(C, name="chan_keep_body_force")
      import :: C_INT, C_INT32_T, C_PTR
      type(C_PTR), value :: ctx
      integer(C_INT32_T), value :: on
    END FUNCTION
    integer(C_INT) FUNCTION chan_get_applied_body_force(ctx, F) bind(C, name="chan_get_applied_body_force")
      import :: C_INT, C_PTR, C_DOUBLE_COMPLEX
      type(C_PTR), value :: ctx
      complex(C_DOUBLE_COMPLEX), intent(out) :: F(*)
    END FUNCTION
    integer(C_INT) FUNCTION chan_sync(ctx) bind(C, name="chan_sync")
      import :: C_INT, C_PTR
      type(C_PTR), value :: ctx
    END FUNCTION
    integer(C_SIZE_T) FUNCTION c_strlen(s) bind(C, name="strlen")
      import :: C_SIZE_T, C_PTR
      type(C_PTR), value :: s
    END FUNCTION
  END INTERFACE

CONTAINS

  ! message of the last failing library call
  FUNCTION chan_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(C_PTR) :: p
    character(kind=C_CHAR), pointer :: s(:)
    integer :: n, i
    p = chan_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  END FUNCTION chan_error_message

  ! fail-stop on a nonzero status, as the reference STOPs on errors
  SUBROUTINE chan_check(rc, what)
    integer(C_INT), intent(in) :: rc
    character(len=*), intent(in) :: what
    IF (rc /= 0) THEN
      WRITE(*,"(A,A,A,I0,A,A)") "ERROR: ", what, " failed (", rc, "): ", chan_error_message()
      ERROR STOP 1
    END IF
  END SUBROUTINE chan_check

END MODULE chanlib
