!
! channel.f90 -- the channel DNS driver on one MI355X.  Reads dns.in and
! Dati.cart.out, runs the RK3 time loop through the HIP library
! (chanlib.f90), writes Runtimedata every step, Dati.cart.N.out every
! dt_field, Dati.cart.out every dt_save and at the end.
!
! Drop-in for the reference's PROGRAM channel (channel.f90:16-196): same
! input files and formats (dns.in: dnsdata.f90:98-125; restart header and
! payload: dnsdata.f90:677-720, 821-848; Runtimedata: dnsdata.f90:876-879),
! same restart semantics (time_from_restart, get_record: channel.f90:52-71,
! dnsdata.f90:181-218) and the same fail-stop behaviour.
!
! Two environment variables stand in for the reference's compile-time
! choices:
!   CHANNEL_BODYFORCE = am_f1 | am_butterfly | coriolis  -- the module that
!     header.h's bodyforce / BODYFORCE_MODULES select (header.h:28-31):
!     config_body_force, set_body_force before every buildrhs, and a
!     Force.cart.N.out beside each Dati.cart.N.out (dnsdata.f90:904-906);
!   CHANNEL_PRECISION = fp64 (default) | split  -- the FP32/FP64 split step.
!
PROGRAM channel
  USE, intrinsic :: iso_c_binding
  USE chanlib
  USE host_io
  IMPLICIT NONE
  type(chan_params) :: prm
  type(chan_scalars) :: sc
  type(C_PTR) :: ctx
  complex(C_DOUBLE_COMPLEX), allocatable, target :: V(:,:,:,:)
  real(C_DOUBLE), allocatable, target :: y(:)
  real(C_DOUBLE) :: row(11), deltat_from_dnsin, time0, selectime, curr_dt, cols(9)
  real(C_DOUBLE), allocatable, target :: gain(:,:,:)
  complex(C_DOUBLE_COMPLEX), allocatable :: F(:,:,:,:)
  integer(C_INT32_T) :: src(3)
  integer(C_INT32_T) :: sizes(5)
  integer(C_SIZE_T) :: istep, ifield
  integer :: io, iy
  logical :: found, rtd_exists, threshold_reached, prev_was_close
  real :: timei, timee
  character(len=40) :: istring

  CALL read_dnsin()
  CALL read_env()
  deltat_from_dnsin = deltat
  prm = params()
  CALL chan_check(chan_create(prm, ctx), "chan_create")
  CALL chan_check(chan_sizes(ctx, sizes), "chan_sizes")
  ALLOCATE(V(-1:ny+1,-nz:nz,0:nx,1:3), y(-1:ny+1))
  V = 0
  CALL chan_check(chan_setup_host(prm, sizes, c_loc(y), C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR), &
                  "chan_setup_host")

  ! Runtimedata (init_memory, dnsdata.f90:159-176)
  INQUIRE(FILE="Runtimedata", EXIST=rtd_exists)
  IF (time_from_restart .AND. rtd_exists) THEN
    WRITE(*,*) 'Found existing Runtimedata...'
    OPEN(UNIT=101, FILE='Runtimedata', ACTION='readwrite', POSITION='rewind')
  ELSE
    IF (time_from_restart) THEN
      WRITE(*,*) 'Runtimedata not found...'
      CALL wanna_continue()
    END IF
    WRITE(*,*) 'Creating new Runtimedata.'
    OPEN(UNIT=101, FILE='Runtimedata', ACTION='write', STATUS='replace')
  END IF

  ! initial field: Dati.cart.out, or laminar 1.5 y(2-y) when it is missing
  CALL read_restart("Dati.cart.out", V, found)
  IF (.NOT. found) THEN
    WRITE(*,*) "Generating initial field..."
    DO iy = -1, ny+1
      V(iy,0,0,1) = 3*0.5*y(iy)*(2-y(iy))
    END DO
  END IF

  ! position in Runtimedata (get_record) or restart the clock
  IF (time_from_restart .AND. rtd_exists) THEN
    threshold_reached = .FALSE.
    io = 0
    DO WHILE (.NOT. threshold_reached .AND. io >= 0)
      READ(101, *, IOSTAT=io) selectime, cols, curr_dt
      threshold_reached = ABS(selectime - time) < (0.5*curr_dt)
    END DO
    IF (io >= 0) THEN
      BACKSPACE(101)
      WRITE(*,*) 'In Runtimedata: starting from time', selectime
      deltat = curr_dt
    ELSE
      WRITE(*,*) 'No instant of time matching restart file has been found in Runtimedata.'
      CALL wanna_continue()
      WRITE(*,*) 'Skipping one line and appending.'
      WRITE(101,*) ''
    END IF
  ELSE IF (.NOT. time_from_restart) THEN
    CALL read_dnsin()
  END IF
  ifield = FLOOR((time+0.5*deltat)/dt_field)
  time0 = time
  IF (cflmax == 0) deltat = deltat_from_dnsin

  WRITE(*,*) " "
  WRITE(*,*) "!====================================================!"
  WRITE(*,*) "!       D   N   S    (MI355X time step)              !"
  WRITE(*,*) "!====================================================!"
  WRITE(*,"(A,I5,A,I5,A,I5)") "   nx =", nx, "   ny =", ny, "   nz =", nz
  WRITE(*,"(A,I5,A,I5)") "   nxd =", sizes(4), "  nzd =", sizes(5)
  IF (precision_sel == CHAN_SPLIT) WRITE(*,"(A)") "   FP32/FP64 split step"
  WRITE(*,*) " "
  IF (LEN_TRIM(bf_name) > 0) THEN
    ALLOCATE(gain(0:2*nz+1, -1:ny+1, 0:2), F(-1:ny+1,-nz:nz,0:nx,1:3))
    CALL body_force_table(y, src, gain)
    CALL chan_check(chan_set_body_force_feedback(ctx, src, gain), "chan_set_body_force_feedback")
    CALL chan_check(chan_keep_body_force(ctx, 1), "chan_keep_body_force")
  END IF

  CALL chan_check(chan_set_field(ctx, V), "chan_set_field")
  CALL chan_check(chan_get_scalars(ctx, sc), "chan_get_scalars")
  sc%time = time
  sc%deltat = deltat
  CALL chan_check(chan_set_scalars(ctx, sc), "chan_set_scalars")

  ! initial CFL, flow rate and CPI (channel.f90:95-115), first Runtimedata line
  CALL chan_check(chan_initial(ctx), "chan_initial")
  istep = 0
  prev_was_close = .FALSE.
  CALL chan_check(chan_outstats(ctx, row), "chan_outstats")
  CALL report(row)

  DO WHILE ((time < t_max-deltat/2.0) .AND. (istep < nstep))
    CALL CPU_TIME(timei)
    istep = istep+1
    CALL chan_check(chan_step(ctx, row), "chan_step")
    CALL report(row)
    CALL CPU_TIME(timee)
    WRITE(*,*) timee-timei
  END DO
  WRITE(*,*) "End of time/iterations loop: writing restart file at time ", time
  CALL chan_check(chan_get_field(ctx, V), "chan_get_field")
  CALL save_restart("Dati.cart.out", V, time)
  CLOSE(101)
  CALL chan_check(chan_destroy(ctx), "chan_destroy")

CONTAINS

  FUNCTION params() result(p)
    type(chan_params) :: p
    p%nx = nx; p%ny = ny; p%nz = nz; p%precision = precision_sel
    p%alfa0 = alfa0; p%beta0 = beta0; p%ni = ni; p%a = a; p%ymin = ymin; p%ymax = ymax
    p%cpi = MERGE(1, 0, CPI); p%cpi_type = CPI_type; p%gamma = gamma
    p%meanpx = meanpx; p%meanpz = meanpz; p%meanflowx = meanflowx; p%meanflowz = meanflowz
    p%u0 = u0; p%uN = uN; p%deltat = deltat; p%cflmax = cflmax; p%time = time
  END FUNCTION params

  ! outstats' output part (dnsdata.f90:875-919): screen and Runtimedata line,
  ! restart and snapshot files on their cadence
  SUBROUTINE report(r)
    real(C_DOUBLE), intent(in) :: r(11)
    character(len=60) :: fn
    time = r(1); deltat = r(11)
    WRITE(*,"(F10.4,3X,4(F11.6,3X),4(F9.4,3X),2(F9.6,3X))") r
    ! one record per line with 17 significant digits (list-directed output
    ! would wrap the 11 values over several lines with flang)
    WRITE(101,"(11(1X,ES24.16E3))") r
    FLUSH(101)
    IF (dt_save > 0) THEN
      IF (((FLOOR((time+0.5*deltat)/dt_save) > FLOOR((time-0.5*deltat)/dt_save)) .AND. (istep > 1))) THEN
        WRITE(*,*) "Writing Dati.cart.out at time ", time
        CALL chan_check(chan_get_field(ctx, V), "chan_get_field")
        CALL save_restart("Dati.cart.out", V, time)
      END IF
    END IF
    IF (time+deltat >= (ifield+1)*dt_field) THEN
      IF (prev_was_close .OR. ((FLOOR((time+0.5*deltat)/dt_field) > FLOOR((time-0.5*deltat)/dt_field)) &
          .AND. (time > time0))) THEN
        ifield = ifield+1
        WRITE(istring,*) ifield
        fn = "Dati.cart."//TRIM(ADJUSTL(istring))//".out"
        WRITE(*,*) "Writing "//TRIM(fn)//" at time ", time
        CALL chan_check(chan_get_field(ctx, V), "chan_get_field")
        CALL save_restart(fn, V, time)
        IF (LEN_TRIM(bf_name) > 0) THEN
          ! the force of the last substep, as F holds it after a step
          fn = "Force.cart."//TRIM(ADJUSTL(istring))//".out"
          WRITE(*,*) "Writing "//TRIM(fn)//" at time ", time
          CALL chan_check(chan_get_applied_body_force(ctx, F), "chan_get_applied_body_force")
          CALL save_restart(fn, F, time)
        END IF
        prev_was_close = .FALSE.
      ELSE
        prev_was_close = .TRUE.
      END IF
    END IF
  END SUBROUTINE report

END PROGRAM channel
