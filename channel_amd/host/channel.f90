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
MODULE host_io
  USE, intrinsic :: iso_c_binding
  USE chanlib
  IMPLICIT NONE
  integer(C_INT) :: nx, ny, nz, npy, CPI_type
  real(C_DOUBLE) :: alfa0, beta0, ni, a, ymin, ymax, deltat, cflmax, time, dt_field, dt_save, t_max, gamma
  real(C_DOUBLE) :: u0, uN, meanpx, meanpz, meanflowx, meanflowz
  logical :: CPI, time_from_restart
  integer(C_SIZE_T) :: nstep
CONTAINS

  SUBROUTINE read_dnsin()
    OPEN(15, file='dns.in', status='old', action='read')
    READ(15, *) nx, ny, nz
    READ(15, *) alfa0, beta0
    READ(15, *) ni
    ni = 1/ni
    READ(15, *) a, ymin, ymax
    READ(15, *) CPI, CPI_type, gamma
    READ(15, *) meanpx, meanpz
    READ(15, *) meanflowx, meanflowz
    READ(15, *) u0, uN
    READ(15, *) deltat, cflmax, time
    READ(15, *) dt_field, dt_save, t_max, time_from_restart
    READ(15, *) nstep
    READ(15, *) npy
    CLOSE(15)
  END SUBROUTINE read_dnsin

  ! restart file: 3 int32 + 7 float64 header, then V(-1:ny+1,-nz:nz,0:nx,1:3)
  SUBROUTINE read_restart(fname, V, found)
    character(len=*), intent(in) :: fname
    complex(C_DOUBLE_COMPLEX), intent(inout) :: V(-1:ny+1,-nz:nz,0:nx,1:3)
    logical, intent(out) :: found
    integer(C_INT) :: r_nx, r_ny, r_nz, io
    real(C_DOUBLE) :: r_alfa0, r_beta0, r_ni, r_a, r_ymin, r_ymax, r_time
    OPEN(UNIT=100, FILE=TRIM(fname), access="stream", status="old", action="read", iostat=io)
    found = (io == 0)
    IF (.NOT. found) RETURN
    WRITE(*,*) "Reading from file "//TRIM(fname)
    READ(100, POS=1) r_nx, r_ny, r_nz, r_alfa0, r_beta0, r_ni, r_a, r_ymin, r_ymax, r_time
    IF (r_nx /= nx .OR. r_ny /= ny .OR. r_nz /= nz .OR. r_alfa0 /= alfa0 .OR. r_beta0 /= beta0 .OR. &
        r_ni /= ni .OR. r_a /= a .OR. r_ymin /= ymin .OR. r_ymax /= ymax) THEN
      WRITE(*,*) "ERROR: mismatch in metadata between restart file and dns.in. Stopping."
      WRITE(*,*) "From .out file:"
      WRITE(*,*) r_nx, r_ny, r_nz, r_alfa0, r_beta0, r_ni, r_a, r_ymin, r_ymax
      WRITE(*,*) "From dns.in:"
      WRITE(*,*) nx, ny, nz, alfa0, beta0, ni, a, ymin, ymax
      STOP
    END IF
    READ(100, POS=69) V
    CLOSE(100)
    time = r_time
  END SUBROUTINE read_restart

  SUBROUTINE save_restart(fname, V, t)
    character(len=*), intent(in) :: fname
    complex(C_DOUBLE_COMPLEX), intent(in) :: V(-1:ny+1,-nz:nz,0:nx,1:3)
    real(C_DOUBLE), intent(in) :: t
    OPEN(UNIT=100, FILE=TRIM(fname), access="stream", status="replace", action="write")
    WRITE(100, POS=1) nx, ny, nz, alfa0, beta0, ni, a, ymin, ymax, t
    WRITE(100, POS=69) V
    CLOSE(100)
  END SUBROUTINE save_restart

  SUBROUTINE wanna_continue()
    character :: yn
    WRITE(*,"(a,$)") "Do you want to continue? (y/n) "
    READ(*,*) yn
    IF (yn == 'n') STOP
  END SUBROUTINE wanna_continue

END MODULE host_io


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
  integer(C_INT32_T) :: sizes(5)
  integer(C_SIZE_T) :: istep, ifield
  integer :: io, iy
  logical :: found, rtd_exists, threshold_reached, prev_was_close
  real :: timei, timee
  character(len=40) :: istring

  CALL read_dnsin()
  deltat_from_dnsin = deltat
  prm = params()
  CALL chan_check(chan_create(prm, ctx), "chan_create")
  CALL chan_check(chan_sizes(ctx, sizes), "chan_sizes")
  ALLOCATE(V(-1:ny+1,-nz:nz,0:nx,1:3), y(-1:ny+1))
  V = 0

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
    CALL chan_check(chan_setup_host(prm, sizes, c_loc(y), C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR), &
                    "chan_setup_host")
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
  WRITE(*,*) " "

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
    p%nx = nx; p%ny = ny; p%nz = nz; p%precision = CHAN_FP64
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
        prev_was_close = .FALSE.
      ELSE
        prev_was_close = .TRUE.
      END IF
    END IF
  END SUBROUTINE report

END PROGRAM channel
