!
! host_io.f90 -- the Fortran host's input and output: dns.in, the restart
! files, the run-time selections and the body-force module tables.
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
  character(len=32) :: bf_name = ''
  integer(C_INT32_T) :: precision_sel = CHAN_FP64
  ! am_pardec.inc:1 (the parameters of am_f1 and am_butterfly)
  real(C_DOUBLE), parameter :: amp = 1000, lambdaz_f = 500
  ! the reference's PI, 2.3 and 1E10 are default-real literals
  ! (dnsdata.f90:26, am_f1.inc:18-22): their single-precision values
  real(C_DOUBLE), parameter :: PI_R = REAL(REAL(3.14159265358979324_C_DOUBLE, C_FLOAT), C_DOUBLE)
  real(C_DOUBLE), parameter :: R23 = REAL(REAL(2.3_C_DOUBLE, C_FLOAT), C_DOUBLE)
  real(C_DOUBLE), parameter :: R1E10 = REAL(REAL(1E10_C_DOUBLE, C_FLOAT), C_DOUBLE)
CONTAINS

  SUBROUTINE read_env()
    character(len=32) :: v
    integer :: n, st
    CALL GET_ENVIRONMENT_VARIABLE("CHANNEL_PRECISION", v, n, st)
    IF (st == 0 .AND. n > 0) THEN
      SELECT CASE (TRIM(v))
      CASE ('fp64')
        precision_sel = CHAN_FP64
      CASE ('split')
        precision_sel = CHAN_SPLIT
      CASE DEFAULT
        WRITE(*,"(A)") "ERROR: CHANNEL_PRECISION must be fp64 or split, not "//TRIM(v)
        ERROR STOP 1
      END SELECT
    END IF
    CALL GET_ENVIRONMENT_VARIABLE("CHANNEL_BODYFORCE", v, n, st)
    IF (st == 0 .AND. n > 0) THEN
      SELECT CASE (TRIM(v))
      CASE ('none')
        bf_name = ''
      CASE ('am_f1', 'am_butterfly', 'coriolis')
        bf_name = TRIM(v)
      CASE DEFAULT
        WRITE(*,"(A)") "ERROR: CHANNEL_BODYFORCE must be am_f1, am_butterfly, coriolis or none, not "//TRIM(v)
        ERROR STOP 1
      END SELECT
    END IF
  END SUBROUTINE read_env

  ! Feedback table of the selected module for chan_set_body_force_feedback:
  ! force component d (0..2) is gain(col,iy,d) times velocity component
  ! src(d+1), col = iz+nz for every mode but (ix,iz) = (0,0), which reads
  ! col = 2nz+1.  The modules (body_forces/*/*.inc) write F = g V on a mask
  ! that depends on iy and iz only.
  SUBROUTINE body_force_table(y, src, gain)
    real(C_DOUBLE), intent(in) :: y(-1:ny+1)
    integer(C_INT32_T), intent(out) :: src(3)
    real(C_DOUBLE), intent(out) :: gain(0:2*nz+1, -1:ny+1, 0:2)
    real(C_DOUBLE) :: yp, lzp, omega2, kz_cutoff, y_thr
    integer :: iy, iz, izf, thr
    logical :: on
    gain = 0
    SELECT CASE (TRIM(bf_name))
    CASE ('am_f1', 'am_butterfly')
      ! F = -amp V on the mask; the mean (0,0) is outside it, so column
      ! 2nz+1 stays zero (am_f1.inc:11-27, am_butterfly.inc:11-27)
      WRITE(*,*) "Using bodyforce, "//TRIM(bf_name)
      src = [0, 1, 2]
      izf = NINT((2*PI_R/lambdaz_f)/(beta0/1000))
      DO iy = -1, ny+1
        ! wall distance in units of 1/1000 of the half height
        yp = 1000*MERGE(ymax-y(iy), y(iy), y(iy) > 1)
        DO iz = -nz, nz
          IF (TRIM(bf_name) == 'am_f1') THEN
            on = .FALSE.
            IF (ABS(iz) <= izf) THEN
              lzp = R1E10
              IF (iz /= 0) lzp = 1000*(2*PI_R/(beta0*ABS(iz)))
              on = lzp > R23*yp**2
            END IF
          ELSE
            on = (ABS(iz) <= izf .AND. yp <= 60) .OR. (ABS(iz) > izf .AND. yp > 60)
          END IF
          IF (on) gain(iz+nz, iy, :) = -amp
        END DO
      END DO
    CASE ('coriolis')
      ! spanwise rotation (coriolis.inc:4-39): F_u = -omega2 v, F_v = omega2 u
      ! on |iz| <= min(nz, floor(kz_cutoff/beta0)) within y_threshold of a
      ! wall, every ix, the mean included
      OPEN(15, file='coriolis.in', status='old', action='read')
      READ(15, *) omega2
      READ(15, *) kz_cutoff
      READ(15, *) y_thr
      CLOSE(15)
      thr = MIN(nz, FLOOR(kz_cutoff/beta0))
      WRITE(*,*) "Using bodyforce, Coriolis force"
      WRITE(*,*) "Ro(tation number)", omega2/2
      WRITE(*,*) "kz_cutoff, iz_cutoff", kz_cutoff, thr
      WRITE(*,*) "y_threshold_bottom", y_thr
      src = [1, 0, -1]
      DO iy = -1, ny+1
        IF (y(iy) <= y_thr .OR. y(iy) >= ymax-y_thr) THEN
          DO iz = -thr, thr
            gain(iz+nz, iy, 0) = -omega2
            gain(iz+nz, iy, 1) = omega2
          END DO
        END IF
      END DO
      IF (thr >= 0) gain(2*nz+1, :, :) = gain(nz, :, :)
    END SELECT
    WRITE(*,*) ""
  END SUBROUTINE body_force_table

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
