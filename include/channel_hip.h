/*
 * channel_hip.h -- C-ABI of the MI355X-native channel DNS time step.
 *
 * The reference has no plugin API: its per-step seam is the module-level
 * subroutine interface of MODULE dnsdata, called from the time loop of
 * PROGRAM channel.  Each entry point below replaces one of those routines and
 * keeps its argument meaning and fail-stop behaviour; the library owns the
 * device-resident state (the role of the module globals, dnsdata.f90:25-90).
 *
 * Conventions
 *  - every function returns int: 0 = success, > 0 = HIP error code,
 *    < 0 = invalid argument / state (see CHAN_E_*); chan_last_error() gives
 *    the message.  A Fortran host prints it and STOPs, as the reference does.
 *  - fields cross the boundary in the reference's own layout and type:
 *    complex128 V(-1:ny+1, -nz:nz, 0:nx, 1:3), Fortran order, as
 *    interleaved doubles (dnsdata.f90:132, README.md:196-216).
 *  - one host thread per handle; calls are not thread-safe.  Calls that
 *    return device-computed scalars synchronise the library's stream.
 *  - no torch types and no ownership transfer: pointers are read or written
 *    during the call only.
 */
#ifndef CHANNEL_HIP_H
#define CHANNEL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHAN_ABI_VERSION 1

/* precision modes */
#define CHAN_FP64 0  /* plain FP64 step: the drop-in */
#define CHAN_SPLIT 1 /* FP32 transforms/products/stencils, FP64 solves and mean mode */

/* status codes below zero */
#define CHAN_E_ARG (-1)     /* invalid argument */
#define CHAN_E_SIZE (-2)    /* unsupported transform size (fftFIT sizes only) */
#define CHAN_E_NOTFINITE (-3) /* CFL reduction is not finite (NaN/Inf in the field) */
#define CHAN_E_CPITYPE (-4) /* CPI_type not 0 or 1 (linsolve_blocking.inc:93-95) */
#define CHAN_E_STATE (-5)   /* call order violated (e.g. step before set_field) */
#define CHAN_E_NODEV (-6)   /* no HIP device */

/* The run-time part of dns.in (dnsdata.f90:110-122) with ni already
 * inverted to the viscosity 1/Re (dnsdata.f90:115), plus the precision mode. */
typedef struct chan_params {
  int32_t nx, ny, nz;
  int32_t precision; /* CHAN_FP64 or CHAN_SPLIT */
  double alfa0, beta0;
  double ni; /* viscosity */
  double a, ymin, ymax;
  int32_t cpi, cpi_type;
  double gamma;
  double meanpx, meanpz;
  double meanflowx, meanflowz;
  double u0, uN;
  double deltat, cflmax, time;
} chan_params;

typedef struct chan_ctx chan_ctx;

/* Scalars of the step (module globals of dnsdata.f90:28-32,68,74,76). */
typedef struct chan_scalars {
  double time, deltat;
  double meanpx, meanpz;
  double corrpx, corrpz;
  double fr[3];
  double cfl; /* running CFL maximum since the last outstats */
} chan_scalars;

/* Library version and the message of the last failing call. */
int chan_abi_version(void);
const char *chan_last_error(void);

/* read_dnsin derived sizes + init_memory + setup_derivatives +
 * setup_boundary_conditions (dnsdata.f90:110-125,129-158,241-308): builds the
 * y grid, compact stencils, boundary rows and the factorised D0 band on the
 * host in FP64, allocates the device state on the current HIP device. */
int chan_create(const chan_params *p, chan_ctx **out);
/* free_memory / free_fft (dnsdata.f90:222-229, ffts.f90:110-115) */
int chan_destroy(chan_ctx *c);

/* sizes[5] = nx, ny, nz, nxd, nzd (nxd, nzd from fftFIT, ffts.f90:78-86) */
int chan_sizes(const chan_ctx *c, int32_t sizes[5]);

/* Set-up products for inspection: y(-1:ny+1) [ny+3]; der(1:ny-1)%d0,d1,d2,d4
 * as [(ny-1)][4][5]; boundary stencils d040,d140,d240,d14m1,d24m1,d04n,d14n,
 * d24n,d14np1,d24np1 as [10][5] (dnsdata.f90:241-276). */
int chan_get_setup(const chan_ctx *c, double *y, double *der, double *bnd);

/* Host-only variant of the set-up query that needs no device and no
 * context: sizes[5] as chan_sizes, then y, der, bnd as chan_get_setup, the
 * eight boundary-condition rows v0bc,v0m1bc,eta0bc,eta0m1bc,vnbc,vnp1bc,
 * etanbc,etanp1bc as bc[8][5] (dnsdata.f90:290-308) and the UL-factorised
 * D0 band as d0fac[(ny+1)][5] (dnsdata.f90:277-285).  Any output may be NULL. */
int chan_setup_host(const chan_params *p, int32_t sizes[5], double *y, double *der, double *bnd, double *bc,
                    double *d0fac);

/* Field I/O in the reference layout (the payload of read_restart_file /
 * save_restart_file, dnsdata.f90:677-720, 821-848; the library does no file
 * I/O).  V has (ny+3)*(2nz+1)*(nx+1)*3 complex128 = 2x doubles. */
int chan_set_field(chan_ctx *c, const double *V);
int chan_get_field(chan_ctx *c, double *V);

/* The right-hand sides left by chan_buildrhs, i.e. what the reference holds
 * in V(1:ny-1, :, :, 1:2) between buildrhs and linsolve (dnsdata.f90:667-671):
 * R(1:ny-1, -nz:nz, 0:nx, 1:2) complex128, Fortran order. */
int chan_get_rhs(chan_ctx *c, double *R);

int chan_get_scalars(chan_ctx *c, chan_scalars *s);
int chan_set_scalars(chan_ctx *c, const chan_scalars *s);

/* channel.f90:95-115: CFL pass over the interior planes (convolutions with
 * compute_cfl), initial flow rates fr(1:2) and the CPI pressure gradient. */
int chan_initial(chan_ctx *c);

/* buildrhs(ODE, compute_cfl) (dnsdata.f90:611-673): products of all planes,
 * RHS stencils and the CN-RK3 time scheme; leaves the right-hand sides of
 * eta and D2v in the state in place of u, v for linsolve. */
int chan_buildrhs(chan_ctx *c, const double ode[3], int32_t compute_cfl);

/* linsolve(lambda) (linsolve_blocking.inc:3-107): per-(kx,kz) pentadiagonal
 * v and eta solves with wall BCs, dv/dy, u and w, and the (0,0) mean-flow
 * mode with flow-rate correction / CPI. */
int chan_linsolve(chan_ctx *c, double lambda);

/* outstats (dnsdata.f90:853-881) minus its file I/O: takes the CFL maximum,
 * sets deltat = cflmax/CFL when cflmax > 0, resets the CFL maximum and fills
 * the 11 Runtimedata columns (time, dudy bottom, -dudy top, dwdy bottom,
 * -dwdy top, fr_x, dp/dx, fr_z, dp/dz, CFL, deltat). */
int chan_outstats(chan_ctx *c, double row[11]);

/* One iteration of the time loop (channel.f90:122-167): wall velocities
 * u0/uN, three substeps of buildrhs+linsolve with time advance, outstats.
 * n iterations write n rows. */
int chan_step(chan_ctx *c, double row[11]);
int chan_steps(chan_ctx *c, int32_t n, double *rows);

/* Wall-normal statistics of the current field for rows iy = 0..ny, the
 * spectral sums of postpro/history/var_history.f90:36-87 (weight 1 at kx = 0,
 * 2 otherwise), out = [7][ny+1]:
 *   rows 0-2  plane means Re V(iy,0,0,c), c = u, v, w;
 *   rows 3-5  variances sum f |V(iy,iz,ix,c) - d m(iy,c)|^2, d = 1 only for
 *             the (0,0) mode;
 *   row  6    Reynolds shear stress sum f Re[(V_u - d m_u) conj(V_v - d m_v)].
 * m = avg[3][ny+1] (a time average, as the reference subtracts) or, when avg
 * is NULL, the field's own plane mean (instantaneous fluctuations). */
int chan_stats(chan_ctx *c, const double *avg, double *out);

/* Body force (the reference's compile-time `bodyforce` option,
 * dnsdata.f90:51-54, 133, 616-628, 643-659).  buildrhs adds the force
 * F(-1:ny+1,-nz:nz,0:nx,1:3) to the explicit terms: -k2 D0(F_v) - i kx D1(F_u)
 * - i kz D1(F_w) in the D2v equation, i kz D0(F_u) - i kx D0(F_w) in the eta
 * equation and (D0 Re F_u, D0 Re F_w) in the mean-flow mode.  As in the
 * reference, F is zero on the wall rows 0 and ny and its ghost rows -1, ny+1
 * make the compact D4 of rows 1 and ny-1 vanish, whatever the caller supplies
 * there.  The force is the sum of two optional parts:
 *
 * chan_set_body_force: a static field F0 in the layout of V (NULL removes it);
 * the role of a config_body_force that fills F once.
 *
 * chan_set_body_force_feedback: F_d += gain[d][iy+1][col] * V_src[d], formed
 * from the field at the start of every buildrhs -- the reference's
 * set_body_force called before each substep (channel.f90:129-157).  src[d] in
 * {-1 (none), 0 (u), 1 (v), 2 (w)}; gain is [3][ny+3][2nz+2] doubles with
 * col = iz+nz for every mode except (ix,iz) = (0,0), which reads col = 2nz+1.
 * This covers body_forces/am_f1 and am_butterfly (src = {0,1,2}, gain -amp on
 * their masks) and body_forces/coriolis (src = {1,0,-1}, gains -omega2 and
 * +omega2); channel_amd.body_force builds those tables.  src or gain NULL, or
 * every src -1, removes it.
 *
 * chan_get_body_force: the force that the next buildrhs applies to the
 * resident field, wall and ghost rows included, in the layout of V (zero
 * when no force is set); the payload of the reference's Force.cart.N.out
 * (dnsdata.f90:904-906) for a static force.
 *
 * chan_keep_body_force(c, 1): from then on every buildrhs that applies a
 * force also stores it (chan_step stores only its third substep's, the one
 * that stays); chan_keep_body_force(c, 0) stops and frees the copy (one
 * field, 3*(ny+3)*(2nz+1)*(nx+1) complex FP64).  chan_get_applied_body_force
 * returns that copy in the layout of V: the reference's F after a time step
 * -- set_body_force applied to the field of the third substep, then the wall
 * rows of buildrhs (dnsdata.f90:616-629) -- which is what Force.cart.N.out
 * holds (dnsdata.f90:904-906) when a module sets F from V.  Zero when that
 * buildrhs applied no force; CHAN_E_STATE when keeping is off or no buildrhs
 * has run since it was turned on. */
int chan_set_body_force(chan_ctx *c, const double *F);
int chan_set_body_force_feedback(chan_ctx *c, const int32_t src[3], const double *gain);
int chan_get_body_force(chan_ctx *c, double *F);
int chan_keep_body_force(chan_ctx *c, int32_t on);
int chan_get_applied_body_force(chan_ctx *c, double *F);

/* Per-kernel timing with HIP events recorded on the library's stream around
 * every launch (off by default).  chan_kernel_times fills up to cap entries of
 * names (NUL-terminated, 32 chars each), launch counts and total ms, and
 * returns the number of kernels in *n. */
int chan_timing(chan_ctx *c, int32_t enable);
int chan_kernel_times(chan_ctx *c, int32_t cap, char *names, int64_t *counts, double *total_ms, int32_t *n);
/* Synchronises the library's stream. */
int chan_sync(chan_ctx *c);

/* Diagnostic: batched c2c FFT of length n (sign -1 forward, +1 backward,
 * unnormalised, FFTW conventions) on `batch` contiguous complex128 rows using
 * the library's device transform. */
int chan_fft_selftest(int32_t n, int32_t sign, int32_t batch, double *inout);

/* Diagnostic: batched out-of-place c2c FFT of length n in {384, 768, 1536}
 * with one of the device transform's build variants (variant & 15: 0-7 bit 0
 * complex LDS exchange, bit 1 padded exchange, bit 2 radix-4/3 plan; 8-11
 * complex exchange padded every 16, 8, 4 elements or XOR-swizzled; variant >> 4
 * extra in-register repetitions of the transform per load/store in the timed
 * launches), repeated `reps` times; inout holds the rows on entry and their
 * single transform on return, *ms the average duration of one timed launch. */
int chan_fft_bench(int32_t n, int32_t sign, int32_t batch, int32_t variant, int32_t reps, double *inout, double *ms);

/* Diagnostic: the four-step 768-point transform of the x pass (one
 * wavefront per row; 12 values per lane, permlane-swap transpose and two LDS
 * exchanges) on `batch` rows of 768 complex128 values in natural order:
 * sign +1 the unnormalised backward transform, -1 the forward one.  inout
 * holds the rows on entry and their transforms on return; *ms the average
 * duration of `reps` timed launches, each of which first runs inner-1
 * backward/forward pairs (1 <= inner <= 8). */
int chan_fft768_fs_test(int32_t sign, int32_t batch, int32_t inner, int32_t reps, double *inout, double *ms);

/* Diagnostic: a plain device-to-device copy of `bytes` (multiple of 32) done
 * `reps` times by the library's own streaming copy kernel; *ms receives the
 * average duration.  Used to calibrate hardware byte counters. */
int chan_copy_selftest(int64_t bytes, int32_t reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
