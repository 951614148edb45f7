/*
 * chan_oracle.h -- CPU restatement of channel's FP64 RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (channel_amd/, include/) never links or calls it.
 *
 * Every routine restates the reference algorithm in plain C with the
 * reference's evaluation order (left-to-right sums, real*complex products,
 * component-wise division by reals).  Fields use the reference's own layout:
 * V(iy,iz,ix,ic) Fortran order, iy in [-1,ny+1], iz in [-nz,nz], ix in [0,nx],
 * ic in 1..3, complex128 interleaved (README.md:196-216, dnsdata.f90:132).
 */
#ifndef CHAN_ORACLE_H
#define CHAN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int32_t nx, ny, nz;          /* dns.in line 1 */
  double alfa0, beta0;         /* dns.in line 2 */
  double ni;                   /* viscosity nu = 1/Re (dnsdata.f90:115 inverts) */
  double a, ymin, ymax;        /* dns.in line 4 */
  int32_t cpi, cpi_type;       /* dns.in line 5 */
  double gamma;
  double meanpx, meanpz;       /* dns.in line 6 */
  double meanflowx, meanflowz; /* dns.in line 7 */
  double u0, uN;               /* dns.in line 8 */
  double deltat, cflmax, time; /* dns.in line 9 */
} orc_params;

typedef struct orc orc_t;

orc_t *orc_create(const orc_params *p);
void orc_destroy(orc_t *o);
void orc_set_threads(int n);
int orc_get_threads(void);
void orc_sizes(const orc_t *o, int32_t out[5]); /* nx ny nz nxd nzd */

void orc_set_field(orc_t *o, const double *V);
void orc_get_field(const orc_t *o, double *V);

/* scalars: [time, deltat, meanpx, meanpz, corrpx, corrpz, fr1, fr2, fr3, cfl] */
void orc_get_scalars(const orc_t *o, double s[10]);
void orc_set_scalars(orc_t *o, const double s[10]);

/* set-up products (dnsdata.f90:148-158, 241-308) */
void orc_get_grid(const orc_t *o, double *y /* ny+3 */, double *dy /* ny-1 */);
void orc_get_der(const orc_t *o, double *der /* (ny-1)*4*5: d0,d1,d2,d4 */);
void orc_get_bnd(const orc_t *o, double *bnd /* 10*5 */);
void orc_get_bc(const orc_t *o, double *bc /* 8*5 */);
void orc_get_d0mat(const orc_t *o, double *d0mat /* (ny+1)*5 factored */);

/* the step and its pieces */
void orc_convolutions(orc_t *o, int iy, int compute_cfl, double *out6);
void orc_buildrhs(orc_t *o, const double ode[3], int compute_cfl);
void orc_linsolve(orc_t *o, double lambda);
void orc_initial(orc_t *o);
void orc_outstats(orc_t *o, double row[11]);
void orc_step(orc_t *o, double row[11]);
void orc_steps(orc_t *o, int n, double *rows);
void orc_time_sample(orc_t *o, int nplanes, int nix, double out[3]);

/* body force (the bodyforce option, dnsdata.f90:51-54, 616-628, 643-659).
 * orc_set_body_force copies a static F in the layout of V (NULL removes it
 * unless a module is set).  orc_set_body_force_module selects a module of
 * body_forces/ that runs at the start of every orc_buildrhs, as the
 * reference calls set_body_force before every buildrhs: kind 0 none,
 * 1 am_f1 and 2 am_butterfly with par = {amp, lambdaz_f, -}, 3 coriolis with
 * par = {omega2, kz_cutoff, y_threshold_bot}.  A module writes over F where
 * its mask holds.  orc_apply_body_force runs the module once;
 * orc_body_force_walls applies buildrhs's wall rows. */
void orc_set_body_force(orc_t *o, const double *F);
void orc_set_body_force_module(orc_t *o, int kind, const double *par);
void orc_apply_body_force(orc_t *o);
void orc_body_force_walls(orc_t *o);
void orc_get_body_force(const orc_t *o, double *F);

/* primitives exposed for pinning tests */
void orc_fft(int n, int sign, double *inout);
void orc_c2r(int m, const double *X, double *x);
void orc_r2c(int m, const double *x, double *X);
void orc_ludecomp5(double *A /* 5x5 row-major */);
void orc_lludiv5(const double *A, const double *b, double *x);
void orc_lu5decomp(double *A /* rows x 5 */, int rows, int first, int last);
void orc_lu5div1(const double *A, int rows, double *x /* complex, rows+2 */);
void orc_lu5div2(const double *A, int rows, double *b /* complex, rows+2 */);
double orc_yintegr(const orc_t *o, const double *f /* ny+3 */);
void orc_complexderiv(const orc_t *o, const double *f0, double *f1);

#ifdef __cplusplus
}
#endif
#endif
