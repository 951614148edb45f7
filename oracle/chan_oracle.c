/*
 * chan_oracle.c -- CPU restatement of channel's FP64 RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it;
 * the product never does.
 *
 * Each function cites the reference routine (path:line under the reference
 * tree) it restates.  Evaluation order follows the Fortran source: sums left
 * to right, real*complex as component products, complex/real component-wise.
 *
 * Pins (what checks each function, see tests/test_oracle_*.py):
 *   setup_derivatives  -- closed-form uniform-grid stencils; exactness on
 *                         polynomials (D1 p = D0 p', D2 p = D0 p'', D4 p = D0 p'''')
 *   boundary stencils  -- exact one-sided derivatives of quartics
 *   LU5 / LeftLU5div   -- residual against numpy dense solve of the same band
 *   ffts               -- brute-force O(N^2) DFT and numpy.fft
 *   convolutions       -- exact truncated convolution of Fourier series
 *   yintegr            -- exact for quadratics on the stretched grid
 *   buildrhs+linsolve  -- laminar Poiseuille fixed point; Orr-Sommerfeld
 *                         growth rate of the leading mode (independent
 *                         Chebyshev solver in the tests)
 *   time loop/outstats -- Runtimedata columns of a laminar run (closed form)
 *   body force         -- laminar Poiseuille kept by a uniform force instead
 *                         of dp/dx (u and w); steady forced Stokes flow of
 *                         oblique and 2-D modes against a Chebyshev solution
 *                         of the forced v-eta equations; damping of an
 *                         Orr-Sommerfeld wave by exactly amp under am_f1;
 *                         module masks against an independent table builder
 */
#include "chan_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ---------------------------------------------------------------- FFT -- */
/* Mixed-radix recursive decimation-in-time DFT.  Plays the part of FFTW's
 * plans in ffts.f90:70-75: unnormalised, sign -1 forward (FFT/HFT), sign +1
 * backward (IFT/RFT). */
#define MAXFAC 40
typedef struct {
  int n, nf, fac[MAXFAC];
  cplx *w; /* w[j] = exp(-2 pi i j / n) */
} fftplan;

static void plan_init(fftplan *P, int n) {
  P->n = n;
  P->nf = 0;
  int m = n;
  static const int pref[] = {4, 2, 3, 5};
  for (int q = 0; q < 4; q++)
    while (m % pref[q] == 0 && (pref[q] != 4 || m % 4 == 0)) {
      P->fac[P->nf++] = pref[q];
      m /= pref[q];
    }
  for (int f = 7; m > 1; f += 2)
    while (m % f == 0) {
      P->fac[P->nf++] = f;
      m /= f;
    }
  P->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int j = 0; j < n; j++) {
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)n;
    P->w[j] = CMPLX((double)cosl(ang), (double)sinl(ang));
  }
}
static void plan_free(fftplan *P) { free(P->w); P->w = NULL; }

static inline cplx tw(const fftplan *P, long e, int sign) {
  cplx v = P->w[e % P->n];
  return sign < 0 ? v : conj(v);
}

static void fft_rec(const fftplan *P, int n, const cplx *x, int xs, cplx *X, int sign, int fi) {
  if (n == 1) {
    X[0] = x[0];
    return;
  }
  int p = P->fac[fi], m = n / p;
  for (int r = 0; r < p; r++) fft_rec(P, m, x + (size_t)r * xs, xs * p, X + (size_t)r * m, sign, fi + 1);
  long step = P->n / n;
  cplx t[64], o[64];
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) t[r] = X[r * m + k] * tw(P, step * r * k, sign);
    for (int q = 0; q < p; q++) {
      cplx s = t[0];
      for (int r = 1; r < p; r++) s += t[r] * tw(P, step * m * ((r * q) % p), sign);
      o[q] = s;
    }
    for (int q = 0; q < p; q++) X[k + q * m] = o[q];
  }
}

/* in-place transform of x[0..n-1] with stride 1; work has n entries */
static void fft_exec(const fftplan *P, cplx *x, cplx *work, int sign) {
  memcpy(work, x, sizeof(cplx) * (size_t)P->n);
  fft_rec(P, P->n, work, 1, x, sign, 0);
}

/* c2r of logical length m (FFTW dft_c2r): x_j = sum_k Z_k e^{+2 pi i jk/m}
 * over the Hermitian extension of X[0..m/2]; Im X_0 and Im X_{m/2} are
 * ignored (the Hermitian extension has them real). */
static void c2r_exec(const fftplan *P, const cplx *X, double *x, cplx *buf, cplx *work) {
  int m = P->n, h = m / 2;
  buf[0] = creal(X[0]);
  for (int k = 1; k < h; k++) {
    buf[k] = X[k];
    buf[m - k] = conj(X[k]);
  }
  buf[h] = creal(X[h]);
  fft_exec(P, buf, work, +1);
  for (int j = 0; j < m; j++) x[j] = creal(buf[j]);
}

/* r2c of length m (FFTW dft_r2c): X_k = sum_j x_j e^{-2 pi i jk/m}, k=0..m/2 */
static void r2c_exec(const fftplan *P, const double *x, cplx *X, cplx *buf, cplx *work) {
  int m = P->n;
  for (int j = 0; j < m; j++) buf[j] = x[j];
  fft_exec(P, buf, work, -1);
  for (int k = 0; k <= m / 2; k++) X[k] = buf[k];
}

void orc_fft(int n, int sign, double *inout) {
  fftplan P;
  plan_init(&P, n);
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  fft_exec(&P, (cplx *)inout, w, sign);
  free(w);
  plan_free(&P);
}
void orc_c2r(int m, const double *X, double *x) {
  fftplan P;
  plan_init(&P, m);
  cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)m * 2);
  c2r_exec(&P, (const cplx *)X, x, b, b + m);
  free(b);
  plan_free(&P);
}
void orc_r2c(int m, const double *x, double *X) {
  fftplan P;
  plan_init(&P, m);
  cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)m * 2);
  r2c_exec(&P, x, (cplx *)X, b, b + m);
  free(b);
  plan_free(&P);
}

/* fftFIT (ffts.f90:78-86): true when n is 2^k or 3*2^k */
static int fftfit(int i) {
  int j = i;
  while (j % 2 == 0) j >>= 1;
  return j == 1 || j == 3;
}

/* ------------------------------------------------------ dense 5x5 LU -- */
/* LUdecomp (rbmat.f90:60-76): in-place, no pivoting, eliminates from the
 * last row up; diagonal stores the reciprocal pivot. */
static void ludecomp(double A[5][5]) {
  for (int i = 4; i >= 1; i--) {
    double piv = 1.0 / A[i][i];
    A[i][i] = piv;
    for (int j = 0; j < i; j++) A[i][j] = A[i][j] * piv;
    for (int k = 0; k < i; k++) {
      piv = A[k][i];
      for (int j = 0; j < i; j++) A[k][j] = A[k][j] - piv * A[i][j];
    }
  }
  A[0][0] = 1.0 / A[0][0];
}
/* LLUdiv, operator .bs. (rbmat.f90:201-215) */
static void lludiv(double A[5][5], const double b[5], double x[5]) {
  x[4] = b[4] * A[4][4];
  for (int i = 3; i >= 0; i--) {
    double s = 0.0;
    for (int j = i + 1; j < 5; j++) s += A[i][j] * x[j];
    x[i] = (b[i] - s) * A[i][i];
  }
  for (int i = 1; i < 5; i++) {
    double s = 0.0;
    for (int j = 0; j < i; j++) s += A[i][j] * x[j];
    x[i] = x[i] - s;
  }
}
void orc_ludecomp5(double *A) { ludecomp((double(*)[5])A); }
void orc_lludiv5(const double *A, const double *b, double *x) {
  double M[5][5];
  memcpy(M, A, sizeof M);
  lludiv(M, b, x);
}

/* ---------------------------------------------------- banded UL steps -- */
/* LU5decompStep, single y-slab (first=last) (rbparmat_blocking.f90:20-51).
 * A is rows x 5, A[r][j+2] = A(r, j); row r <-> iy=r+1. */
static void lu5decomp(double (*A)[5], int rows) {
  int HI1 = rows - 1;
  A[HI1 - 2][3] = 0.0;
  A[HI1 - 2][4] = 0.0;
  A[HI1 - 3][4] = 0.0;
  for (int i = HI1 - 2; i >= 0; i--) {
    for (int k = 2; k >= 1; k--) {
      double piv = A[i][k + 2];
      for (int j = -1; j >= -2; j--) A[i][j + k + 2] = A[i][j + k + 2] - piv * A[i + k][j + 2];
    }
    double piv = 1.0 / A[i][2];
    A[i][2] = piv;
    A[i][0] = A[i][0] * piv;
    A[i][1] = A[i][1] * piv;
  }
  A[0][0] = 0.0;
  A[0][1] = 0.0;
  A[1][0] = 0.0;
}
/* LeftLU5divStep1 in place (rbparmat_blocking.f90:57-77).  x points at
 * element "-2" of the reference's x(-2:), i.e. x[i+2] = x(i). */
static void lu5div1(double (*A)[5], int rows, cplx *x) {
  int HI1 = rows - 1;
  for (int i = HI1 - 2; i >= 0; i--) {
    cplx s = A[i][3] * x[i + 3] + A[i][4] * x[i + 4];
    x[i + 2] = (x[i + 2] - s) * A[i][2];
  }
}
/* LeftLU5divStep2 (rbparmat_blocking.f90:82-100) */
static void lu5div2(double (*A)[5], int rows, cplx *b) {
  int HI1 = rows - 1;
  for (int i = 0; i <= HI1; i++) {
    cplx s = A[i][0] * b[i] + A[i][1] * b[i + 1];
    b[i + 2] = b[i + 2] - s;
  }
}
void orc_lu5decomp(double *A, int rows, int first, int last) {
  (void)first;
  (void)last;
  lu5decomp((double(*)[5])A, rows);
}
void orc_lu5div1(const double *A, int rows, double *x) { lu5div1((double(*)[5])A, rows, (cplx *)x); }
void orc_lu5div2(const double *A, int rows, double *b) { lu5div2((double(*)[5])A, rows, (cplx *)b); }

/* ------------------------------------------------------------ state ---- */
typedef struct {
  double d0[5], d1[5], d2[5], d4[5];
} Di; /* typedef.f90:28-30 */
typedef struct {
  cplx u, v, w, vy, eta;
} BCOND; /* typedef.f90:32-34 */

struct orc {
  orc_params p;
  int nx, ny, nz, nxd, nzd, Nx, Nz, Ny;
  double PI, dx, dz, factor;
  double *y;  /* y[iy+1], iy=-1..ny+1 */
  double *dy; /* dy[iy], iy=1..ny-1 */
  Di *der;    /* der[iy], iy=1..ny-1 */
  double d040[5], d140[5], d240[5], d14m1[5], d24m1[5];
  double d04n[5], d14n[5], d24n[5], d14np1[5], d24np1[5];
  double v0bc[5], v0m1bc[5], eta0bc[5], eta0m1bc[5];
  double vnbc[5], vnp1bc[5], etanbc[5], etanp1bc[5];
  double (*D0mat)[5]; /* rows iy=1..ny+1 */
  cplx *V;            /* Fortran layout (iy,iz,ix,ic) */
  cplx *oldrhs;       /* [iy-1][ix][iz][2]: eta, d2v */
  cplx *P;            /* products [iy+1][k][ix][iz] */
  BCOND *bc0, *bcn;   /* [ix][iz] */
  double time, deltat, meanpx, meanpz, corrpx, corrpz, fr[3], cfl;
  fftplan pz, px;
  cplx *F;          /* body force, layout of V; NULL when off */
  int bf_kind;      /* 0 none, 1 am_f1, 2 am_butterfly, 3 coriolis */
  double bf_par[3];
  int bf_iz;        /* iz_f (am_*) or iz_thr (coriolis) */
};

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }

#define VIX(o, iy, iz, ix, c) \
  ((size_t)((iy) + 1) + (size_t)(o)->Ny * ((size_t)((iz) + (o)->nz) + (size_t)(o)->Nz * ((size_t)(ix) + (size_t)(o)->Nx * (size_t)(c))))
#define PIX(o, iy, k, ix, iz) \
  ((((size_t)((iy) + 1) * 6 + (size_t)(k)) * (size_t)(o)->Nx + (size_t)(ix)) * (size_t)(o)->Nz + (size_t)((iz) + (o)->nz))
#define OIX(o, iy, ix, iz) ((((size_t)((iy)-1) * (size_t)(o)->Nx + (size_t)(ix)) * (size_t)(o)->Nz + (size_t)((iz) + (o)->nz)) * 2)
#define Y(o, iy) ((o)->y[(iy) + 1])

/* i*a*X computed as the reference's complex product (0,a)*(Xr,Xi):
 * real part 0*Xr - a*Xi, imaginary part 0*Xi + a*Xr, i.e. exactly (-a Xi, a Xr) */
static inline cplx imul(double a, cplx X) { return CMPLX(-(a * cimag(X)), a * creal(X)); }

/* setup_derivatives (dnsdata.f90:241-286) */
static void setup_derivatives(orc_t *o) {
  double M[5][5], t[5], h[5];
  int ny = o->ny;
  for (int iy = 1; iy <= ny - 1; iy++) {
    Di *d = &o->der[iy];
    for (int j = 0; j < 5; j++) h[j] = Y(o, iy - 2 + j) - Y(o, iy);
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = pow(h[j], 4.0 - i);
    ludecomp(M);
    memset(t, 0, sizeof t);
    t[0] = 24;
    lludiv(M, t, d->d4);
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = (5.0 - i) * (6.0 - i) * (7.0 - i) * (8.0 - i) * pow(h[j], 4.0 - i);
    for (int i = 0; i < 5; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += d->d4[j] * pow(h[j], 8.0 - i);
      t[i] = s;
    }
    ludecomp(M);
    lludiv(M, t, d->d0);
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) M[i][j] = pow(h[j], 4.0 - i);
    ludecomp(M);
    memset(t, 0, sizeof t);
    for (int i = 0; i <= 2; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += d->d0[j] * (4.0 - i) * (3.0 - i) * pow(h[j], 2.0 - i);
      t[i] = s;
    }
    lludiv(M, t, d->d2);
    memset(t, 0, sizeof t);
    for (int i = 0; i <= 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 5; j++) s += d->d0[j] * (4.0 - i) * pow(h[j], 3.0 - i);
      t[i] = s;
    }
    lludiv(M, t, d->d1);
  }
  /* wall and ghost stencils (first) */
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) M[i][j] = pow(Y(o, -1 + j) - Y(o, 0), 4.0 - i);
  ludecomp(M);
  memset(t, 0, sizeof t);
  t[3] = 1.0;
  lludiv(M, t, o->d140);
  memset(t, 0, sizeof t);
  t[2] = 2.0;
  lludiv(M, t, o->d240);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) M[i][j] = pow(Y(o, -1 + j) - Y(o, -1), 4.0 - i);
  ludecomp(M);
  memset(t, 0, sizeof t);
  t[3] = 1.0;
  lludiv(M, t, o->d14m1);
  memset(t, 0, sizeof t);
  t[2] = 2.0;
  lludiv(M, t, o->d24m1);
  memset(o->d040, 0, sizeof o->d040);
  o->d040[1] = 1.0;
  /* (last) */
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) M[i][j] = pow(Y(o, ny - 3 + j) - Y(o, ny), 4.0 - i);
  ludecomp(M);
  memset(t, 0, sizeof t);
  t[3] = 1.0;
  lludiv(M, t, o->d14n);
  memset(t, 0, sizeof t);
  t[2] = 2.0;
  lludiv(M, t, o->d24n);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) M[i][j] = pow(Y(o, ny - 3 + j) - Y(o, ny + 1), 4.0 - i);
  ludecomp(M);
  memset(t, 0, sizeof t);
  t[3] = 1.0;
  lludiv(M, t, o->d14np1);
  memset(t, 0, sizeof t);
  t[2] = 2.0;
  lludiv(M, t, o->d24np1);
  memset(o->d04n, 0, sizeof o->d04n);
  o->d04n[3] = 1.0;
  /* D0mat rows 1..ny-1 = d0, rows ny..ny+1 zero (dnsdata.f90:277,284) */
  memset(o->D0mat, 0, sizeof(double) * 5 * (size_t)(ny + 1));
  for (int iy = 1; iy <= ny - 1; iy++)
    for (int j = 0; j < 5; j++) o->D0mat[iy - 1][j] = o->der[iy].d0[j];
  lu5decomp(o->D0mat, ny + 1);
}

/* setup_boundary_conditions (dnsdata.f90:290-308), full channel */
static void setup_bc(orc_t *o) {
  int ny = o->ny;
  memcpy(o->v0bc, o->d040, sizeof o->v0bc);
  memcpy(o->v0m1bc, o->d140, sizeof o->v0m1bc);
  memcpy(o->eta0bc, o->d040, sizeof o->eta0bc);
  memcpy(o->eta0m1bc, o->der[1].d4, sizeof o->eta0m1bc);
  for (int j = 1; j < 5; j++) o->v0bc[j] = o->v0bc[j] - o->v0bc[0] * o->v0m1bc[j] / o->v0m1bc[0];
  for (int j = 1; j < 5; j++) o->eta0bc[j] = o->eta0bc[j] - o->eta0bc[0] * o->eta0m1bc[j] / o->eta0m1bc[0];
  memcpy(o->vnbc, o->d04n, sizeof o->vnbc);
  memcpy(o->vnp1bc, o->d14n, sizeof o->vnp1bc);
  memcpy(o->etanbc, o->d04n, sizeof o->etanbc);
  memcpy(o->etanp1bc, o->der[ny - 1].d4, sizeof o->etanp1bc);
  for (int j = 0; j < 4; j++) o->vnbc[j] = o->vnbc[j] - o->vnbc[4] * o->vnp1bc[j] / o->vnp1bc[4];
  for (int j = 0; j < 4; j++) o->etanbc[j] = o->etanbc[j] - o->etanbc[4] * o->etanp1bc[j] / o->etanp1bc[4];
}

orc_t *orc_create(const orc_params *p) {
  orc_t *o = (orc_t *)calloc(1, sizeof *o);
  o->p = *p;
  o->nx = p->nx;
  o->ny = p->ny;
  o->nz = p->nz;
  /* read_dnsin (dnsdata.f90:110-114,124) */
  o->nxd = 3 * (p->nx + 1) / 2;
  o->nzd = 3 * p->nz;
  while (!fftfit(o->nxd)) o->nxd++;
  while (!fftfit(o->nzd)) o->nzd++;
  o->Nx = o->nx + 1;
  o->Nz = 2 * o->nz + 1;
  o->Ny = o->ny + 3;
  /* PI is initialised from a default-real literal (dnsdata.f90:26) */
  o->PI = (double)(float)3.14159265358979323846;
  o->dx = o->PI / (p->alfa0 * o->nxd);
  o->dz = 2.0 * o->PI / (p->beta0 * o->nzd);
  o->factor = 1.0 / (2.0 * o->nxd * o->nzd);
  /* y grid (dnsdata.f90:153): 2.0d0*real(iy)/real(ny) is evaluated left to
   * right, i.e. in double precision */
  o->y = (double *)calloc((size_t)o->Ny, sizeof(double));
  o->dy = (double *)calloc((size_t)o->ny + 1, sizeof(double));
  for (int iy = -1; iy <= o->ny + 1; iy++) {
    double r = 2.0 * (double)(float)iy / (double)(float)o->ny;
    o->y[iy + 1] = p->ymin + 0.5 * (p->ymax - p->ymin) * (tanh(p->a * (r - 1)) / tanh(p->a) + 1);
  }
  for (int iy = 1; iy <= o->ny - 1; iy++) o->dy[iy] = 0.5 * (Y(o, iy + 1) - Y(o, iy - 1));
  o->der = (Di *)calloc((size_t)o->ny + 1, sizeof(Di));
  o->D0mat = (double(*)[5])calloc((size_t)o->ny + 1, sizeof(double[5]));
  setup_derivatives(o);
  setup_bc(o);
  size_t nv = (size_t)o->Ny * o->Nz * o->Nx * 3;
  o->V = (cplx *)calloc(nv, sizeof(cplx));
  o->oldrhs = (cplx *)calloc((size_t)(o->ny - 1) * o->Nx * o->Nz * 2, sizeof(cplx));
  o->P = (cplx *)calloc((size_t)o->Ny * 6 * o->Nx * o->Nz, sizeof(cplx));
  o->bc0 = (BCOND *)calloc((size_t)o->Nx * o->Nz, sizeof(BCOND));
  o->bcn = (BCOND *)calloc((size_t)o->Nx * o->Nz, sizeof(BCOND));
  o->time = p->time;
  o->deltat = p->deltat;
  o->meanpx = p->meanpx;
  o->meanpz = p->meanpz;
  plan_init(&o->pz, o->nzd);
  plan_init(&o->px, 2 * o->nxd);
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  free(o->y);
  free(o->dy);
  free(o->der);
  free(o->D0mat);
  free(o->V);
  free(o->oldrhs);
  free(o->P);
  free(o->bc0);
  free(o->bcn);
  free(o->F);
  plan_free(&o->pz);
  plan_free(&o->px);
  free(o);
}

void orc_sizes(const orc_t *o, int32_t out[5]) {
  out[0] = o->nx;
  out[1] = o->ny;
  out[2] = o->nz;
  out[3] = o->nxd;
  out[4] = o->nzd;
}
void orc_set_field(orc_t *o, const double *V) {
  memcpy(o->V, V, sizeof(cplx) * (size_t)o->Ny * o->Nz * o->Nx * 3);
}
void orc_get_field(const orc_t *o, double *V) {
  memcpy(V, o->V, sizeof(cplx) * (size_t)o->Ny * o->Nz * o->Nx * 3);
}
void orc_get_scalars(const orc_t *o, double s[10]) {
  s[0] = o->time;
  s[1] = o->deltat;
  s[2] = o->meanpx;
  s[3] = o->meanpz;
  s[4] = o->corrpx;
  s[5] = o->corrpz;
  s[6] = o->fr[0];
  s[7] = o->fr[1];
  s[8] = o->fr[2];
  s[9] = o->cfl;
}
void orc_set_scalars(orc_t *o, const double s[10]) {
  o->time = s[0];
  o->deltat = s[1];
  o->meanpx = s[2];
  o->meanpz = s[3];
  o->corrpx = s[4];
  o->corrpz = s[5];
  o->fr[0] = s[6];
  o->fr[1] = s[7];
  o->fr[2] = s[8];
  o->cfl = s[9];
}
void orc_get_grid(const orc_t *o, double *y, double *dy) {
  memcpy(y, o->y, sizeof(double) * (size_t)o->Ny);
  for (int iy = 1; iy <= o->ny - 1; iy++) dy[iy - 1] = o->dy[iy];
}
void orc_get_der(const orc_t *o, double *der) {
  for (int iy = 1; iy <= o->ny - 1; iy++) {
    const Di *d = &o->der[iy];
    double *q = der + (size_t)(iy - 1) * 20;
    memcpy(q, d->d0, 40);
    memcpy(q + 5, d->d1, 40);
    memcpy(q + 10, d->d2, 40);
    memcpy(q + 15, d->d4, 40);
  }
}
void orc_get_bnd(const orc_t *o, double *b) {
  const double *s[10] = {o->d040, o->d140, o->d240, o->d14m1, o->d24m1, o->d04n, o->d14n, o->d24n, o->d14np1, o->d24np1};
  for (int i = 0; i < 10; i++) memcpy(b + 5 * i, s[i], 40);
}
void orc_get_bc(const orc_t *o, double *b) {
  const double *s[8] = {o->v0bc, o->v0m1bc, o->eta0bc, o->eta0m1bc, o->vnbc, o->vnp1bc, o->etanbc, o->etanp1bc};
  for (int i = 0; i < 8; i++) memcpy(b + 5 * i, s[i], 40);
}
void orc_get_d0mat(const orc_t *o, double *d) { memcpy(d, o->D0mat, sizeof(double) * 5 * (size_t)(o->ny + 1)); }

/* yintegr (dnsdata.f90:312-324), f indexed f[iy+1] */
static double yintegr(const orc_t *o, const double *f) {
  double II = 0.0;
  for (int iy = 1; iy <= o->ny - 1; iy += 2) {
    double yp1 = Y(o, iy + 1) - Y(o, iy), ym1 = Y(o, iy - 1) - Y(o, iy);
    double a1 = -(1.0 / 3.0 * ym1) + 1.0 / 6.0 * yp1 + 1.0 / 6.0 * yp1 * yp1 / ym1;
    double a3 = +(1.0 / 3.0 * yp1) - 1.0 / 6.0 * ym1 - 1.0 / 6.0 * ym1 * ym1 / yp1;
    double a2 = yp1 - ym1 - a1 - a3;
    II = II + a1 * f[iy] + a2 * f[iy + 1] + a3 * f[iy + 2];
  }
  return II;
}
double orc_yintegr(const orc_t *o, const double *f) { return yintegr(o, f); }

/* COMPLEXderiv (dnsdata.f90:337-373): f0, f1 point at element iy=-1 */
static void complexderiv(const orc_t *o, const cplx *f0, cplx *f1) {
  int ny = o->ny;
#define F0(i) f0[(i) + 1]
#define F1(i) f1[(i) + 1]
  cplx s;
  s = 0;
  for (int j = 0; j < 5; j++) s += o->d140[j] * F0(-1 + j);
  F1(0) = s;
  s = 0;
  for (int j = 0; j < 5; j++) s += o->d14m1[j] * F0(-1 + j);
  F1(-1) = s;
  s = 0;
  for (int j = 0; j < 5; j++) s += o->d14n[j] * F0(ny - 3 + j);
  F1(ny) = s;
  s = 0;
  for (int j = 0; j < 5; j++) s += o->d14np1[j] * F0(ny - 3 + j);
  F1(ny + 1) = s;
  for (int iy = 1; iy <= ny - 1; iy++) {
    s = 0;
    for (int j = 0; j < 5; j++) s += o->der[iy].d1[j] * F0(iy - 2 + j);
    F1(iy) = s;
  }
  F1(1) = F1(1) - (o->der[1].d0[1] * F1(0) + o->der[1].d0[0] * F1(-1));
  F1(2) = F1(2) - o->der[2].d0[0] * F1(0);
  F1(ny - 1) = F1(ny - 1) - (o->der[ny - 1].d0[3] * F1(ny) + o->der[ny - 1].d0[4] * F1(ny + 1));
  F1(ny - 2) = F1(ny - 2) - o->der[ny - 2].d0[4] * F1(ny);
  lu5div1(o->D0mat, ny + 1, f1);
#undef F0
#undef F1
}
void orc_complexderiv(const orc_t *o, const double *f0, double *f1) {
  complexderiv(o, (const cplx *)f0, (cplx *)f1);
  lu5div2(o->D0mat, o->ny + 1, (cplx *)f1);
}

/* ------------------------------------------------------ convolutions -- */
typedef struct {
  cplx *vz;   /* 3 x Nx x nzd : z-physical, x-spectral velocities */
  cplx *pzb;  /* 6 x Nx x nzd : z-physical, x-spectral products */
  cplx *col;  /* nzd or 2nxd */
  cplx *work; /* 2*max */
  cplx *buf;  /* 2*max */
  double *phys; /* 9 x 2nxd */
} convws;

static void ws_alloc(const orc_t *o, convws *w) {
  int mx = o->nzd > 2 * o->nxd ? o->nzd : 2 * o->nxd;
  w->vz = (cplx *)malloc(sizeof(cplx) * 3 * (size_t)o->Nx * o->nzd);
  w->pzb = (cplx *)malloc(sizeof(cplx) * 6 * (size_t)o->Nx * o->nzd);
  w->col = (cplx *)malloc(sizeof(cplx) * (size_t)mx);
  w->work = (cplx *)malloc(sizeof(cplx) * (size_t)mx * 2);
  w->buf = (cplx *)malloc(sizeof(cplx) * (size_t)mx * 2);
  w->phys = (double *)malloc(sizeof(double) * 9 * 2 * (size_t)o->nxd);
}
static void ws_free(convws *w) {
  free(w->vz);
  free(w->pzb);
  free(w->col);
  free(w->work);
  free(w->buf);
  free(w->phys);
}

/* convolutions (dnsdata.f90:487-602) for one plane; products go to
 * out[k][ix][iz+nz] (k = uu,vv,ww,uv,vw,uw); returns the plane's CFL max
 * (dnsdata.f90:552-556) when do_cfl. */
static double conv_plane(const orc_t *o, int iy, int do_cfl, cplx *out, convws *w) {
  int nx = o->nx, nz = o->nz, nzd = o->nzd, nxd = o->nxd, Nx = o->Nx, Nz = o->Nz, M = 2 * nxd;
  double cfl = 0.0;
  /* pad in z and IFT (dnsdata.f90:504-510) */
  for (int c = 0; c < 3; c++)
    for (int ix = 0; ix <= nx; ix++) {
      cplx *col = w->vz + ((size_t)c * Nx + ix) * nzd;
      for (int z = 0; z < nzd; z++) col[z] = 0;
      for (int iz = 0; iz <= nz; iz++) col[iz] = o->V[VIX(o, iy, iz, ix, c)];
      for (int iz = -nz; iz <= -1; iz++) col[nzd + iz] = o->V[VIX(o, iy, iz, ix, c)];
      fft_exec(&o->pz, col, w->work, +1);
    }
  /* per physical z: zTOx, zero x modes nx+1..nxd, RFT, CFL, products, HFT, xTOz */
  cplx *X = w->col;
  double *u = w->phys, *pr = w->phys + 3 * (size_t)M;
  for (int z = 0; z < nzd; z++) {
    for (int c = 0; c < 3; c++) {
      for (int ix = 0; ix <= nxd; ix++) X[ix] = ix <= nx ? w->vz[((size_t)c * Nx + ix) * nzd + z] : 0;
      c2r_exec(&o->px, X, u + (size_t)c * M, w->buf, w->work);
    }
    const double *uu = u, *vv = u + M, *ww = u + 2 * M;
    if (do_cfl) {
      double m = 0.0;
      for (int x = 0; x < M; x++) {
        double c = fabs(uu[x]) / o->dx + fabs(vv[x]) / o->dy[iy] + fabs(ww[x]) / o->dz;
        if (c > m) m = c;
      }
      if (m > cfl) cfl = m;
    }
    double f = o->factor;
    for (int x = 0; x < M; x++) {
      pr[3 * M + x] = uu[x] * vv[x] * f;
      pr[4 * M + x] = vv[x] * ww[x] * f;
      pr[5 * M + x] = uu[x] * ww[x] * f;
      pr[0 * M + x] = uu[x] * uu[x] * f;
      pr[1 * M + x] = vv[x] * vv[x] * f;
      pr[2 * M + x] = ww[x] * ww[x] * f;
    }
    for (int k = 0; k < 6; k++) {
      r2c_exec(&o->px, pr + (size_t)k * M, X, w->buf, w->work);
      for (int ix = 0; ix <= nx; ix++) w->pzb[((size_t)k * Nx + ix) * nzd + z] = X[ix];
    }
  }
  /* FFT in z, keep modes -nz..nz (izd, dnsdata.f90:156) */
  for (int k = 0; k < 6; k++)
    for (int ix = 0; ix <= nx; ix++) {
      cplx *col = w->pzb + ((size_t)k * Nx + ix) * nzd;
      fft_exec(&o->pz, col, w->work, -1);
      cplx *dst = out + ((size_t)k * Nx + ix) * Nz;
      for (int iz = -nz; iz <= nz; iz++) dst[iz + nz] = col[iz >= 0 ? iz : nzd + iz];
    }
  return cfl;
}

void orc_convolutions(orc_t *o, int iy, int compute_cfl, double *out6) {
  convws w;
  ws_alloc(o, &w);
  double c = conv_plane(o, iy, compute_cfl && iy >= 1 && iy <= o->ny - 1, (cplx *)out6, &w);
  if (c > o->cfl) o->cfl = c;
  ws_free(&w);
}

/* all planes -1..ny+1 into o->P (the reference computes the same planes in a
 * 5-slot ring, dnsdata.f90:630-632; the values are identical) */
static void all_products(orc_t *o, int compute_cfl, int lo, int hi) {
  int n = hi - lo + 1;
  double *cfls = (double *)calloc((size_t)n, sizeof(double));
#pragma omp parallel num_threads(g_threads)
  {
    convws w;
    ws_alloc(o, &w);
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < n; i++) {
      int iy = lo + i;
      cfls[i] = conv_plane(o, iy, compute_cfl && iy >= 1 && iy <= o->ny - 1, o->P + PIX(o, iy, 0, 0, -o->nz), &w);
    }
    ws_free(&w);
  }
  for (int i = 0; i < n; i++)
    if (cfls[i] > o->cfl) o->cfl = cfls[i];
  free(cfls);
}

/* ----------------------------------------------------------- buildrhs -- */
/* buildrhs (dnsdata.f90:611-673) with the DD stencil (:609-610), OS/SQ
 * (:476-477) and timescheme (:486).  New rhs goes to V(1..ny-1, :, :, 1:2). */
static void rhs_rows(orc_t *o, const double ode[3], int lo, int hi, cplx *newrhs) {
  int nx = o->nx, nz = o->nz, Nz = o->Nz, Nx = o->Nx;
  const double ni = o->p.ni, dt = o->deltat;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int iy = lo; iy <= hi; iy++) {
    const Di *d = &o->der[iy];
    for (int iz = -nz; iz <= nz; iz++)
      for (int ix = 0; ix <= nx; ix++) {
        double al = o->p.alfa0 * ix, be = o->p.beta0 * iz;
        double k2 = (o->p.alfa0 * ix) * (o->p.alfa0 * ix) + (o->p.beta0 * iz) * (o->p.beta0 * iz);
#define PP(k, j) o->P[PIX(o, iy - 2 + (j), (k)-1, ix, iz)]
#define DD(f, k) (d->f[0] * PP(k, 0) + d->f[1] * PP(k, 1) + d->f[2] * PP(k, 2) + d->f[3] * PP(k, 3) + d->f[4] * PP(k, 4))
#define VV(j, c) o->V[VIX(o, iy - 2 + (j), iz, ix, (c)-1)]
        cplx DD0_6 = DD(d0, 6), DD1_6 = DD(d1, 6);
        cplx rhsu = -imul(al, DD(d0, 1)) - DD(d1, 4) - imul(be, DD0_6);
        cplx rhsv = -imul(al, DD(d0, 4)) - DD(d1, 2) - imul(be, DD(d0, 5));
        cplx rhsw = -imul(al, DD0_6) - DD(d1, 5) - imul(be, DD(d0, 3));
        cplx expl = imul(al, imul(al, DD(d1, 1)) + DD(d2, 4) + imul(be, DD1_6)) +
                    imul(be, imul(al, DD1_6) + DD(d2, 5) + imul(be, DD(d1, 3))) - k2 * rhsv;
        /* body-force terms (dnsdata.f90:643-645, 650-652, 657-659): D0 and
         * D1 of F(iy-2:iy+2) by the D0/D1 macros (:330-331) */
        cplx D0F[3] = {0, 0, 0}, D1F[3] = {0, 0, 0};
        if (o->F) {
          for (int c = 0; c < 3; c++) {
            double r0 = 0, i0 = 0, r1 = 0, i1 = 0;
            for (int j = 0; j < 5; j++) {
              cplx f = o->F[VIX(o, iy - 2 + j, iz, ix, c)];
              r0 += d->d0[j] * creal(f);
              i0 += d->d0[j] * cimag(f);
              r1 += d->d1[j] * creal(f);
              i1 += d->d1[j] * cimag(f);
            }
            D0F[c] = CMPLX(r0, i0);
            D1F[c] = CMPLX(r1, i1);
          }
          expl = expl - k2 * D0F[1] - imul(al, D1F[0]) - imul(be, D1F[2]);
        }
        size_t oi = OIX(o, iy, ix, iz);
        size_t ni_ = (((size_t)(iy - 1) * Nx + ix) * Nz + (iz + nz)) * 2;
        /* D2v equation */
        {
          double d2r = 0, d2i = 0, d0r = 0, d0i = 0;
          for (int j = 0; j < 5; j++) {
            d2r += d->d2[j] * creal(VV(j, 2));
            d2i += d->d2[j] * cimag(VV(j, 2));
          }
          for (int j = 0; j < 5; j++) {
            d0r += d->d0[j] * creal(VV(j, 2));
            d0i += d->d0[j] * cimag(VV(j, 2));
          }
          cplx unkn = CMPLX(d2r, d2i) - k2 * CMPLX(d0r, d0i);
          cplx impl = 0;
          for (int j = 0; j < 5; j++) {
            double os = ni * (d->d4[j] - 2.0 * k2 * d->d2[j] + k2 * k2 * d->d0[j]);
            impl += os * VV(j, 2);
          }
          cplx old = o->oldrhs[oi + 1];
          newrhs[ni_ + 1] = ode[0] * unkn / dt + impl + ode[1] * expl - ode[2] * old;
          o->oldrhs[oi + 1] = expl;
        }
        if (ix == 0 && iz == 0) {
          cplx ex = CMPLX(creal(rhsu) + o->meanpx, creal(rhsw) + o->meanpz);
          if (o->F) ex = ex + CMPLX(creal(D0F[0]), creal(D0F[2])); /* rD0(F,1,3) */
          double a1 = 0, a3 = 0, b1 = 0, b3 = 0;
          for (int j = 0; j < 5; j++) {
            a1 += d->d0[j] * creal(VV(j, 1));
            a3 += d->d0[j] * creal(VV(j, 3));
          }
          for (int j = 0; j < 5; j++) {
            b1 += d->d2[j] * creal(VV(j, 1));
            b3 += d->d2[j] * creal(VV(j, 3));
          }
          cplx unkn = CMPLX(a1, a3);
          cplx impl = ni * CMPLX(b1, b3);
          cplx old = o->oldrhs[oi + 0];
          newrhs[ni_ + 0] = ode[0] * unkn / dt + impl + ode[1] * ex - ode[2] * old;
          o->oldrhs[oi + 0] = ex;
        } else {
          cplx ex = imul(be, rhsu) - imul(al, rhsw);
          if (o->F) ex = ex + imul(be, D0F[0]) - imul(al, D0F[2]);
          double r1 = 0, i1 = 0, r3 = 0, i3 = 0;
          for (int j = 0; j < 5; j++) {
            r1 += d->d0[j] * creal(VV(j, 1));
            i1 += d->d0[j] * cimag(VV(j, 1));
          }
          for (int j = 0; j < 5; j++) {
            r3 += d->d0[j] * creal(VV(j, 3));
            i3 += d->d0[j] * cimag(VV(j, 3));
          }
          cplx unkn = imul(be, CMPLX(r1, i1)) - imul(al, CMPLX(r3, i3));
          cplx impl = 0;
          for (int j = 0; j < 5; j++) {
            double sq = ni * (d->d2[j] - k2 * d->d0[j]);
            impl += sq * (imul(be, VV(j, 1)) - imul(al, VV(j, 3)));
          }
          cplx old = o->oldrhs[oi + 0];
          newrhs[ni_ + 0] = ode[0] * unkn / dt + impl + ode[1] * ex - ode[2] * old;
          o->oldrhs[oi + 0] = ex;
        }
#undef PP
#undef DD
#undef VV
      }
  }
}

/* ---------------------------------------------------------- body force -- */
/* The three modules of body_forces/, each a set_body_force that writes
 * F(iy,iz,ix,pc) = g V(iy,iz,ix,c) where its mask holds and leaves F alone
 * elsewhere. */
static void bf_am(orc_t *o, int butterfly) {
  /* am_f1.inc:11-27 and am_butterfly.inc:11-27; amp = bf_par[0].  am_f1
   * loops over iz = -iz_f..iz_f, capped here at nz (the last mode). */
  const double amp = o->bf_par[0];
  const int nz = o->nz, izf = o->bf_iz;
  const int lo = butterfly ? -nz : -(izf < nz ? izf : nz);
  for (int iV = 0; iV < 3; iV++)
    for (int ix = 0; ix <= o->nx; ix++)
      for (int iz = lo; iz <= -lo; iz++) {
        double lzp = iz == 0 ? (double)1E10f : 2 * o->PI / (o->p.beta0 * abs(iz)) * 1000;
        for (int iy = -1; iy <= o->ny + 1; iy++) {
          double yp = (Y(o, iy) > 1 ? o->p.ymax - Y(o, iy) : Y(o, iy)) * 1000;
          int on;
          if (!butterfly)
            on = lzp > (double)2.3f * (yp * yp) && !(ix == 0 && iz == 0);
          else
            on = (abs(iz) <= izf && yp <= 60 && !(ix == 0 && iz == 0)) || (abs(iz) > izf && yp > 60);
          if (on) o->F[VIX(o, iy, iz, ix, iV)] = -(amp * o->V[VIX(o, iy, iz, ix, iV)]);
        }
      }
}

static void bf_coriolis(orc_t *o) {
  /* coriolis.inc:26-40: omega2, y_threshold_bot = bf_par[0], bf_par[2] */
  const double omega2 = o->bf_par[0], bot = o->bf_par[2], top = o->p.ymax - bot;
  for (int c = 1; c <= 2; c++) {
    int pc = c % 2 + 1, zeichen = 2 * pc - 3;
    for (int ix = 0; ix <= o->nx; ix++)
      for (int iz = -o->bf_iz; iz <= o->bf_iz; iz++)
        for (int iy = -1; iy <= o->ny + 1; iy++)
          if (Y(o, iy) <= bot || Y(o, iy) >= top)
            o->F[VIX(o, iy, iz, ix, pc - 1)] = (zeichen * omega2) * o->V[VIX(o, iy, iz, ix, c - 1)];
  }
}

static void bf_alloc(orc_t *o) {
  if (!o->F) o->F = (cplx *)calloc((size_t)o->Ny * o->Nz * o->Nx * 3, sizeof(cplx));
}

void orc_set_body_force(orc_t *o, const double *F) {
  if (!F) {
    if (!o->bf_kind) {
      free(o->F);
      o->F = NULL;
    }
    return;
  }
  bf_alloc(o);
  memcpy(o->F, F, sizeof(cplx) * (size_t)o->Ny * o->Nz * o->Nx * 3);
}

void orc_set_body_force_module(orc_t *o, int kind, const double *par) {
  o->bf_kind = kind;
  if (!kind) return;
  bf_alloc(o);
  memcpy(o->bf_par, par, sizeof o->bf_par);
  if (kind == 1 || kind == 2) /* iz_f = NINT((2 PI / lambdaz_f) / (beta0/1000)), am_f1.inc:7 */
    o->bf_iz = (int)lround((2.0 * o->PI / par[1]) / (o->p.beta0 / 1000.0));
  else { /* iz_thr = min(nz, floor(kz_cutoff / beta0)), coriolis.inc:15 */
    double t = floor(par[1] / o->p.beta0);
    o->bf_iz = t < o->nz ? (int)t : o->nz;
  }
}

void orc_apply_body_force(orc_t *o) {
  if (o->bf_kind == 1 || o->bf_kind == 2) bf_am(o, o->bf_kind == 2);
  if (o->bf_kind == 3) bf_coriolis(o);
}

void orc_get_body_force(const orc_t *o, double *F) {
  size_t n = (size_t)o->Ny * o->Nz * o->Nx * 3;
  if (o->F)
    memcpy(F, o->F, sizeof(cplx) * n);
  else
    memset(F, 0, sizeof(cplx) * n);
}

/* the wall rows of F at the start of buildrhs (dnsdata.f90:616-628): zero on
 * the walls, ghost rows from D4(F) = 0 on rows 1 and ny-1 (D4 macro :333) */
void orc_body_force_walls(orc_t *o) {
  if (!o->F) return;
  const int ny = o->ny;
  for (int c = 0; c < 3; c++)
    for (int ix = 0; ix <= o->nx; ix++)
      for (int iz = -o->nz; iz <= o->nz; iz++) {
        o->F[VIX(o, -1, iz, ix, c)] = 0;
        o->F[VIX(o, 0, iz, ix, c)] = 0;
        o->F[VIX(o, ny, iz, ix, c)] = 0;
        o->F[VIX(o, ny + 1, iz, ix, c)] = 0;
      }
  for (int side = 0; side < 2; side++) {
    const int iy = side ? ny - 1 : 1, g = side ? ny + 1 : -1;
    const double *d4 = o->der[iy].d4, piv = side ? d4[4] : d4[0];
    for (int c = 0; c < 3; c++)
      for (int ix = 0; ix <= o->nx; ix++)
        for (int iz = -o->nz; iz <= o->nz; iz++) {
          double re = 0, im = 0;
          for (int j = 0; j < 5; j++) {
            cplx f = o->F[VIX(o, iy - 2 + j, iz, ix, c)];
            re += d4[j] * creal(f);
            im += d4[j] * cimag(f);
          }
          o->F[VIX(o, g, iz, ix, c)] = CMPLX(-re / piv, -im / piv);
        }
  }
}

void orc_buildrhs(orc_t *o, const double ode[3], int compute_cfl) {
  int ny = o->ny, nx = o->nx, nz = o->nz, Nz = o->Nz, Nx = o->Nx;
  /* set_body_force runs right before every buildrhs (channel.f90:129-157) */
  orc_apply_body_force(o);
  orc_body_force_walls(o);
  all_products(o, compute_cfl, -1, ny + 1);
  cplx *newrhs = (cplx *)malloc(sizeof(cplx) * (size_t)(ny - 1) * Nx * Nz * 2);
  rhs_rows(o, ode, 1, ny - 1, newrhs);
  /* write-back of the rhs into V(iy,:,:,1:2) (dnsdata.f90:667-671) */
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int iy = 1; iy <= ny - 1; iy++)
    for (int ix = 0; ix <= nx; ix++)
      for (int iz = -nz; iz <= nz; iz++) {
        size_t ni_ = (((size_t)(iy - 1) * Nx + ix) * Nz + (iz + nz)) * 2;
        o->V[VIX(o, iy, iz, ix, 0)] = newrhs[ni_ + 0];
        o->V[VIX(o, iy, iz, ix, 1)] = newrhs[ni_ + 1];
      }
  free(newrhs);
}

/* ----------------------------------------------------------- linsolve -- */
/* applybc_0 / applybc_n (dnsdata.f90:458-472); EQ rows are indexed by iy */
static void applybc_0(double (*EQ)[5], const double *bc0, const double *bc0m1) {
  double *r1 = EQ[0], *r2 = EQ[1]; /* iy=1, iy=2 */
  double e = r1[0];
  for (int j = 1; j <= 4; j++) r1[j] = r1[j] - e * bc0m1[j] / bc0m1[0];
  e = r1[1];
  for (int j = 2; j <= 4; j++) r1[j] = r1[j] - e * bc0[j] / bc0[1];
  e = r2[0];
  for (int j = 1; j <= 3; j++) r2[j] = r2[j] - e * bc0[j + 1] / bc0[1];
}
static void applybc_n(double (*EQ)[5], int ny, const double *bcn, const double *bcnp1) {
  double *r1 = EQ[ny - 2], *r2 = EQ[ny - 3]; /* iy=ny-1, iy=ny-2 */
  double e = r1[4];
  for (int j = 0; j <= 3; j++) r1[j] = r1[j] - e * bcnp1[j] / bcnp1[4];
  e = r1[3];
  for (int j = 0; j <= 2; j++) r1[j] = r1[j] - e * bcn[j] / bcn[3];
  e = r2[4];
  for (int j = 1; j <= 3; j++) r2[j] = r2[j] - e * bcn[j - 1] / bcn[3];
}

/* linsolve (linsolve_blocking.inc:3-107) over the x modes ixlo..ixhi */
static void linsolve_cols(orc_t *o, double lambda, int ixlo, int ixhi) {
  int ny = o->ny, nz = o->nz, Ny = o->Ny;
  const double ni = o->p.ni;
#pragma omp parallel num_threads(g_threads)
  {
    double(*D2v)[5] = (double(*)[5])malloc(sizeof(double[5]) * (size_t)(ny + 1));
    double(*eta)[5] = (double(*)[5])malloc(sizeof(double[5]) * (size_t)(ny + 1));
    cplx *temp = (cplx *)malloc(sizeof(cplx) * (size_t)Ny);
    cplx *ucor = (cplx *)malloc(sizeof(cplx) * (size_t)Ny);
    double *fre = (double *)malloc(sizeof(double) * (size_t)Ny);
#pragma omp for schedule(dynamic, 1)
    for (int ix = ixlo; ix <= ixhi; ix++)
      for (int iz = -nz; iz <= nz; iz++) {
        double al = o->p.alfa0 * ix, be = o->p.beta0 * iz;
        double k2 = (o->p.alfa0 * ix) * (o->p.alfa0 * ix) + (o->p.beta0 * iz) * (o->p.beta0 * iz);
        cplx *v1 = o->V + VIX(o, -1, iz, ix, 0); /* element iy=-1 of each component */
        cplx *v2 = o->V + VIX(o, -1, iz, ix, 1);
        cplx *v3 = o->V + VIX(o, -1, iz, ix, 2);
#define C1(i) v1[(i) + 1]
#define C2(i) v2[(i) + 1]
#define C3(i) v3[(i) + 1]
        memset(D2v, 0, sizeof(double[5]) * (size_t)(ny + 1));
        memset(eta, 0, sizeof(double[5]) * (size_t)(ny + 1));
        for (int iy = 1; iy <= ny - 1; iy++) {
          const Di *d = &o->der[iy];
          for (int j = 0; j < 5; j++) {
            double os = ni * (d->d4[j] - 2.0 * k2 * d->d2[j] + k2 * k2 * d->d0[j]);
            double sq = ni * (d->d2[j] - k2 * d->d0[j]);
            D2v[iy - 1][j] = lambda * (d->d2[j] - k2 * d->d0[j]) - os;
            eta[iy - 1][j] = lambda * d->d0[j] - sq;
          }
        }
        BCOND *b0 = &o->bc0[(size_t)ix * o->Nz + iz + nz], *bn = &o->bcn[(size_t)ix * o->Nz + iz + nz];
        /* bottom wall (first) */
        if (ix == 0 && iz == 0) {
          b0->v = 0;
          b0->vy = 0;
          b0->eta = CMPLX(creal(b0->u) - cimag(b0->w), cimag(b0->u) + creal(b0->w));
        } else {
          b0->vy = -imul(al, b0->u) - imul(be, b0->w);
          b0->eta = imul(be, b0->u) - imul(al, b0->w);
        }
        b0->v = b0->v - o->v0bc[0] * b0->vy / o->v0m1bc[0];
        applybc_0(D2v, o->v0bc, o->v0m1bc);
        C2(1) = C2(1) - D2v[0][0] * b0->vy / o->v0m1bc[0] - D2v[0][1] * b0->v / o->v0bc[1];
        C2(2) = C2(2) - D2v[1][0] * b0->v / o->v0bc[1];
        applybc_0(eta, o->eta0bc, o->eta0m1bc);
        C1(1) = C1(1) - eta[0][1] * b0->eta / o->eta0bc[1];
        C1(2) = C1(2) - eta[1][0] * b0->eta / o->eta0bc[1];
        /* top wall (last) */
        if (ix == 0 && iz == 0) {
          bn->v = 0;
          bn->vy = 0;
          bn->eta = CMPLX(creal(bn->u) - cimag(bn->w), cimag(bn->u) + creal(bn->w));
        } else {
          bn->vy = -imul(al, bn->u) - imul(be, bn->w);
          bn->eta = imul(be, bn->u) - imul(al, bn->w);
        }
        bn->v = bn->v - o->vnbc[4] * bn->vy / o->vnp1bc[4];
        applybc_n(D2v, ny, o->vnbc, o->vnp1bc);
        C2(ny - 1) = C2(ny - 1) - D2v[ny - 2][4] * bn->vy / o->vnp1bc[4] - D2v[ny - 2][3] * bn->v / o->vnbc[3];
        C2(ny - 2) = C2(ny - 2) - D2v[ny - 3][4] * bn->v / o->vnbc[3];
        applybc_n(eta, ny, o->etanbc, o->etanp1bc);
        C1(ny - 1) = C1(ny - 1) - eta[ny - 2][3] * bn->eta / o->etanbc[3];
        C1(ny - 2) = C1(ny - 2) - eta[ny - 3][4] * bn->eta / o->etanbc[3];
        /* UL factorisation and the two sweeps */
        lu5decomp(D2v, ny + 1);
        lu5decomp(eta, ny + 1);
        lu5div1(D2v, ny + 1, v2);
        lu5div1(eta, ny + 1, v1);
        lu5div2(D2v, ny + 1, v2);
        lu5div2(eta, ny + 1, v1);
        /* ghost and wall values (linsolve_blocking.inc:50-61) */
        {
          const double *a = o->v0bc, *b = o->v0m1bc, *e = o->eta0bc, *f = o->eta0m1bc;
          C2(0) = (b0->v - (C2(1) * a[2] + C2(2) * a[3] + C2(3) * a[4])) / a[1];
          C2(-1) = (b0->vy - (C2(0) * b[1] + C2(1) * b[2] + C2(2) * b[3] + C2(3) * b[4])) / b[0];
          C1(0) = (b0->eta - (C1(1) * e[2] + C1(2) * e[3] + C1(3) * e[4])) / e[1];
          C1(-1) = -((C1(0) * f[1] + C1(1) * f[2] + C1(2) * f[3] + C1(3) * f[4]) / f[0]);
        }
        {
          const double *a = o->vnbc, *b = o->vnp1bc, *e = o->etanbc, *f = o->etanp1bc;
          C2(ny) = (bn->v - (C2(ny - 3) * a[0] + C2(ny - 2) * a[1] + C2(ny - 1) * a[2])) / a[3];
          C2(ny + 1) = (bn->vy - (C2(ny - 3) * b[0] + C2(ny - 2) * b[1] + C2(ny - 1) * b[2] + C2(ny) * b[3])) / b[4];
          C1(ny) = (bn->eta - (C1(ny - 3) * e[0] + C1(ny - 2) * e[1] + C1(ny - 1) * e[2])) / e[3];
          C1(ny + 1) = -((C1(ny - 3) * f[0] + C1(ny - 2) * f[1] + C1(ny - 1) * f[2] + C1(ny) * f[3]) / f[4]);
        }
        if (ix == 0 && iz == 0) {
          /* mean flow (linsolve_blocking.inc:62-97) */
          for (int iy = -1; iy <= ny + 1; iy++) {
            C3(iy) = CMPLX(cimag(C1(iy)), 0.0);
            C1(iy) = CMPLX(creal(C1(iy)), 0.0);
          }
          ucor[0] = 0;
          ucor[1] = 0;
          for (int iy = 1; iy <= ny - 1; iy++) ucor[iy + 1] = 1;
          ucor[ny + 1] = 0;
          ucor[ny + 2] = 0;
          lu5div1(eta, ny + 1, ucor);
          lu5div2(eta, ny + 1, ucor);
#define UC(i) ucor[(i) + 1]
          {
            const double *e = o->eta0bc, *f = o->eta0m1bc;
            UC(0) = -((UC(1) * e[2] + UC(2) * e[3] + UC(3) * e[4]) / e[1]);
            UC(-1) = -((UC(0) * f[1] + UC(1) * f[2] + UC(2) * f[3] + UC(3) * f[4]) / f[0]);
          }
          {
            const double *e = o->etanbc, *f = o->etanp1bc;
            UC(ny) = -((UC(ny - 3) * e[0] + UC(ny - 2) * e[1] + UC(ny - 1) * e[2]) / e[3]);
            UC(ny + 1) = -((UC(ny - 3) * f[0] + UC(ny - 2) * f[1] + UC(ny - 1) * f[2] + UC(ny) * f[3]) / f[4]);
          }
          for (int iy = -1; iy <= ny + 1; iy++) fre[iy + 1] = creal(C1(iy));
          o->fr[0] = yintegr(o, fre);
          for (int iy = -1; iy <= ny + 1; iy++) fre[iy + 1] = creal(C3(iy));
          o->fr[1] = yintegr(o, fre);
          for (int iy = -1; iy <= ny + 1; iy++) fre[iy + 1] = creal(UC(iy));
          o->fr[2] = yintegr(o, fre);
          if (fabs(o->p.meanflowx) > 1.0e-7 && !o->p.cpi) {
            o->corrpx = (o->p.meanflowx - o->fr[0]) / o->fr[2];
            for (int iy = -1; iy <= ny + 1; iy++) C1(iy) = CMPLX(creal(C1(iy)) + o->corrpx * creal(UC(iy)), cimag(C1(iy)));
          }
          if (fabs(o->p.meanflowz) > 1.0e-7 && !o->p.cpi) {
            o->corrpz = (o->p.meanflowz - o->fr[1]) / o->fr[2];
            for (int iy = -1; iy <= ny + 1; iy++) C3(iy) = CMPLX(creal(C3(iy)) + o->corrpz * creal(UC(iy)), cimag(C3(iy)));
          }
          if (o->p.cpi) {
            if (o->p.cpi_type == 0)
              o->meanpx = (1 - o->p.gamma) * 6 * ni / o->fr[0];
            else
              o->meanpx = (1.5 / o->p.gamma) * o->fr[0] * ni;
          }
#undef UC
        } else {
          /* dv/dy and u,w (linsolve_blocking.inc:98-104) */
          complexderiv(o, v2, v3);
          lu5div2(o->D0mat, ny + 1, v3);
          for (int iy = -1; iy <= ny + 1; iy++) {
            cplx t = imul(al, C3(iy)) - imul(be, C1(iy));
            temp[iy + 1] = CMPLX(creal(t) / k2, cimag(t) / k2);
            cplx w = imul(be, C3(iy)) + imul(al, C1(iy));
            C3(iy) = CMPLX(creal(w) / k2, cimag(w) / k2);
            C1(iy) = temp[iy + 1];
          }
        }
#undef C1
#undef C2
#undef C3
      }
    free(D2v);
    free(eta);
    free(temp);
    free(ucor);
    free(fre);
  }
}

void orc_linsolve(orc_t *o, double lambda) { linsolve_cols(o, lambda, 0, o->nx); }

static double wall_s(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/* Wall-clock cost of the three parts of a substep measured on a bounded
 * sample of the workload: out[0] seconds per plane of products (planes
 * 1..nplanes), out[1] seconds per row of rhs assembly (rows 3..nplanes-2),
 * out[2] seconds per x-mode column of the implicit solve (ix = 1..nix).  The
 * state is left modified: call it on a scratch instance. */
void orc_time_sample(orc_t *o, int nplanes, int nix, double out[3]) {
  static const double ode[3] = {120.0 / 32.0, 2.0, 0.0};
  if (o->deltat == 0) o->deltat = 1e-3;
  double t0 = wall_s();
  all_products(o, 0, 1, nplanes);
  double t1 = wall_s();
  cplx *newrhs = (cplx *)malloc(sizeof(cplx) * (size_t)(o->ny - 1) * o->Nx * o->Nz * 2);
  rhs_rows(o, ode, 3, nplanes - 2, newrhs);
  double t2 = wall_s();
  free(newrhs);
  linsolve_cols(o, ode[0] / o->deltat, 1, nix);
  double t3 = wall_s();
  out[0] = (t1 - t0) / nplanes;
  out[1] = (t2 - t1) / (nplanes - 4);
  out[2] = (t3 - t2) / nix;
}

/* channel.f90:95-115: pre-loop CFL pass and initial flow rate / CPI */
void orc_initial(orc_t *o) {
  int ny = o->ny;
  if (o->deltat == 0) o->deltat = 1.0;
  /* products of the interior planes are computed and discarded; only the
   * CFL maximum survives */
  all_products(o, 1, 1, ny - 1);
  double *f = (double *)malloc(sizeof(double) * (size_t)o->Ny);
  for (int iy = -1; iy <= ny + 1; iy++) f[iy + 1] = creal(o->V[VIX(o, iy, 0, 0, 0)]);
  o->fr[0] = yintegr(o, f);
  for (int iy = -1; iy <= ny + 1; iy++) f[iy + 1] = creal(o->V[VIX(o, iy, 0, 0, 2)]);
  o->fr[1] = yintegr(o, f);
  o->fr[2] = 0.0; /* frl(3) is never assigned before this Allreduce */
  free(f);
  if (o->p.cpi) {
    if (o->p.cpi_type == 0)
      o->meanpx = (1 - o->p.gamma) * 6 * o->p.ni / o->fr[0];
    else
      o->meanpx = (1.5 / o->p.gamma) * o->fr[0] * o->p.ni;
  }
}

/* outstats (dnsdata.f90:853-881): CFL -> deltat, wall gradients, one
 * Runtimedata row */
void orc_outstats(orc_t *o, double row[11]) {
  int ny = o->ny;
  double rg = o->cfl;
  o->cfl = 0;
  if (o->p.cflmax > 0) o->deltat = o->p.cflmax / rg;
  double d11 = 0, d21 = 0, d12 = 0, d22 = 0;
  for (int j = 0; j < 5; j++) {
    d12 += o->d14n[j] * creal(o->V[VIX(o, ny - 3 + j, 0, 0, 0)]);
    d22 += o->d14n[j] * creal(o->V[VIX(o, ny - 3 + j, 0, 0, 2)]);
  }
  d12 = -d12;
  d22 = -d22;
  for (int j = 0; j < 5; j++) {
    d11 += o->d140[j] * creal(o->V[VIX(o, -1 + j, 0, 0, 0)]);
    d21 += o->d140[j] * creal(o->V[VIX(o, -1 + j, 0, 0, 2)]);
  }
  row[0] = o->time;
  row[1] = d11;
  row[2] = d12;
  row[3] = d21;
  row[4] = d22;
  row[5] = o->fr[0] + o->corrpx * o->fr[2];
  row[6] = o->meanpx + o->corrpx;
  row[7] = o->fr[1] + o->corrpz * o->fr[2];
  row[8] = o->meanpz + o->corrpz;
  row[9] = rg * o->deltat;
  row[10] = o->deltat;
}

static const double RK1[3] = {120.0 / 32.0, 2.0, 0.0};
static const double RK2[3] = {120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0};
static const double RK3[3] = {120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0};

/* one iteration of the time loop (channel.f90:122-167) */
void orc_step(orc_t *o, double row[11]) {
  o->bc0[0 * o->Nz + o->nz].u = o->p.u0;
  o->bcn[0 * o->Nz + o->nz].u = o->p.uN;
  o->time = o->time + 2.0 / RK1[0] * o->deltat;
  orc_buildrhs(o, RK1, 0);
  orc_linsolve(o, RK1[0] / o->deltat);
  o->time = o->time + 2.0 / RK2[0] * o->deltat;
  orc_buildrhs(o, RK2, 0);
  orc_linsolve(o, RK2[0] / o->deltat);
  o->time = o->time + 2.0 / RK3[0] * o->deltat;
  orc_buildrhs(o, RK3, 1);
  orc_linsolve(o, RK3[0] / o->deltat);
  orc_outstats(o, row);
}

void orc_steps(orc_t *o, int n, double *rows) {
  for (int s = 0; s < n; s++) orc_step(o, rows + 11 * (size_t)s);
}
