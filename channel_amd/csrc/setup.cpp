// setup.cpp -- host-side FP64 set-up of the time step: dealiased sizes, the
// stretched y grid, the compact fourth-order y-derivative stencils, the wall
// and ghost closures, the boundary-condition rows and the UL-factorised D0
// band.  Everything here runs once per chan_create.
//
// Reference: read_dnsin (dnsdata.f90:110-124), init_memory (:153-158),
// setup_derivatives (:241-286) with LUdecomp / .bs. (rbmat.f90:60-76,
// 201-215), setup_boundary_conditions (:290-308), LU5decompStep
// (rbparmat_blocking.f90:20-51).
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "setup.hpp"

namespace chan {

// fftFIT (ffts.f90:78-86): sizes of the form 2^k or 3*2^k.
int fftfit(int i) {
  if (i <= 0) return 0;
  while ((i & 1) == 0) i >>= 1;
  return i == 1 || i == 3;
}

namespace {

// 5x5 solve used for the stencil coefficients.  The elimination runs from the
// last row upwards without pivoting and stores reciprocal pivots, and the
// solve is a backward then forward sweep: the order of rbmat.f90's LUdecomp
// and LLUdiv, kept so the coefficients round the same way.
struct Dense5 {
  double a[5][5];
  void factor() {
    for (int i = 4; i >= 1; --i) {
      const double r = 1.0 / a[i][i];
      a[i][i] = r;
      for (int j = 0; j < i; ++j) a[i][j] *= r;
      for (int k = 0; k < i; ++k) {
        const double f = a[k][i];
        for (int j = 0; j < i; ++j) a[k][j] -= f * a[i][j];
      }
    }
    a[0][0] = 1.0 / a[0][0];
  }
  void solve(const double b[5], double x[5]) const {
    x[4] = b[4] * a[4][4];
    for (int i = 3; i >= 0; --i) {
      double s = 0.0;
      for (int j = i + 1; j < 5; ++j) s += a[i][j] * x[j];
      x[i] = (b[i] - s) * a[i][i];
    }
    for (int i = 1; i < 5; ++i) {
      double s = 0.0;
      for (int j = 0; j < i; ++j) s += a[i][j] * x[j];
      x[i] -= s;
    }
  }
};

// Vandermonde-type system for a 5-point stencil on offsets h[0..4]:
// row i holds scale_i * h_j^(4-i).
void vander(Dense5 &M, const double h[5], bool d0rows) {
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      const double p = std::pow(h[j], 4.0 - i);
      M.a[i][j] = d0rows ? (5.0 - i) * (6.0 - i) * (7.0 - i) * (8.0 - i) * p : p;
    }
  M.factor();
}

// Point stencil at node c of the five nodes y[k0..k0+4] for derivative order
// `order` (1 or 2): the right-hand side has a single nonzero, order! at the
// row of power `order`.
void point_stencil(const double *y, int k0, int c, int order, double out[5]) {
  double h[5], t[5] = {0, 0, 0, 0, 0};
  for (int j = 0; j < 5; ++j) h[j] = y[k0 + j] - y[c];
  Dense5 M;
  vander(M, h, false);
  t[4 - order] = order == 1 ? 1.0 : 2.0;
  M.solve(t, out);
}

}  // namespace

void setup_build(HostSetup &s, int nx, int ny, int nz, double alfa0, double beta0, double a, double ymin, double ymax) {
  std::memset(&s, 0, sizeof s);
  s.nx = nx;
  s.ny = ny;
  s.nz = nz;
  s.nxd = 3 * (nx + 1) / 2;
  s.nzd = 3 * nz;
  while (!fftfit(s.nxd)) ++s.nxd;
  while (!fftfit(s.nzd)) ++s.nzd;
  // PI is a default-kind (single precision) literal in the reference
  // (dnsdata.f90:26), so dx, dz and the CFL number carry float(pi).
  s.PI = (double)(float)3.14159265358979323846;
  s.dx = s.PI / (alfa0 * s.nxd);
  s.dz = 2.0 * s.PI / (beta0 * s.nzd);
  s.factor = 1.0 / (2.0 * s.nxd * s.nzd);

  // y(iy) for iy = -1..ny+1; 2*iy/ny is evaluated in double precision
  s.y = (double *)std::calloc(ny + 3, sizeof(double));
  s.dy = (double *)std::calloc(ny + 1, sizeof(double));
  const double ta = std::tanh(a);
  for (int iy = -1; iy <= ny + 1; ++iy) {
    const double r = 2.0 * (double)iy / (double)ny;
    s.y[iy + 1] = ymin + 0.5 * (ymax - ymin) * (std::tanh(a * (r - 1)) / ta + 1);
  }
  const double *Y = s.y + 1;  // Y[iy]
  for (int iy = 1; iy <= ny - 1; ++iy) s.dy[iy] = 0.5 * (Y[iy + 1] - Y[iy - 1]);

  // interior compact stencils
  s.der = (double *)std::calloc((size_t)(ny - 1) * 20, sizeof(double));
  for (int iy = 1; iy <= ny - 1; ++iy) {
    double *d0 = s.der + (size_t)(iy - 1) * 20, *d1 = d0 + 5, *d2 = d0 + 10, *d4 = d0 + 15;
    double h[5], t[5];
    for (int j = 0; j < 5; ++j) h[j] = Y[iy - 2 + j] - Y[iy];
    Dense5 M;
    // d4: exact fourth derivative on quartics
    vander(M, h, false);
    double e4[5] = {24, 0, 0, 0, 0};
    M.solve(e4, d4);
    // d0: D0[f''''] = D4[f] for f = h^4 .. h^8
    vander(M, h, true);
    for (int i = 0; i < 5; ++i) {
      double acc = 0.0;
      for (int j = 0; j < 5; ++j) acc += d4[j] * std::pow(h[j], 8.0 - i);
      t[i] = acc;
    }
    M.solve(t, d0);
    // d2, d1: D2[f] = D0[f''], D1[f] = D0[f'] on quartics
    vander(M, h, false);
    for (int i = 0; i < 5; ++i) {
      double acc = 0.0;
      if (i <= 2)
        for (int j = 0; j < 5; ++j) acc += d0[j] * (4.0 - i) * (3.0 - i) * std::pow(h[j], 2.0 - i);
      t[i] = acc;
    }
    M.solve(t, d2);
    for (int i = 0; i < 5; ++i) {
      double acc = 0.0;
      if (i <= 3)
        for (int j = 0; j < 5; ++j) acc += d0[j] * (4.0 - i) * std::pow(h[j], 3.0 - i);
      t[i] = acc;
    }
    M.solve(t, d1);
  }

  // one-sided wall and ghost stencils on y(-1..3) and y(ny-3..ny+1)
  point_stencil(Y, -1, 0, 1, s.bnd[B_D140]);
  point_stencil(Y, -1, 0, 2, s.bnd[B_D240]);
  point_stencil(Y, -1, -1, 1, s.bnd[B_D14M1]);
  point_stencil(Y, -1, -1, 2, s.bnd[B_D24M1]);
  point_stencil(Y, ny - 3, ny, 1, s.bnd[B_D14N]);
  point_stencil(Y, ny - 3, ny, 2, s.bnd[B_D24N]);
  point_stencil(Y, ny - 3, ny + 1, 1, s.bnd[B_D14NP1]);
  point_stencil(Y, ny - 3, ny + 1, 2, s.bnd[B_D24NP1]);
  s.bnd[B_D040][1] = 1.0;
  s.bnd[B_D04N][3] = 1.0;

  // boundary-condition rows with the ghost eliminated (full channel)
  auto elim = [](double *row, const double *ghost, int g, int lo, int hi) {
    const double f = row[g];
    for (int j = lo; j <= hi; ++j) row[j] = row[j] - f * ghost[j] / ghost[g];
  };
  std::memcpy(s.bc[BC_V0], s.bnd[B_D040], 40);
  std::memcpy(s.bc[BC_V0M1], s.bnd[B_D140], 40);
  std::memcpy(s.bc[BC_ETA0], s.bnd[B_D040], 40);
  std::memcpy(s.bc[BC_ETA0M1], s.der + 15, 40);  // der(1)%d4
  elim(s.bc[BC_V0], s.bc[BC_V0M1], 0, 1, 4);
  elim(s.bc[BC_ETA0], s.bc[BC_ETA0M1], 0, 1, 4);
  std::memcpy(s.bc[BC_VN], s.bnd[B_D04N], 40);
  std::memcpy(s.bc[BC_VNP1], s.bnd[B_D14N], 40);
  std::memcpy(s.bc[BC_ETAN], s.bnd[B_D04N], 40);
  std::memcpy(s.bc[BC_ETANP1], s.der + (size_t)(ny - 2) * 20 + 15, 40);  // der(ny-1)%d4
  elim(s.bc[BC_VN], s.bc[BC_VNP1], 4, 0, 3);
  elim(s.bc[BC_ETAN], s.bc[BC_ETANP1], 4, 0, 3);

  // D0 band rows iy=1..ny-1 (rows ny, ny+1 zero), UL-factorised bottom-up
  // with reciprocal pivots; wall couplings are zeroed as the single-slab
  // LU5decompStep does.
  const int rows = ny + 1;
  s.d0fac = (double *)std::calloc((size_t)rows * 5, sizeof(double));
  double(*A)[5] = (double(*)[5])s.d0fac;
  for (int iy = 1; iy <= ny - 1; ++iy)
    for (int j = 0; j < 5; ++j) A[iy - 1][j] = s.der[(size_t)(iy - 1) * 20 + j];
  const int top = rows - 1;
  A[top - 2][3] = A[top - 2][4] = 0.0;
  A[top - 3][4] = 0.0;
  for (int i = top - 2; i >= 0; --i) {
    double p2 = A[i][4];
    A[i][3] -= p2 * A[i + 2][1];
    A[i][2] -= p2 * A[i + 2][0];
    double p1 = A[i][3];
    A[i][2] -= p1 * A[i + 1][1];
    A[i][1] -= p1 * A[i + 1][0];
    const double r = 1.0 / A[i][2];
    A[i][2] = r;
    A[i][0] *= r;
    A[i][1] *= r;
  }
  A[0][0] = A[0][1] = 0.0;
  A[1][0] = 0.0;
}

void setup_free(HostSetup &s) {
  std::free(s.y);
  std::free(s.dy);
  std::free(s.der);
  std::free(s.d0fac);
  s.y = s.dy = s.der = s.d0fac = nullptr;
}

}  // namespace chan
