// setup.hpp -- the host-side set-up of the time step (setup.cpp).  Plain
// C++ with no HIP dependency, so that setup.cpp also builds with a host
// compiler alone (the sanitizer build of the CPU tests).
#pragma once

namespace chan {

// Set-up products computed on the host (setup.cpp) and mirrored to device.
struct HostSetup {
  int nx, ny, nz, nxd, nzd;
  double PI, dx, dz, factor;
  double *y;      // [ny+3], y[iy+1]
  double *dy;     // [ny+1], dy[iy] for iy=1..ny-1
  double *der;    // [(ny-1)][4][5]: d0,d1,d2,d4 of rows iy=1..ny-1
  double *d0fac;  // [(ny+1)][5]: D0 band after the UL factorisation
  double bnd[10][5];  // d040,d140,d240,d14m1,d24m1,d04n,d14n,d24n,d14np1,d24np1
  double bc[8][5];    // v0bc,v0m1bc,eta0bc,eta0m1bc,vnbc,vnp1bc,etanbc,etanp1bc
};
enum { B_D040, B_D140, B_D240, B_D14M1, B_D24M1, B_D04N, B_D14N, B_D24N, B_D14NP1, B_D24NP1 };
enum { BC_V0, BC_V0M1, BC_ETA0, BC_ETA0M1, BC_VN, BC_VNP1, BC_ETAN, BC_ETANP1 };

int fftfit(int i);
void setup_build(HostSetup &s, int nx, int ny, int nz, double alfa0, double beta0, double a, double ymin, double ymax);
void setup_free(HostSetup &s);

}  // namespace chan
