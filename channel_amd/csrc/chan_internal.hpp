// chan_internal.hpp -- state and kernel interfaces shared by the library's
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "fft.hpp"
#include "setup.hpp"

namespace chan {

// Intermediate layouts of the transforms use blocks of ZB consecutive
// physical-z points so that the z-transform kernels read and write ZB*16
// contiguous bytes per x-mode.
constexpr int ZB = 4;

// Geometry passed by value to kernels.
struct Geom {
  int nx, ny, nz;   // dns.in sizes
  int Nx, Nz, Ny;   // nx+1, 2nz+1, ny+3
  int nxd, nzd;     // dealiased sizes, physical grid 2nxd x nzd
  int npair;        // Nx*Nz wavenumber pairs, pair p = ix*Nz + (iz+nz)
  size_t ps;        // plane stride (pairs, padded)
  int ncu;          // compute units of the device (persistent grids)
  double alfa0, beta0;
};

// Scalars that live on the device between kernels (the (0,0) mode writes
// them in linsolve, buildrhs reads meanpx/meanpz).
struct DevScalars {
  double meanpx, meanpz, corrpx, corrpz;
  double fr[3];
  unsigned long long cfl_bits;  // running CFL maximum as the bits of a non-negative double
  double pad[8];
};

// ---------------------------------------------------------- kernel args --
template <typename R>
struct RhsArgs {
  const cx<R> *P;        // [6][Ny][ps] products
  const cx<double> *V;   // [3][Ny][ps]
  cx<R> *old;            // [2][ny-1][ps] previous explicit terms (eta, d2v); in the split
                         // mode FP32, as the products they are formed from
  cx<double> *rhs;       // [2][ny-1][ps]
  const double *der;     // [(ny-1)][4][5]
  const DevScalars *sc;  // meanpx, meanpz
  bool force;            // rhs holds the body-force terms (k_force_terms) on entry
  Geom g;
  double ode0, ode1, ode2, dt, ni;
  int chunk;
};

// Body force F_d = F0_d + gain[d][plane][col] * V_src[d]; either part may be
// absent (F0 null, src[d] < 0).  col = iz+nz, or Nz for the mode (0,0).
struct ForceArgs {
  const cx<double> *V, *F0;  // [3][Ny][ps]; F0 may be null
  const double *gain, *der;  // gain [3][Ny][Nz+1]
  int src[3];
  cx<double> *out;  // k_force: [3][Ny][ps]; k_force_terms: [2][ny-1][ps]
  Geom g;
  int chunk;
};

struct SolveArgs {
  cx<double> *V;           // [3][Ny][ps]; components 1,2 receive eta and v, pair
                           // slot npair of component 1 the ucor response
  const cx<double> *rhs;   // [2][ny-1][ps]
  double2 *L;              // [2][ny-1][gs] lower factors of the UL decomposition, one
                           // row per (kx, |kz|) group, gs <= ps
  const double *der;
  double bc[8][5];
  Geom g;
  double lambda, ni, u0, uN;
};

struct MeanArgs {
  cx<double> *V;
  const double *y;
  DevScalars *sc;
  double bc[8][5];
  Geom g;
  double meanflowx, meanflowz, gamma, ni;
  int cpi, cpi_type;
};

struct UwArgs {
  cx<double> *V;
  const double *der;
  const double *d0fac;
  double bnd[10][5];
  Geom g;
};

// records the message of a failing C-ABI call (chan_last_error) and returns code
int set_error(int code, const char *fmt, ...);

// ----------------------------------------------------------- launchers --
int fft_size_supported(int N);
template <typename R>
hipError_t launch_zinv(int NZ, hipStream_t st, const cx<double> *V, cx<R> *A, const cx<R> *tw, const Geom &g, int plane0,
                       int nplanes);
template <typename R>
hipError_t launch_xpass(int M, hipStream_t st, const cx<R> *A, cx<R> *B, const cx<R> *tw, const Geom &g, int plane0,
                        int nplanes, bool products, int cfl_on, const double *dy, double dx, double dz, R factor,
                        unsigned long long *cfl_bits);
template <typename R>
hipError_t launch_zfwd(int NZ, hipStream_t st, const cx<R> *B, cx<R> *P, const cx<R> *tw, const Geom &g, int plane0,
                       int nplanes);
hipError_t launch_fft_test(int N, int sign, hipStream_t st, cx<double> *data, const cx<double> *tw, int batch);
hipError_t launch_fs768_test(const cx<double> *in, cx<double> *out, const cx<double> *tw, int batch, int sign, int inner);
hipError_t launch_fft_bench(int N, int sign, int variant, const cx<double> *in, cx<double> *out, const cx<double> *tw,
                            int batch, int inner);
template <typename R>
hipError_t launch_rhs(hipStream_t st, const RhsArgs<R> &a);
hipError_t launch_force(hipStream_t st, const ForceArgs &a);
hipError_t launch_force_terms(hipStream_t st, const ForceArgs &a);
hipError_t launch_stats(hipStream_t st, const cx<double> *V, const double *avg, double *out, const Geom &g);
hipError_t launch_solve(hipStream_t st, const SolveArgs &a);
hipError_t launch_mean(hipStream_t st, const MeanArgs &a);
hipError_t launch_uw(hipStream_t st, const UwArgs &a);
hipError_t launch_flowrate(hipStream_t st, const MeanArgs &a);
hipError_t launch_to_device_layout(hipStream_t st, const cx<double> *in, cx<double> *out, const Geom &g);
hipError_t launch_to_host_layout(hipStream_t st, const cx<double> *in, cx<double> *out, const Geom &g);

}  // namespace chan
