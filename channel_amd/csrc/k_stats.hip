// k_stats.hip -- wall-normal statistics profiles of the current field
// (postpro/history/var_history.f90:36-87): plane means, variances about a
// given mean and the Reynolds shear stress, as spectral sums over the
// Hermitian half plane (weight 1 at kx = 0, 2 otherwise: Parseval).
//
// One block per row iy = 0..ny reads the three components' rows (contiguous
// over pairs) once and reduces in LDS in a fixed order, so results are
// reproducible from run to run.
#include "chan_internal.hpp"

namespace chan {

__global__ __launch_bounds__(256) void k_stats(const cx<double> *__restrict__ V, const double *__restrict__ avg,
                                               double *__restrict__ out, Geom g) {
  __shared__ double red[6][256];
  const int iy = blockIdx.x;  // 0..ny
  const int r = iy + 1;       // row index in the device layout
  const size_t ps = g.ps, cs = (size_t)g.Ny * ps;
  const int n1 = g.ny + 1;
  const cx<double> *u = V + (size_t)r * ps, *v = u + cs, *w = v + cs;
  // the mean subtracted from the (0,0) mode: the given average, else the
  // field's own plane mean
  double m[3];
#pragma unroll
  for (int c = 0; c < 3; c++) m[c] = avg ? avg[c * n1 + iy] : V[c * cs + (size_t)r * ps + g.nz].x;
  double s[4] = {0.0, 0.0, 0.0, 0.0};  // var u, var v, var w, uv
  for (int p = threadIdx.x; p < g.npair; p += 256) {
    const double f = p < g.Nz ? 1.0 : 2.0;  // ix = 0 for the first Nz pairs
    cx<double> a = u[p], b = v[p], c = w[p];
    if (p == g.nz) {
      a.x -= m[0];
      b.x -= m[1];
      c.x -= m[2];
    }
    s[0] += f * (a.x * a.x + a.y * a.y);
    s[1] += f * (b.x * b.x + b.y * b.y);
    s[2] += f * (c.x * c.x + c.y * c.y);
    s[3] += f * (a.x * b.x + a.y * b.y);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 3; c++) out[c * n1 + iy] = V[c * cs + (size_t)r * ps + g.nz].x;
#pragma unroll
    for (int k = 0; k < 4; k++) out[(3 + k) * n1 + iy] = red[k][0];
  }
}

hipError_t launch_stats(hipStream_t st, const cx<double> *V, const double *avg, double *out, const Geom &g) {
  hipLaunchKernelGGL(k_stats, dim3((unsigned)(g.ny + 1)), dim3(256), 0, st, V, avg, out, g);
  return hipGetLastError();
}

}  // namespace chan
