// k_les.hip -- the learned subgrid-stress closure inside a coarse channel
// step, on the device (include/channel_closure.h): closure features of the
// resident field, the MLP, and the body force -div(tau) written straight
// into the context's static force, so that an LES step with the closure
// never leaves the GPU.  The reference has no subgrid model.
//
// Grid: rows iy = 0..ny of the dealiased physical grid Mx x Mz = 2nxd x nzd,
// x_j = j Lx / Mx, z_k = k Lz / Mz.  Spectral <-> physical transforms here
// are separable direct sums in FP32 (the LES grids are small: at 31x64x31 a
// transform of one row is 32 x 63 modes <-> 96 x 96 points); x and z
// derivatives are spectral, y derivatives the second-order differences of the
// stretched grid (one-sided three-point rows at the walls), applied to the
// spectral coefficients.
//
// Features of sample (iy, x, z), 46 values: the nine gradients du_c/dx_d
// (index 3c + d, d = x, y, z) at rows iy, iy-2, iy-1, iy+1, iy+2 (rows
// clamped to 0..ny), then y+ = min(y - ymin, ymax - y) u_tau / nu with u_tau
// from the mean wall shear of both walls (the wall-gradient stencils of
// outstats, dnsdata.f90:861-871).  Stresses: tau = (xx, yy, zz, xy, xz, yz).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/channel_closure.h"
#include "ctx_internal.hpp"
#include "mlp_internal.hpp"

namespace chan {

struct LesState {
  int n = 0;                  // samples (ny+1) Mx Mz
  float *Gp = nullptr;        // [n][9] physical gradients
  float *Xf = nullptr;        // [n][46] features (host query)
  __bf16 *Xb = nullptr;       // [n][64] normalised BF16 MLP inputs
  float *Y = nullptr;         // [n][6] MLP outputs
  float2 *T = nullptr;        // [6][ny+1][npair] spectral stresses
  float2 *twz = nullptr;      // [Mz] exp(+2 pi i k / Mz)
  float2 *twx = nullptr;      // [Mx] exp(+2 pi i k / Mx)
  double *ydc = nullptr;      // [ny+1][3] y-difference coefficients of rows iy-1, iy, iy+1
  int *ydr = nullptr;         // [ny+1] first row of each stencil
  double *yplus = nullptr;    // [ny+1] wall distance in wall units
  float *norm = nullptr;      // [46 xmean][46 xscale][6 tmean][6 tscale]
  double *wall = nullptr;     // [2][5] wall-gradient stencils d140, d14n
};

void les_free(LesState *s) {
  if (!s) return;
  void *b[] = {s->Gp, s->Xf, s->Xb, s->Y, s->T, s->twz, s->twx, s->ydc, s->ydr, s->yplus, s->norm, s->wall};
  for (void *p : b)
    if (p) (void)hipFree(p);
  delete s;
}

namespace {

struct LesGeom {
  int nx, nz, ny, Nx, Nz, Ny, Mx, Mz, npair;
  size_t ps;
  double alfa0, beta0;
};

LesGeom les_geom(const chan_ctx *c) {
  LesGeom q;
  q.nx = c->g.nx;
  q.nz = c->g.nz;
  q.ny = c->g.ny;
  q.Nx = c->g.Nx;
  q.Nz = c->g.Nz;
  q.Ny = c->g.Ny;
  q.Mx = 2 * c->g.nxd;
  q.Mz = c->g.nzd;
  q.npair = c->g.npair;
  q.ps = c->g.ps;
  q.alfa0 = c->g.alfa0;
  q.beta0 = c->g.beta0;
  return q;
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// The three derivatives (d = x, y, z) of component c on row iy, in physical
// space: Gp[(iy Mx + x) Mz + z][3c + d].  One block per (row, component);
// LDS holds the spectral coefficients of one derivative and its z transform.
__global__ __launch_bounds__(256) void k_les_grad(const cx<double> *__restrict__ V, LesGeom q,
                                                  const double *__restrict__ ydc, const int *__restrict__ ydr,
                                                  const float2 *__restrict__ twz, const float2 *__restrict__ twx,
                                                  float *__restrict__ Gp) {
  extern __shared__ float2 sm[];
  float2 *S = sm;             // [npair]
  float2 *H = sm + q.npair;   // [Nx][Mz]
  const int iy = blockIdx.x, c = blockIdx.y;
  const size_t cs = (size_t)c * q.Ny * q.ps;
  for (int d = 0; d < 3; d++) {
    for (int p = threadIdx.x; p < q.npair; p += 256) {
      const int ix = p / q.Nz, iz = p % q.Nz - q.nz;
      float2 s;
      if (d == 1) {
        const int r0 = ydr[iy];
        double re = 0.0, im = 0.0;
        for (int k = 0; k < 3; k++) {
          const cx<double> v = V[cs + (size_t)(r0 + k + 1) * q.ps + p];
          re += ydc[3 * iy + k] * v.x;
          im += ydc[3 * iy + k] * v.y;
        }
        s = make_float2((float)re, (float)im);
      } else {
        const cx<double> v = V[cs + (size_t)(iy + 1) * q.ps + p];
        const double k = d == 0 ? q.alfa0 * ix : q.beta0 * iz;  // i k v
        s = make_float2((float)(-k * v.y), (float)(k * v.x));
      }
      S[p] = s;
    }
    __syncthreads();
    // z: H[ix][zk] = sum_iz S e^{+2 pi i iz zk / Mz}
    for (int o = threadIdx.x; o < q.Nx * q.Mz; o += 256) {
      const int ix = o / q.Mz, zk = o % q.Mz;
      float2 acc = make_float2(0.f, 0.f);
      for (int j = 0; j < q.Nz; j++) {
        const int iz = j - q.nz;
        const int e = (int)(((long)((iz % q.Mz) + q.Mz) * zk) % q.Mz);
        const float2 t = cmulf(S[ix * q.Nz + j], twz[e]);
        acc.x += t.x;
        acc.y += t.y;
      }
      H[o] = acc;
    }
    __syncthreads();
    // x: Hermitian sum over kx
    for (int o = threadIdx.x; o < q.Mx * q.Mz; o += 256) {
      const int xj = o / q.Mz, zk = o % q.Mz;
      float f = H[zk].x, g = 0.f;
      for (int ix = 1; ix < q.Nx; ix++) {
        const float2 t = cmulf(H[ix * q.Mz + zk], twx[(ix * xj) % q.Mx]);
        g += t.x;
      }
      f += 2.f * g;
      Gp[((size_t)iy * q.Mx * q.Mz + o) * 9 + 3 * c + d] = f;
    }
    __syncthreads();
  }
}

// u_tau from the mean wall shear of both walls, y+ of every row
__global__ void k_les_yplus(const cx<double> *__restrict__ V, LesGeom q, const double *__restrict__ wall,
                            const double *__restrict__ y, double ymin, double ymax, double ni,
                            double *__restrict__ yplus) {
  double db = 0.0, dt = 0.0;
  const int p0 = q.nz;  // pair (0, 0), component u
  for (int j = 0; j < 5; j++) {
    db += wall[j] * V[(size_t)j * q.ps + p0].x;                      // rows -1..3
    dt += wall[5 + j] * V[(size_t)(q.ny - 2 + j) * q.ps + p0].x;     // rows ny-3..ny+1
  }
  const double utau = sqrt(ni * 0.5 * (fabs(db) + fabs(dt)));
  for (int iy = threadIdx.x; iy <= q.ny; iy += blockDim.x) {
    const double yy = y[iy + 1];
    yplus[iy] = fmin(yy - ymin, ymax - yy) * utau / ni;
  }
}

// One thread per sample: the 46 features, raw into Xf (when given) and
// normalised BF16 into Xb (when given)
__global__ __launch_bounds__(256) void k_les_features(const float *__restrict__ Gp, LesGeom q,
                                                      const double *__restrict__ yplus, const float *__restrict__ norm,
                                                      float *__restrict__ Xf, __bf16 *__restrict__ Xb) {
  const int plane = q.Mx * q.Mz;
  const int n = (q.ny + 1) * plane;
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const int iy = s / plane, o = s % plane;
  const int rows[5] = {iy, iy - 2, iy - 1, iy + 1, iy + 2};
  float f[mlp::KP];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int r = min(max(rows[k], 0), q.ny);
    const float *g = Gp + ((size_t)r * plane + o) * 9;
#pragma unroll
    for (int j = 0; j < 9; j++) f[9 * k + j] = g[j];
  }
  f[45] = (float)yplus[iy];
  if (Xf)
    for (int j = 0; j < 46; j++) Xf[(size_t)s * 46 + j] = f[j];
  if (Xb) {
#pragma unroll
    for (int j = 0; j < mlp::KP; j++)
      Xb[(size_t)s * mlp::KP + j] = (__bf16)(j < 46 ? (f[j] - norm[j]) / norm[46 + j] : 0.0f);
  }
}

// Spectral stresses of row iy, component qq: T = (1/(Mx Mz)) sum tau e^{-i(kx x + kz z)}
__global__ __launch_bounds__(256) void k_les_tau(const float *__restrict__ Y, LesGeom q, const float *__restrict__ norm,
                                                 const float2 *__restrict__ twz, const float2 *__restrict__ twx,
                                                 float2 *__restrict__ T) {
  extern __shared__ float2 sm[];
  float *P = reinterpret_cast<float *>(sm);        // [Mx][Mz]
  float2 *H = sm + (q.Mx * q.Mz + 1) / 2;         // [Nx][Mz]
  const int iy = blockIdx.x, qq = blockIdx.y;
  const int plane = q.Mx * q.Mz;
  const float tm = norm[92 + qq], ts = norm[98 + qq];
  for (int o = threadIdx.x; o < plane; o += 256) P[o] = Y[((size_t)iy * plane + o) * 6 + qq] * ts + tm;
  __syncthreads();
  for (int o = threadIdx.x; o < q.Nx * q.Mz; o += 256) {
    const int ix = o / q.Mz, zk = o % q.Mz;
    float2 acc = make_float2(0.f, 0.f);
    for (int xj = 0; xj < q.Mx; xj++) {
      const float2 w = twx[(ix * xj) % q.Mx];
      const float v = P[xj * q.Mz + zk];
      acc.x += v * w.x;
      acc.y -= v * w.y;  // conj
    }
    H[o] = make_float2(acc.x / q.Mx, acc.y / q.Mx);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < q.npair; p += 256) {
    const int ix = p / q.Nz, iz = p % q.Nz - q.nz;
    float2 acc = make_float2(0.f, 0.f);
    for (int zk = 0; zk < q.Mz; zk++) {
      const int e = (int)(((long)((iz % q.Mz) + q.Mz) * zk) % q.Mz);
      const float2 w = make_float2(twz[e].x, -twz[e].y);
      const float2 t = cmulf(H[ix * q.Mz + zk], w);
      acc.x += t.x;
      acc.y += t.y;
    }
    T[((size_t)qq * (q.ny + 1) + iy) * q.npair + p] = make_float2(acc.x / q.Mz, acc.y / q.Mz);
  }
}

// F_i = -(d_x tau_xi + d_y tau_yi + d_z tau_zi) on rows 0..ny; ghost rows 0
__global__ __launch_bounds__(256) void k_les_div(const float2 *__restrict__ T, LesGeom q,
                                                 const double *__restrict__ ydc, const int *__restrict__ ydr,
                                                 cx<double> *__restrict__ F) {
  const int p = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;  // device row: iy + 1
  if (p >= q.npair) return;
  const int ix = p / q.Nz, iz = p % q.Nz - q.nz;
  const double kx = q.alfa0 * ix, kz = q.beta0 * iz;
  const size_t cs = (size_t)q.Ny * q.ps;
  if (row == 0 || row == q.ny + 2) {
    for (int i = 0; i < 3; i++) F[i * cs + (size_t)row * q.ps + p] = cx<double>{0.0, 0.0};
    return;
  }
  const int iy = row - 1;
  auto Tq = [&](int qq, int r) {
    const float2 t = T[((size_t)qq * (q.ny + 1) + r) * q.npair + p];
    return cx<double>{(double)t.x, (double)t.y};
  };
  auto dyq = [&](int qq) {
    const int r0 = ydr[iy];
    cx<double> a{0.0, 0.0};
    for (int k = 0; k < 3; k++) {
      const cx<double> t = Tq(qq, r0 + k);
      a.x += ydc[3 * iy + k] * t.x;
      a.y += ydc[3 * iy + k] * t.y;
    }
    return a;
  };
  // (x-stress, y-stress, z-stress) of each momentum component i
  const int qx[3] = {0, 3, 4}, qy[3] = {3, 1, 5}, qz[3] = {4, 5, 2};
  for (int i = 0; i < 3; i++) {
    const cx<double> a = Tq(qx[i], iy), b = dyq(qy[i]), cz = Tq(qz[i], iy);
    // i kx a + b + i kz c
    const double re = -kx * a.y + b.x - kz * cz.y, im = kx * a.x + b.y + kz * cz.x;
    F[i * cs + (size_t)row * q.ps + p] = cx<double>{-re, -im};
  }
}

template <class T>
int lmal(T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(T) + 256);
  if (e != hipSuccess) return set_error((int)e, "hipMalloc of %zu bytes failed", n * sizeof(T));
  return 0;
}

#define LCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return set_error((int)e_, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// the closure buffers of c, made on the first call
int les_state(chan_ctx *c, LesState **out) {
  if (c->les) {
    *out = c->les;
    return 0;
  }
  const LesGeom q = les_geom(c);
  LesState *s = new LesState();
  s->n = (q.ny + 1) * q.Mx * q.Mz;
  int rc;
  auto bail = [&](int code) {
    les_free(s);
    return code;
  };
  const size_t n = (size_t)s->n;
  if ((rc = lmal(&s->Gp, n * 9)) || (rc = lmal(&s->Xf, n * 46)) || (rc = lmal(&s->Xb, n * mlp::KP)) ||
      (rc = lmal(&s->Y, n * 6)) || (rc = lmal(&s->T, (size_t)6 * (q.ny + 1) * q.npair)) ||
      (rc = lmal(&s->twz, (size_t)q.Mz)) || (rc = lmal(&s->twx, (size_t)q.Mx)) ||
      (rc = lmal(&s->ydc, (size_t)3 * (q.ny + 1))) || (rc = lmal(&s->ydr, (size_t)q.ny + 1)) ||
      (rc = lmal(&s->yplus, (size_t)q.ny + 1)) || (rc = lmal(&s->norm, (size_t)104)) || (rc = lmal(&s->wall, (size_t)10)))
    return bail(rc);
  // twiddles
  std::vector<float2> tz(q.Mz), tx(q.Mx);
  const double pi2 = 2.0 * 3.14159265358979323846;
  for (int k = 0; k < q.Mz; k++) tz[k] = make_float2((float)cos(pi2 * k / q.Mz), (float)sin(pi2 * k / q.Mz));
  for (int k = 0; k < q.Mx; k++) tx[k] = make_float2((float)cos(pi2 * k / q.Mx), (float)sin(pi2 * k / q.Mx));
  // second-order differences on the stretched rows 0..ny (numpy.gradient, edge_order 2)
  const double *y = c->hs.y + 1;  // y[iy], iy = 0..ny
  const int ny = q.ny;
  std::vector<double> dc(3 * (ny + 1));
  std::vector<int> dr(ny + 1);
  for (int i = 0; i <= ny; i++) {
    double *w = &dc[3 * i];
    if (i == 0) {
      const double h1 = y[1] - y[0], h2 = y[2] - y[1];
      w[0] = -(2 * h1 + h2) / (h1 * (h1 + h2));
      w[1] = (h1 + h2) / (h1 * h2);
      w[2] = -h1 / (h2 * (h1 + h2));
      dr[i] = 0;
    } else if (i == ny) {
      const double h1 = y[ny - 1] - y[ny - 2], h2 = y[ny] - y[ny - 1];
      w[0] = h2 / (h1 * (h1 + h2));
      w[1] = -(h1 + h2) / (h1 * h2);
      w[2] = (2 * h2 + h1) / (h2 * (h1 + h2));
      dr[i] = ny - 2;
    } else {
      const double h1 = y[i] - y[i - 1], h2 = y[i + 1] - y[i];
      w[0] = -h2 / (h1 * (h1 + h2));
      w[1] = (h2 - h1) / (h1 * h2);
      w[2] = h1 / (h2 * (h1 + h2));
      dr[i] = i - 1;
    }
  }
  double wall[10];
  std::memcpy(wall, c->hs.bnd[B_D140], 5 * sizeof(double));
  std::memcpy(wall + 5, c->hs.bnd[B_D14N], 5 * sizeof(double));
  if (hipMemcpy(s->twz, tz.data(), tz.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->twx, tx.data(), tx.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->ydc, dc.data(), dc.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->ydr, dr.data(), dr.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->wall, wall, sizeof wall, hipMemcpyHostToDevice) != hipSuccess)
    return bail(set_error(CHAN_E_STATE, "upload of the closure set-up failed"));
  c->les = s;
  *out = s;
  return 0;
}

// gradients, y+ and features of the resident field (Xf and/or Xb)
int les_features(chan_ctx *c, LesState *s, bool raw, bool bf16) {
  const LesGeom q = les_geom(c);
  const size_t lds_g = sizeof(float2) * ((size_t)q.npair + (size_t)q.Nx * q.Mz);
  if (lds_g > 64 * 1024) return set_error(CHAN_E_SIZE, "closure grid too large (%zu bytes of LDS per row)", lds_g);
  hipLaunchKernelGGL(k_les_grad, dim3(q.ny + 1, 3), dim3(256), lds_g, c->st, c->V, q, s->ydc, s->ydr, s->twz, s->twx,
                     s->Gp);
  LCHK(hipGetLastError());
  hipLaunchKernelGGL(k_les_yplus, dim3(1), dim3(256), 0, c->st, c->V, q, s->wall, c->y, c->prm.ymin, c->prm.ymax,
                     c->prm.ni, s->yplus);
  LCHK(hipGetLastError());
  hipLaunchKernelGGL(k_les_features, dim3((s->n + 255) / 256), dim3(256), 0, c->st, s->Gp, q, s->yplus, s->norm,
                     raw ? s->Xf : nullptr, bf16 ? s->Xb : nullptr);
  LCHK(hipGetLastError());
  return 0;
}

}  // namespace
}  // namespace chan

using namespace chan;

extern "C" {

int chan_les_features(chan_ctx *c, float *X) {
  if (!c || !X) return set_error(CHAN_E_ARG, "null argument");
  if (!c->field_set) return set_error(CHAN_E_STATE, "closure features before set_field");
  LesState *s;
  int rc;
  if ((rc = les_state(c, &s))) return rc;
  if ((rc = les_features(c, s, true, false))) return rc;
  LCHK(hipMemcpyAsync(X, s->Xf, sizeof(float) * 46 * (size_t)s->n, hipMemcpyDeviceToHost, c->st));
  LCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_les_samples(const chan_ctx *c, int64_t *n) {
  if (!c || !n) return set_error(CHAN_E_ARG, "null argument");
  *n = (int64_t)(c->g.ny + 1) * 2 * c->g.nxd * c->g.nzd;
  return 0;
}

int chan_les_closure_force(chan_ctx *c, chan_mlp *m, const float *norm) {
  if (!c || !m || !norm) return set_error(CHAN_E_ARG, "null argument");
  if (!c->field_set) return set_error(CHAN_E_STATE, "closure force before set_field");
  LesState *s;
  int rc;
  if ((rc = les_state(c, &s))) return rc;
  for (int j = 0; j < 46; j++)
    if (!(norm[46 + j] > 0.0f)) return set_error(CHAN_E_ARG, "feature scale %d is not positive", j);
  LCHK(hipMemcpyAsync(s->norm, norm, sizeof(float) * 104, hipMemcpyHostToDevice, c->st));
  if ((rc = les_features(c, s, false, true))) return rc;
  LCHK(mlp::forward_device(m, s->Xb, s->n, s->Y, c->st));
  const LesGeom q = les_geom(c);
  const size_t lds_t = sizeof(float) * ((size_t)q.Mx * q.Mz + 1) + sizeof(float2) * (size_t)q.Nx * q.Mz + 16;
  if (lds_t > 64 * 1024) return set_error(CHAN_E_SIZE, "closure grid too large (%zu bytes of LDS per row)", lds_t);
  hipLaunchKernelGGL(k_les_tau, dim3(q.ny + 1, 6), dim3(256), lds_t, c->st, s->Y, q, s->norm, s->twz, s->twx, s->T);
  LCHK(hipGetLastError());
  if (!c->F0) {
    if ((rc = lmal(&c->F0, 3 * (size_t)q.Ny * q.ps))) return rc;
    LCHK(hipMemsetAsync(c->F0, 0, sizeof(cx<double>) * 3 * (size_t)q.Ny * q.ps, c->st));
  }
  hipLaunchKernelGGL(k_les_div, dim3((q.npair + 255) / 256, q.Ny), dim3(256), 0, c->st, s->T, q, s->ydc, s->ydr, c->F0);
  LCHK(hipGetLastError());
  LCHK(hipStreamSynchronize(c->st));
  return 0;
}

}  // extern "C"
