// api.hip -- the C-ABI of include/channel_hip.h: device-resident state of
// one channel DNS, set-up, field exchange in the reference layout and the
// launch sequences of buildrhs / linsolve / the RK3 time loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/channel_hip.h"
#include "chan_internal.hpp"
#include "ctx_internal.hpp"

using namespace chan;

namespace {

thread_local char g_err[1024] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace

// error reporting shared with the closure library part (k_mlp.hip)
int chan::set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

namespace {

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail((int)e_, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

// RK3 coefficients of Rai & Moin (dnsdata.f90:70-72)
const double RK1[3] = {120.0 / 32.0, 2.0, 0.0};
const double RK2[3] = {120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0};
const double RK3[3] = {120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0};

// ---------------------------------------------------- layout transposes --
// host layout: [c][pair][row] (the reference's V(iy,iz,ix,c), iy fastest);
// device layout: [c][row][pair] with plane stride ps.
constexpr int TT = 32;
__global__ __launch_bounds__(TT * 8) void k_to_dev(const cx<double> *__restrict__ in, cx<double> *__restrict__ out, Geom g) {
  __shared__ cx<double> tile[TT][TT + 1];
  const int c = blockIdx.z;
  const int p0 = blockIdx.x * TT, r0 = blockIdx.y * TT;
  const cx<double> *src = in + (size_t)c * g.npair * g.Ny;
  cx<double> *dst = out + (size_t)c * g.Ny * g.ps;
  for (int k = threadIdx.y; k < TT; k += 8) {
    const int p = p0 + k, r = r0 + threadIdx.x;
    if (p < g.npair && r < g.Ny) tile[k][threadIdx.x] = src[(size_t)p * g.Ny + r];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < TT; k += 8) {
    const int r = r0 + k, p = p0 + threadIdx.x;
    if (p < g.npair && r < g.Ny) dst[(size_t)r * g.ps + p] = tile[threadIdx.x][k];
  }
}
__global__ __launch_bounds__(TT * 8) void k_to_host(const cx<double> *__restrict__ in, cx<double> *__restrict__ out, Geom g) {
  __shared__ cx<double> tile[TT][TT + 1];
  const int c = blockIdx.z;
  const int p0 = blockIdx.x * TT, r0 = blockIdx.y * TT;
  const cx<double> *src = in + (size_t)c * g.Ny * g.ps;
  cx<double> *dst = out + (size_t)c * g.npair * g.Ny;
  for (int k = threadIdx.y; k < TT; k += 8) {
    const int r = r0 + k, p = p0 + threadIdx.x;
    if (p < g.npair && r < g.Ny) tile[k][threadIdx.x] = src[(size_t)r * g.ps + p];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < TT; k += 8) {
    const int p = p0 + k, r = r0 + threadIdx.x;
    if (p < g.npair && r < g.Ny) dst[(size_t)p * g.Ny + r] = tile[threadIdx.x][k];
  }
}

// outstats staging: the (0,0) columns of all three components as out[c][row],
// then a copy of the device scalars; the running CFL maximum restarts at 0.
__global__ void k_gather00(const cx<double> *__restrict__ V, cx<double> *__restrict__ out, DevScalars *sc, Geom g) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) {
    DevScalars *stash = reinterpret_cast<DevScalars *>(out + 3 * g.Ny);
    *stash = *sc;
    sc->cfl_bits = 0ull;
  }
  if (r >= g.Ny) return;
  for (int c = 0; c < 3; c++) out[c * g.Ny + r] = V[((size_t)c * g.Ny + r) * g.ps + g.nz];
}
constexpr size_t STASH = (sizeof(DevScalars) + sizeof(cx<double>) - 1) / sizeof(cx<double>);

// grid-stride streaming copy of n 32-byte elements
__global__ __launch_bounds__(256) void k_copy(const double4 *__restrict__ a, double4 *__restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

template <typename T>
void twiddles(std::vector<cx<T>> &w, int n) {
  w.resize((size_t)fft::tw_table_size(n));
  fft::build_twiddles(w.data(), n);
}
// the x pass's table: 768-point rows use the four-step transform, 1536-point
// rows its two-wave form
template <typename T>
void twiddles_x(std::vector<cx<T>> &w, int n) {
  if (n == 1536) {
    w.resize((size_t)fft::FS2_TW);
    fft::build_fs1536_twiddles(w.data());
    return;
  }
  if (n != 768) return twiddles(w, n);
  w.resize((size_t)fft::FS_TW);
  fft::build_fs768_twiddles(w.data());
}

}  // namespace

namespace chan {
hipError_t launch_to_device_layout(hipStream_t st, const cx<double> *in, cx<double> *out, const Geom &g) {
  dim3 grid((unsigned)((g.npair + TT - 1) / TT), (unsigned)((g.Ny + TT - 1) / TT), 3);
  hipLaunchKernelGGL(k_to_dev, grid, dim3(TT, 8), 0, st, in, out, g);
  return hipGetLastError();
}
hipError_t launch_to_host_layout(hipStream_t st, const cx<double> *in, cx<double> *out, const Geom &g) {
  dim3 grid((unsigned)((g.npair + TT - 1) / TT), (unsigned)((g.Ny + TT - 1) / TT), 3);
  hipLaunchKernelGGL(k_to_host, grid, dim3(TT, 8), 0, st, in, out, g);
  return hipGetLastError();
}
}  // namespace chan

namespace {

int kid(chan_ctx *c, const char *name) {
  for (size_t i = 0; i < c->names.size(); i++)
    if (c->names[i] == name) return (int)i;
  c->names.push_back(name);
  c->counts.push_back(0);
  c->ms.push_back(0.0);
  return (int)c->names.size() - 1;
}

hipEvent_t get_event(chan_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int harvest(chan_ctx *c) {
  for (auto &r : c->pend) {
    HIPCHK(hipEventSynchronize(r.b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
    c->counts[r.id] += 1;
    c->ms[r.id] += t;
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->pend.clear();
  return 0;
}

// runs `launch` on the context's stream, bracketed by events when timing
template <class F>
int timed(chan_ctx *c, const char *name, F launch) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->timing) {
    if (c->pend.size() > 8192) {
      int rc = harvest(c);
      if (rc) return rc;
    }
    a = get_event(c);
    b = get_event(c);
    if (!a || !b) return fail(CHAN_E_STATE, "hipEventCreate failed");
    HIPCHK(hipEventRecord(a, c->st));
  }
  hipError_t e = launch();
  if (e != hipSuccess) return fail((int)e, "launch of %s failed: %s", name, hipGetErrorString(e));
  if (c->timing) {
    HIPCHK(hipEventRecord(b, c->st));
    c->pend.push_back({kid(c, name), a, b});
  }
  return 0;
}

int dalloc_bytes(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes + 256);
  if (e != hipSuccess) return fail((int)e, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  return 0;
}

template <class T>
int dalloc(T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(T) + 256);
  if (e != hipSuccess) return fail((int)e, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  return 0;
}

ForceArgs force_args(const chan_ctx *c) {
  ForceArgs a;
  a.V = c->V;
  a.F0 = c->F0;
  a.gain = c->gain;
  a.der = c->der;
  for (int k = 0; k < 3; k++) a.src[k] = c->gain ? c->src[k] : -1;
  a.out = nullptr;
  a.g = c->g;
  a.chunk = 48;
  return a;
}

double cfl_value(const DevScalars &s) {
  double v;
  std::memcpy(&v, &s.cfl_bits, sizeof v);
  return v;
}

}  // namespace

extern "C" {

int chan_abi_version(void) { return CHAN_ABI_VERSION; }
const char *chan_last_error(void) { return g_err; }

int chan_create(const chan_params *p, chan_ctx **out) {
  if (!p || !out) return fail(CHAN_E_ARG, "null argument");
  *out = nullptr;
  if (p->nx < 1 || p->nz < 1 || p->ny < 8) return fail(CHAN_E_ARG, "need nx >= 1, nz >= 1, ny >= 8 (got %d %d %d)", p->nx, p->ny, p->nz);
  if (p->precision != CHAN_FP64 && p->precision != CHAN_SPLIT)
    return fail(CHAN_E_ARG, "precision mode %d is not available", p->precision);
  if (p->cpi && p->cpi_type != 0 && p->cpi_type != 1) return fail(CHAN_E_CPITYPE, "Wrong selection of CPI_Type (%d)", p->cpi_type);
  if (!(p->alfa0 > 0) || !(p->beta0 > 0) || !(p->ni > 0) || !(p->ymax > p->ymin) || !(p->a > 0))
    return fail(CHAN_E_ARG, "alfa0, beta0, ni, a must be positive and ymax > ymin");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CHAN_E_NODEV, "no HIP device");
  chan_ctx *c = new chan_ctx();
  c->prm = *p;
  c->split = p->precision == CHAN_SPLIT;
  setup_build(c->hs, p->nx, p->ny, p->nz, p->alfa0, p->beta0, p->a, p->ymin, p->ymax);
  const HostSetup &h = c->hs;
  // the x pass holds at most (E+2)/3 x modes per thread, i.e. Nx <= 2nxd/3,
  // which fftFIT's nxd >= 3(nx+1)/2 guarantees
  if (!fft_size_supported(h.nzd) || !fft_size_supported(2 * h.nxd) || h.nzd % ZB != 0 || 3 * (p->nx + 1) > 2 * h.nxd) {
    int rc;
    if (3 * (p->nx + 1) > 2 * h.nxd)
      rc = fail(CHAN_E_SIZE,
                "nx=%d: the x pass needs 3(nx+1) <= 2nxd, but fftFIT keeps nxd=%d (3(nx+1)/2 rounded down is "
                "already 2^k or 3*2^k), so 3(nx+1)=%d > 2nxd=%d; choose another nx (nx=%d or %d)",
                p->nx, h.nxd, 3 * (p->nx + 1), 2 * h.nxd, p->nx - 1, p->nx + 1);
    else
      rc = fail(CHAN_E_SIZE, "transform sizes nzd=%d, 2nxd=%d are not supported", h.nzd, 2 * h.nxd);
    setup_free(c->hs);
    delete c;
    return rc;
  }
  Geom &g = c->g;
  g.nx = p->nx;
  g.ny = p->ny;
  g.nz = p->nz;
  g.Nx = p->nx + 1;
  g.Nz = 2 * p->nz + 1;
  g.Ny = p->ny + 3;
  g.nxd = h.nxd;
  g.nzd = h.nzd;
  g.npair = g.Nx * g.Nz;
  g.ps = ((size_t)g.npair + 1 + 63) / 64 * 64;  // one spare slot: the ucor solve
  g.alfa0 = p->alfa0;
  g.beta0 = p->beta0;
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
      ncu = 256;
    g.ncu = ncu;
  }
  c->time = p->time;
  c->deltat = p->deltat;

  int rc = 0;
  auto bail = [&](int code) {
    chan_destroy(c);
    return code;
  };
  if (hipStreamCreate(&c->st) != hipSuccess) return bail(fail(CHAN_E_STATE, "hipStreamCreate failed"));
  const size_t plane = g.ps, NyP = (size_t)g.Ny * plane;
  const size_t slab = (size_t)g.nzd * g.Nx;  // one component of one plane in physical z
  if ((rc = dalloc(&c->V, 3 * NyP))) return bail(rc);
  if ((rc = dalloc(&c->rhs, 2 * (size_t)(g.ny - 1) * plane))) return bail(rc);
  if ((rc = dalloc(&c->L, 2 * (size_t)(g.ny - 1) * plane))) return bail(rc);
  if ((rc = dalloc(&c->col00, 3 * (size_t)g.Ny + STASH))) return bail(rc);
  const size_t esz = c->split ? sizeof(cx<float>) : sizeof(cx<double>);
  if ((rc = dalloc_bytes(&c->old, 2 * (size_t)(g.ny - 1) * plane * esz))) return bail(rc);
  // B doubles as the staging area of the field and rhs transposes, which
  // needs 3*Ny*npair complex FP64 <= 6*Ny*slab*esz
  if ((rc = dalloc_bytes(&c->A, 3 * (size_t)g.Ny * slab * esz))) return bail(rc);
  if ((rc = dalloc_bytes(&c->B, 6 * (size_t)g.Ny * slab * esz))) return bail(rc);
  if ((rc = dalloc_bytes(&c->P, 6 * NyP * esz))) return bail(rc);
  if ((rc = dalloc_bytes(&c->twz, (size_t)fft::tw_table_size(g.nzd) * esz))) return bail(rc);
  if ((rc = dalloc_bytes(&c->twx, (size_t)fft::tw_table_size(2 * g.nxd) * esz))) return bail(rc);
  if ((rc = dalloc(&c->der, (size_t)(g.ny - 1) * 20))) return bail(rc);
  if ((rc = dalloc(&c->d0fac, (size_t)(g.ny + 1) * 5))) return bail(rc);
  if ((rc = dalloc(&c->y, (size_t)g.Ny))) return bail(rc);
  if ((rc = dalloc(&c->dy, (size_t)g.ny + 1))) return bail(rc);
  if ((rc = dalloc(&c->sc, 1))) return bail(rc);
  if ((rc = dalloc(&c->stat, (size_t)10 * (g.ny + 1)))) return bail(rc);

  auto upload_tw = [&](auto proto) {
    std::vector<decltype(proto)> tw;
    twiddles(tw, g.nzd);
    if (hipMemcpy(c->twz, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice) != hipSuccess) return 1;
    twiddles_x(tw, 2 * g.nxd);
    if (hipMemcpy(c->twx, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice) != hipSuccess) return 1;
    return 0;
  };
  if (c->split ? upload_tw(cx<float>{}) : upload_tw(cx<double>{})) return bail(fail(CHAN_E_STATE, "upload failed"));
  if (hipMemcpy(c->der, h.der, sizeof(double) * 20 * (g.ny - 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d0fac, h.d0fac, sizeof(double) * 5 * (g.ny + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->y, h.y, sizeof(double) * g.Ny, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->dy, h.dy, sizeof(double) * (g.ny + 1), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(CHAN_E_STATE, "upload of the set-up failed"));
  DevScalars s0;
  std::memset(&s0, 0, sizeof s0);
  s0.meanpx = p->meanpx;
  s0.meanpz = p->meanpz;
  if (hipMemcpy(c->sc, &s0, sizeof s0, hipMemcpyHostToDevice) != hipSuccess) return bail(fail(CHAN_E_STATE, "upload failed"));
  // stored explicit terms start at zero (ODE3 = 0 in the first substep)
  if (hipMemset(c->old, 0, esz * 2 * (g.ny - 1) * plane) != hipSuccess ||
      hipMemset(c->V, 0, sizeof(cx<double>) * 3 * NyP) != hipSuccess)
    return bail(fail(CHAN_E_STATE, "hipMemset failed"));
  *out = c;
  return 0;
}

int chan_destroy(chan_ctx *c) {
  if (!c) return 0;
  if (c->st) hipStreamSynchronize(c->st);
  if (c->les) les_free(c->les);
  for (auto &r : c->pend) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : c->pool) hipEventDestroy(e);
  void *bufs[] = {c->V, c->old, c->rhs, c->L, c->col00, c->A, c->B, c->P, c->twz, c->twx,
                  c->der, c->d0fac, c->y, c->dy, c->sc, c->stat, c->F0, c->gain, c->Fkeep};
  for (void *b : bufs)
    if (b) hipFree(b);
  if (c->st) hipStreamDestroy(c->st);
  setup_free(c->hs);
  delete c;
  return 0;
}

int chan_sizes(const chan_ctx *c, int32_t s[5]) {
  if (!c || !s) return fail(CHAN_E_ARG, "null argument");
  s[0] = c->g.nx;
  s[1] = c->g.ny;
  s[2] = c->g.nz;
  s[3] = c->g.nxd;
  s[4] = c->g.nzd;
  return 0;
}

int chan_get_setup(const chan_ctx *c, double *y, double *der, double *bnd) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  if (y) std::memcpy(y, c->hs.y, sizeof(double) * c->g.Ny);
  if (der) std::memcpy(der, c->hs.der, sizeof(double) * 20 * (c->g.ny - 1));
  if (bnd) std::memcpy(bnd, c->hs.bnd, sizeof(double) * 50);
  return 0;
}

int chan_setup_host(const chan_params *p, int32_t sizes[5], double *y, double *der, double *bnd, double *bc,
                    double *d0fac) {
  if (!p) return fail(CHAN_E_ARG, "null argument");
  if (p->nx < 1 || p->nz < 1 || p->ny < 8) return fail(CHAN_E_ARG, "need nx >= 1, nz >= 1, ny >= 8");
  HostSetup h;
  setup_build(h, p->nx, p->ny, p->nz, p->alfa0, p->beta0, p->a, p->ymin, p->ymax);
  if (sizes) {
    sizes[0] = p->nx;
    sizes[1] = p->ny;
    sizes[2] = p->nz;
    sizes[3] = h.nxd;
    sizes[4] = h.nzd;
  }
  if (y) std::memcpy(y, h.y, sizeof(double) * (p->ny + 3));
  if (der) std::memcpy(der, h.der, sizeof(double) * 20 * (p->ny - 1));
  if (bnd) std::memcpy(bnd, h.bnd, sizeof(double) * 50);
  if (bc) std::memcpy(bc, h.bc, sizeof(double) * 40);
  if (d0fac) std::memcpy(d0fac, h.d0fac, sizeof(double) * 5 * (p->ny + 1));
  setup_free(h);
  return 0;
}

int chan_set_field(chan_ctx *c, const double *V) {
  if (!c || !V) return fail(CHAN_E_ARG, "null argument");
  const size_t n = (size_t)3 * c->g.npair * c->g.Ny;
  cx<double> *stage = static_cast<cx<double> *>(c->B);
  HIPCHK(hipMemcpyAsync(stage, V, n * sizeof(cx<double>), hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_to_device_layout(c->st, stage, c->V, c->g));
  HIPCHK(hipStreamSynchronize(c->st));
  c->field_set = true;
  return 0;
}

int chan_get_field(chan_ctx *c, double *V) {
  if (!c || !V) return fail(CHAN_E_ARG, "null argument");
  const size_t n = (size_t)3 * c->g.npair * c->g.Ny;
  cx<double> *stage = static_cast<cx<double> *>(c->B);
  HIPCHK(launch_to_host_layout(c->st, c->V, stage, c->g));
  HIPCHK(hipMemcpyAsync(V, stage, n * sizeof(cx<double>), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_set_body_force(chan_ctx *c, const double *F) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  if (!F) {
    if (c->F0) HIPCHK(hipFree(c->F0));
    c->F0 = nullptr;
    return 0;
  }
  int rc;
  if (!c->F0 && (rc = dalloc(&c->F0, 3 * (size_t)c->g.Ny * c->g.ps))) return rc;
  const size_t n = (size_t)3 * c->g.npair * c->g.Ny;
  cx<double> *stage = static_cast<cx<double> *>(c->B);
  HIPCHK(hipMemcpyAsync(stage, F, n * sizeof(cx<double>), hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_to_device_layout(c->st, stage, c->F0, c->g));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_set_body_force_feedback(chan_ctx *c, const int32_t src[3], const double *gain) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  bool on = src && gain;
  if (on) {
    on = false;
    for (int k = 0; k < 3; k++) {
      if (src[k] < -1 || src[k] > 2) return fail(CHAN_E_ARG, "feedback source component %d of force component %d is not -1, 0, 1 or 2", src[k], k);
      on = on || src[k] >= 0;
    }
  }
  HIPCHK(hipStreamSynchronize(c->st));
  if (!on) {
    if (c->gain) HIPCHK(hipFree(c->gain));
    c->gain = nullptr;
    return 0;
  }
  const size_t n = (size_t)3 * c->g.Ny * (c->g.Nz + 1);
  int rc;
  if (!c->gain && (rc = dalloc(&c->gain, n))) return rc;
  HIPCHK(hipMemcpy(c->gain, gain, n * sizeof(double), hipMemcpyHostToDevice));
  for (int k = 0; k < 3; k++) c->src[k] = src[k];
  return 0;
}

int chan_get_body_force(chan_ctx *c, double *F) {
  if (!c || !F) return fail(CHAN_E_ARG, "null argument");
  const size_t n = (size_t)3 * c->g.npair * c->g.Ny;
  cx<double> *stage = static_cast<cx<double> *>(c->B);
  if (!c->F0 && !c->gain) {
    std::memset(F, 0, n * sizeof(cx<double>));
    return 0;
  }
  // the products buffer (dead between calls; 6*Ny*ps elements of at least
  // 8 bytes) holds the force in the device layout, 3*Ny*ps complex FP64
  cx<double> *dev = static_cast<cx<double> *>(c->P);
  ForceArgs a = force_args(c);
  a.out = dev;
  HIPCHK(launch_force(c->st, a));
  HIPCHK(launch_to_host_layout(c->st, dev, stage, c->g));
  HIPCHK(hipMemcpyAsync(F, stage, n * sizeof(cx<double>), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_keep_body_force(chan_ctx *c, int32_t on) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  c->kept_state = 0;
  if (!on) {
    if (c->Fkeep) HIPCHK(hipFree(c->Fkeep));
    c->Fkeep = nullptr;
    return 0;
  }
  int rc;
  if (!c->Fkeep && (rc = dalloc(&c->Fkeep, 3 * (size_t)c->g.Ny * c->g.ps))) return rc;
  return 0;
}

int chan_get_applied_body_force(chan_ctx *c, double *F) {
  if (!c || !F) return fail(CHAN_E_ARG, "null argument");
  if (!c->Fkeep) return fail(CHAN_E_STATE, "get_applied_body_force without keep_body_force");
  if (!c->kept_state) return fail(CHAN_E_STATE, "get_applied_body_force before a buildrhs");
  const size_t n = (size_t)3 * c->g.npair * c->g.Ny;
  if (c->kept_state == 2) {
    std::memset(F, 0, n * sizeof(cx<double>));
    return 0;
  }
  cx<double> *stage = static_cast<cx<double> *>(c->B);
  HIPCHK(launch_to_host_layout(c->st, c->Fkeep, stage, c->g));
  HIPCHK(hipMemcpyAsync(F, stage, n * sizeof(cx<double>), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_get_rhs(chan_ctx *c, double *R) {
  if (!c || !R) return fail(CHAN_E_ARG, "null argument");
  // rhs is [2][ny-1][ps]: the same transpose as the field with ny-1 rows
  Geom g = c->g;
  g.Ny = g.ny - 1;
  cx<double> *stage = static_cast<cx<double> *>(c->B);
  for (int k = 0; k < 2; k++) {
    Geom g1 = g;
    // one component at a time: the transpose kernel runs over 3 components,
    // so reuse it with a single-component view
    dim3 grid((unsigned)((g1.npair + 31) / 32), (unsigned)((g1.Ny + 31) / 32), 1);
    hipLaunchKernelGGL(k_to_host, grid, dim3(32, 8), 0, c->st, c->rhs + (size_t)k * g1.Ny * g1.ps,
                       stage + (size_t)k * g1.npair * g1.Ny, g1);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(R, stage, sizeof(cx<double>) * 2 * (size_t)g.npair * g.Ny, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_get_scalars(chan_ctx *c, chan_scalars *s) {
  if (!c || !s) return fail(CHAN_E_ARG, "null argument");
  DevScalars d;
  HIPCHK(hipMemcpyAsync(&d, c->sc, sizeof d, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  s->time = c->time;
  s->deltat = c->deltat;
  s->meanpx = d.meanpx;
  s->meanpz = d.meanpz;
  s->corrpx = d.corrpx;
  s->corrpz = d.corrpz;
  s->fr[0] = d.fr[0];
  s->fr[1] = d.fr[1];
  s->fr[2] = d.fr[2];
  s->cfl = cfl_value(d);
  return 0;
}

int chan_set_scalars(chan_ctx *c, const chan_scalars *s) {
  if (!c || !s) return fail(CHAN_E_ARG, "null argument");
  DevScalars d;
  std::memset(&d, 0, sizeof d);
  d.meanpx = s->meanpx;
  d.meanpz = s->meanpz;
  d.corrpx = s->corrpx;
  d.corrpz = s->corrpz;
  d.fr[0] = s->fr[0];
  d.fr[1] = s->fr[1];
  d.fr[2] = s->fr[2];
  const double cf = s->cfl < 0 ? 0.0 : s->cfl;
  std::memcpy(&d.cfl_bits, &cf, sizeof cf);
  c->time = s->time;
  c->deltat = s->deltat;
  HIPCHK(hipMemcpyAsync(c->sc, &d, sizeof d, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C++" {
// K1-K3 with transform arithmetic and intermediates in R (double, or float
// in the split mode)
template <typename R>
static int products_t(chan_ctx *c, int plane0, int nplanes, bool prods, int cfl) {
  const Geom &g = c->g;
  const HostSetup &h = c->hs;
  // the transform kernels index A and B from the launch's first plane: the
  // base pointers are offset so that every plane keeps its own slices
  const size_t slab = (size_t)g.nzd * g.Nx;
  cx<R> *A = static_cast<cx<R> *>(c->A) + (size_t)plane0 * 3 * slab;
  cx<R> *B = static_cast<cx<R> *>(c->B) + (size_t)plane0 * 6 * slab;
  cx<R> *P = static_cast<cx<R> *>(c->P);
  const cx<R> *twz = static_cast<const cx<R> *>(c->twz), *twx = static_cast<const cx<R> *>(c->twx);
  int rc;
  if ((rc = timed(c, "zinv", [&] { return launch_zinv<R>(g.nzd, c->st, c->V, A, twz, g, plane0, nplanes); }))) return rc;
  if ((rc = timed(c, prods ? "xpass" : "xpass_cfl", [&] {
         return launch_xpass<R>(2 * g.nxd, c->st, A, B, twx, g, plane0, nplanes, prods, cfl, c->dy, h.dx, h.dz,
                                (R)h.factor, &c->sc->cfl_bits);
       })))
    return rc;
  if (prods)
    if ((rc = timed(c, "zfwd", [&] { return launch_zfwd<R>(g.nzd, c->st, B, P, twz, g, plane0, nplanes); }))) return rc;
  return 0;
}

static int products(chan_ctx *c, int plane0, int nplanes, bool prods, int cfl) {
  return c->split ? products_t<float>(c, plane0, nplanes, prods, cfl) : products_t<double>(c, plane0, nplanes, prods, cfl);
}

template <typename R>
static int rhs_t(chan_ctx *c, const double ode[3], bool keep) {
  RhsArgs<R> a;
  a.P = static_cast<const cx<R> *>(c->P);
  a.V = c->V;
  a.old = static_cast<cx<R> *>(c->old);
  a.rhs = c->rhs;
  a.der = c->der;
  a.sc = c->sc;
  a.g = c->g;
  a.ode0 = ode[0];
  a.ode1 = ode[1];
  a.ode2 = ode[2];
  a.dt = c->deltat;
  a.ni = c->prm.ni;
  a.chunk = 48;
  a.force = c->F0 || c->gain;
  if (keep) c->kept_state = a.force ? 1 : 2;
  if (a.force) {
    ForceArgs f = force_args(c);
    int rc;
    if (keep) {
      f.out = c->Fkeep;
      if ((rc = timed(c, "force_keep", [&] { return launch_force(c->st, f); }))) return rc;
    }
    f.out = c->rhs;
    if ((rc = timed(c, "force", [&] { return launch_force_terms(c->st, f); }))) return rc;
  }
  return timed(c, "rhs", [&] { return launch_rhs<R>(c->st, a); });
}

// keep: store the applied force when keeping is on (chan_step asks only in
// its last substep, whose force is the one that stays)
static int buildrhs(chan_ctx *c, const double ode[3], bool compute_cfl, bool keep) {
  const Geom &g = c->g;
  int rc;
  if ((rc = products(c, 0, g.Ny, true, compute_cfl ? 1 : 0))) return rc;
  keep = keep && c->Fkeep;
  return c->split ? rhs_t<float>(c, ode, keep) : rhs_t<double>(c, ode, keep);
}
}  // extern "C++"

int chan_buildrhs(chan_ctx *c, const double ode[3], int32_t compute_cfl) {
  if (!c || !ode) return fail(CHAN_E_ARG, "null argument");
  if (!c->field_set) return fail(CHAN_E_STATE, "buildrhs before set_field");
  return buildrhs(c, ode, compute_cfl != 0, true);
}

int chan_linsolve(chan_ctx *c, double lambda) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  if (!c->field_set) return fail(CHAN_E_STATE, "linsolve before set_field");
  const Geom &g = c->g;
  int rc;
  SolveArgs s;
  s.V = c->V;
  s.rhs = c->rhs;
  s.L = c->L;
  s.der = c->der;
  std::memcpy(s.bc, c->hs.bc, sizeof s.bc);
  s.g = g;
  s.lambda = lambda;
  s.ni = c->prm.ni;
  s.u0 = c->prm.u0;
  s.uN = c->prm.uN;
  if ((rc = timed(c, "solve_ve", [&] { return launch_solve(c->st, s); }))) return rc;
  MeanArgs m;
  m.V = c->V;
  m.y = c->y;
  m.sc = c->sc;
  std::memcpy(m.bc, c->hs.bc, sizeof m.bc);
  m.g = g;
  m.meanflowx = c->prm.meanflowx;
  m.meanflowz = c->prm.meanflowz;
  m.gamma = c->prm.gamma;
  m.ni = c->prm.ni;
  m.cpi = c->prm.cpi;
  m.cpi_type = c->prm.cpi_type;
  if ((rc = timed(c, "mean", [&] { return launch_mean(c->st, m); }))) return rc;
  UwArgs u;
  u.V = c->V;
  u.der = c->der;
  u.d0fac = c->d0fac;
  std::memcpy(u.bnd, c->hs.bnd, sizeof u.bnd);
  u.g = g;
  return timed(c, "uw", [&] { return launch_uw(c->st, u); });
}

int chan_initial(chan_ctx *c) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  if (!c->field_set) return fail(CHAN_E_STATE, "initial before set_field");
  if (c->deltat == 0) c->deltat = 1.0;
  const Geom &g = c->g;
  int rc;
  // interior planes iy = 1..ny-1 are plane indices 2..ny
  if ((rc = products(c, 2, g.ny - 1, false, 1))) return rc;
  MeanArgs m;
  std::memset(&m, 0, sizeof m);
  m.V = c->V;
  m.y = c->y;
  m.sc = c->sc;
  m.g = g;
  m.gamma = c->prm.gamma;
  m.ni = c->prm.ni;
  m.cpi = c->prm.cpi;
  m.cpi_type = c->prm.cpi_type;
  return timed(c, "flowrate", [&] { return launch_flowrate(c->st, m); });
}

int chan_stats(chan_ctx *c, const double *avg, double *out) {
  if (!c || !out) return fail(CHAN_E_ARG, "null argument");
  if (!c->field_set) return fail(CHAN_E_STATE, "stats before set_field");
  const Geom &g = c->g;
  const size_t n1 = (size_t)g.ny + 1;
  double *dav = nullptr;
  if (avg) {
    dav = c->stat;
    HIPCHK(hipMemcpyAsync(dav, avg, sizeof(double) * 3 * n1, hipMemcpyHostToDevice, c->st));
  }
  int rc;
  if ((rc = timed(c, "stats", [&] { return launch_stats(c->st, c->V, dav, c->stat + 3 * n1, g); }))) return rc;
  HIPCHK(hipMemcpyAsync(out, c->stat + 3 * n1, sizeof(double) * 7 * n1, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_outstats(chan_ctx *c, double row[11]) {
  if (!c || !row) return fail(CHAN_E_ARG, "null argument");
  const Geom &g = c->g;
  std::vector<cx<double>> col(3 * (size_t)g.Ny + STASH);
  DevScalars d;
  // one kernel stages the columns and scalars and resets the running CFL
  // maximum (cfl=0 in outstats); one copy brings them back
  hipLaunchKernelGGL(k_gather00, dim3((g.Ny + 127) / 128), dim3(128), 0, c->st, c->V, c->col00, c->sc, g);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(col.data(), c->col00, col.size() * sizeof(col[0]), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  std::memcpy(&d, col.data() + 3 * (size_t)g.Ny, sizeof d);
  const double rg = cfl_value(d);
  if (!std::isfinite(rg)) return fail(CHAN_E_NOTFINITE, "CFL maximum is not finite (%g): the field holds NaN or Inf", rg);
  if (c->prm.cflmax > 0) c->deltat = c->prm.cflmax / rg;
  const double *b140 = c->hs.bnd[B_D140], *b14n = c->hs.bnd[B_D14N];
  const int ny = g.ny;
  auto re = [&](int comp, int iy) { return col[(size_t)comp * g.Ny + iy + 1].x; };
  double d11 = 0, d21 = 0, d12 = 0, d22 = 0;
  for (int j = 0; j < 5; j++) {
    d12 += b14n[j] * re(0, ny - 3 + j);
    d22 += b14n[j] * re(2, ny - 3 + j);
  }
  d12 = -d12;
  d22 = -d22;
  for (int j = 0; j < 5; j++) {
    d11 += b140[j] * re(0, -1 + j);
    d21 += b140[j] * re(2, -1 + j);
  }
  row[0] = c->time;
  row[1] = d11;
  row[2] = d12;
  row[3] = d21;
  row[4] = d22;
  row[5] = d.fr[0] + d.corrpx * d.fr[2];
  row[6] = d.meanpx + d.corrpx;
  row[7] = d.fr[1] + d.corrpz * d.fr[2];
  row[8] = d.meanpz + d.corrpz;
  row[9] = rg * c->deltat;
  row[10] = c->deltat;
  return 0;
}

int chan_step(chan_ctx *c, double row[11]) {
  if (!c || !row) return fail(CHAN_E_ARG, "null argument");
  int rc;
  const double *ode[3] = {RK1, RK2, RK3};
  for (int k = 0; k < 3; k++) {
    c->time = c->time + 2.0 / ode[k][0] * c->deltat;
    if (!c->field_set) return fail(CHAN_E_STATE, "step before set_field");
    if ((rc = buildrhs(c, ode[k], k == 2, k == 2))) return rc;
    if ((rc = chan_linsolve(c, ode[k][0] / c->deltat))) return rc;
  }
  return chan_outstats(c, row);
}

int chan_steps(chan_ctx *c, int32_t n, double *rows) {
  if (!c || (n > 0 && !rows)) return fail(CHAN_E_ARG, "null argument");
  for (int32_t i = 0; i < n; i++) {
    int rc = chan_step(c, rows + 11 * (size_t)i);
    if (rc) return rc;
  }
  return 0;
}

int chan_timing(chan_ctx *c, int32_t enable) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  if (!enable) {
    int rc = harvest(c);
    if (rc) return rc;
  }
  c->timing = enable != 0;
  if (enable) {
    int rc = harvest(c);
    if (rc) return rc;
    for (size_t i = 0; i < c->names.size(); i++) {
      c->counts[i] = 0;
      c->ms[i] = 0;
    }
  }
  return 0;
}

int chan_kernel_times(chan_ctx *c, int32_t cap, char *names, int64_t *counts, double *total_ms, int32_t *n) {
  if (!c || !n) return fail(CHAN_E_ARG, "null argument");
  int rc = harvest(c);
  if (rc) return rc;
  *n = (int32_t)c->names.size();
  for (int32_t i = 0; i < *n && i < cap; i++) {
    if (names) {
      std::strncpy(names + 32 * i, c->names[i].c_str(), 31);
      names[32 * i + 31] = 0;
    }
    if (counts) counts[i] = c->counts[i];
    if (total_ms) total_ms[i] = c->ms[i];
  }
  return 0;
}

int chan_sync(chan_ctx *c) {
  if (!c) return fail(CHAN_E_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int chan_fft_selftest(int32_t n, int32_t sign, int32_t batch, double *inout) {
  if (!inout || batch < 1 || (sign != 1 && sign != -1)) return fail(CHAN_E_ARG, "bad argument");
  if (!fft_size_supported(n)) return fail(CHAN_E_SIZE, "size %d not supported", n);
  std::vector<cx<double>> tw;
  twiddles(tw, n);
  cx<double> *d = nullptr, *t = nullptr;
  const size_t bytes = sizeof(cx<double>) * (size_t)n * batch;
  HIPCHK(hipMalloc(&d, bytes));
  hipError_t e = hipMalloc(&t, sizeof(cx<double>) * tw.size());
  if (e == hipSuccess) e = hipMemcpy(d, inout, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t, tw.data(), sizeof(cx<double>) * tw.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_fft_test(n, sign, 0, d, t, batch);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(inout, d, bytes, hipMemcpyDeviceToHost);
  hipFree(d);
  if (t) hipFree(t);
  if (e != hipSuccess) return fail((int)e, "fft selftest: %s", hipGetErrorString(e));
  return 0;
}

int chan_fft768_fs_test(int32_t sign, int32_t batch, int32_t inner, int32_t reps, double *inout, double *ms) {
  if (!inout || !ms || batch < 1 || reps < 1 || inner < 1 || inner > 8 || (sign != 1 && sign != -1))
    return fail(CHAN_E_ARG, "bad argument");
  cx<double> tw[fft::FS_TW];
  fft::build_fs768_twiddles(tw);
  const size_t bytes = sizeof(cx<double>) * 768 * (size_t)batch;
  cx<double> *a = nullptr, *b = nullptr, *t = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  hipError_t e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMalloc(&t, sizeof tw);
  if (e == hipSuccess) e = hipMemcpy(a, inout, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t, tw, sizeof tw, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_fs768_test(a, b, t, batch, sign, 1);
  if (e == hipSuccess) e = hipMemcpy(inout, b, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_fs768_test(a, b, t, batch, sign, inner);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ft = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ft, e0, e1);
  *ms = ft / reps;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(a);
  if (b) hipFree(b);
  if (t) hipFree(t);
  if (e != hipSuccess) return fail((int)e, "chan_fft768_fs_test: %s", hipGetErrorString(e));
  return 0;
}

int chan_fft_bench(int32_t n, int32_t sign, int32_t batch, int32_t variant, int32_t reps, double *inout, double *ms) {
  const int inner = 1 + (variant >> 4);  // transforms per load/store pair
  if (!inout || !ms || batch < 1 || reps < 1 || variant < 0 || inner > 64 || (sign != 1 && sign != -1))
    return fail(CHAN_E_ARG, "bad argument");
  if (n != 384 && n != 768 && n != 1536) return fail(CHAN_E_SIZE, "size %d not benchmarked", n);
  std::vector<cx<double>> tw;
  twiddles(tw, n);
  const size_t bytes = sizeof(cx<double>) * (size_t)n * batch;
  cx<double> *a = nullptr, *b = nullptr, *t = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  hipError_t e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMalloc(&t, sizeof(cx<double>) * tw.size());
  if (e == hipSuccess) e = hipMemcpy(a, inout, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t, tw.data(), sizeof(cx<double>) * tw.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_fft_bench(n, sign, variant & 15, a, b, t, batch, 1);
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_fft_bench(n, sign, variant & 15, a, b, t, batch, inner);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ft = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ft, e0, e1);
  if (e == hipSuccess) e = hipMemcpy(inout, b, bytes, hipMemcpyDeviceToHost);
  *ms = ft / reps;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(a);
  if (b) hipFree(b);
  if (t) hipFree(t);
  if (e != hipSuccess) return fail((int)e, "fft bench: %s", hipGetErrorString(e));
  return 0;
}

int chan_copy_selftest(int64_t bytes, int32_t reps, double *ms) {
  if (bytes <= 0 || bytes % 32 != 0 || reps < 1 || !ms) return fail(CHAN_E_ARG, "bad argument");
  double4 *a = nullptr, *b = nullptr;
  const size_t n = (size_t)bytes / sizeof(double4);
  HIPCHK(hipMalloc(&a, n * sizeof(double4)));
  hipError_t e = hipMalloc(&b, n * sizeof(double4));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipMemset(a, 0, n * sizeof(double4));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, 0, a, b, n);  // warm-up
    e = hipEventRecord(e0, 0);
  }
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, 0, a, b, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  *ms = t / reps;
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(a);
  if (b) hipFree(b);
  if (e != hipSuccess) return fail((int)e, "copy selftest: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
