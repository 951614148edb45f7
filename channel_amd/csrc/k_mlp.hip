// k_mlp.hip -- the learned subgrid-stress closure: an MLP 46 -> 64 -> 64 ->
// 64 -> 6 (ReLU) with BF16 weights and activations and FP32 accumulation on
// the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16).  The reference has no
// counterpart; the network maps box-filtered velocity gradients to the six
// subgrid stresses.
//
// Matrix-core fragments (16x16x32, one wavefront):
//   A (16 x 32): lane l holds A[l % 16][8 (l / 16) + j], j = 0..7
//   B (32 x 16): lane l holds B[8 (l / 16) + j][l % 16]
//   D (16 x 16): lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
// Layers are padded to 64 inputs (46 features + zeros) and 16 outputs (6 +
// zeros).  Weights are kept twice in BF16: Woi[l][out][in] feeds the forward
// pass (8 consecutive inputs per lane), Wio[l][in][out] the backward data
// pass; biases stay FP32.
//
// Kernels: k_mlp_fwd (inference, or training forward storing activations and
// the loss gradient), k_mlp_bwd (dz of one layer from the next), k_mlp_wgrad
// (weight and bias gradients of one layer as per-block partial sums over
// samples), k_mlp_reduce (sum of the partials in block order: deterministic)
// and k_mlp_adam (Adam on FP32 master weights, refreshing the BF16 copies).
#include "chan_internal.hpp"
#include "mlp_internal.hpp"

namespace chan {
namespace mlp {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 load8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; j++) z[j] = (__bf16)0.0f;
  return z;
}

// ------------------------------------------------------------ forward --
// One wave per 16-sample tile.  TRAIN: stores the hidden activations act[l]
// ([N][64] BF16, l = 0..2 for layers 1..3) and dy = 2 (y - t) / (6 N) as BF16
// [N][16], and the block's squared-error sum; otherwise stores y [N][6].
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_fwd(Weights w, const __bf16 *__restrict__ X, const float *__restrict__ Tg,
                                                 int n, float inv_count, float *__restrict__ Y,
                                                 __bf16 *__restrict__ act, __bf16 *__restrict__ dy,
                                                 float *__restrict__ loss_part) {
  __shared__ __align__(16) __bf16 woi[WSIZE];
  __shared__ __align__(16) __bf16 tile[4][16 * 64];
  __shared__ float lred[4];
  for (int i = threadIdx.x; i < WSIZE / 8; i += 256)
    reinterpret_cast<bf16x8 *>(woi)[i] = reinterpret_cast<const bf16x8 *>(w.woi)[i];
  __syncthreads();
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, c16 = l % 16, q = l / 16;
  __bf16 *tl = tile[wv];
  float lsum = 0.0f;
  const int ntile = (n + 15) / 16;
  for (int tt = blockIdx.x * 4 + wv; tt < ntile; tt += gridDim.x * 4) {
    const int s0 = tt * 16;
    // input fragments straight from global memory (rows past n read row n-1, masked later)
    bf16x8 a[2];
    {
      const int row = min(s0 + c16, n - 1);
      a[0] = load8(X + (size_t)row * KP + 8 * q);
      a[1] = load8(X + (size_t)row * KP + 32 + 8 * q);
    }
#pragma unroll
    for (int L = 0; L < 4; L++) {
      const int nout = L < 3 ? H : OP;
      const __bf16 *W = woi + WOFF[L];  // [out][64]
      const float *b = w.bias + BOFF[L];
      f32x4 acc[4];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        if (nt * 16 >= nout) break;
        acc[nt] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) acc[nt] = mma(a[ks], load8(W + (nt * 16 + c16) * KP + ks * 32 + 8 * q), acc[nt]);
      }
      if (L < 3) {
        // bias, ReLU, BF16; the tile goes through LDS into the A layout
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
          const int col = nt * 16 + c16;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const float z = acc[nt][i] + b[col];
            tl[(4 * q + i) * 64 + col] = (__bf16)(z > 0.0f ? z : 0.0f);
          }
        }
        wave_sync();
        a[0] = load8(tl + c16 * 64 + 8 * q);
        a[1] = load8(tl + c16 * 64 + 32 + 8 * q);
        if constexpr (TRAIN) {
          const int row = s0 + c16;
          if (row < n) {
            __bf16 *dst = act + ((size_t)L * n + row) * KP;
            *reinterpret_cast<bf16x8 *>(dst + 8 * q) = a[0];
            *reinterpret_cast<bf16x8 *>(dst + 32 + 8 * q) = a[1];
          }
        }
        wave_sync();
      } else {
        // output layer: lane holds y[4q + i][c16]
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int row = s0 + 4 * q + i;
          const float y = acc[0][i] + b[c16];
          if (row < n && c16 < OUT) {
            if constexpr (TRAIN) {
              const float d = y - Tg[(size_t)row * OUT + c16];
              lsum += d * d;
              dy[(size_t)row * OP + c16] = (__bf16)(2.0f * d * inv_count);
            } else {
              Y[(size_t)row * OUT + c16] = y;
            }
          } else if (TRAIN && row < n) {
            dy[(size_t)row * OP + c16] = (__bf16)0.0f;
          }
        }
      }
    }
  }
  if constexpr (TRAIN) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (l == 0) lred[wv] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (lred[0] + lred[1]) + (lred[2] + lred[3]);
  }
}

// ------------------------------------------------------ backward (data) --
// dz of hidden layer L (0..2): dh = dz_next W_next^T, masked by act_L > 0.
// dz_next is dy ([N][16]) for L = 2, else dz of layer L+1 ([N][64]).
__global__ __launch_bounds__(256) void k_mlp_bwd(Weights w, int L, int n, const __bf16 *__restrict__ act,
                                                 const __bf16 *__restrict__ dnext, __bf16 *__restrict__ dz) {
  __shared__ __align__(16) __bf16 wio[64 * 64];
  const int Lw = L + 1;                    // the layer whose weights carry the gradient back
  const int kout = Lw < 3 ? H : OP;        // its outputs: the K of this product
  for (int i = threadIdx.x; i < 64 * kout / 8; i += 256)
    reinterpret_cast<bf16x8 *>(wio)[i] = reinterpret_cast<const bf16x8 *>(w.wio + WOFF[Lw])[i];  // [in][kout]
  __syncthreads();
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, c16 = l % 16, q = l / 16;
  const int ntile = (n + 15) / 16;
  const int dstride = Lw < 3 ? KP : OP;
  for (int tt = blockIdx.x * 4 + wv; tt < ntile; tt += gridDim.x * 4) {
    const int s0 = tt * 16;
    const int row = min(s0 + c16, n - 1);
    bf16x8 a[2];
    a[0] = 8 * q < kout ? load8(dnext + (size_t)row * dstride + 8 * q) : zero8();
    a[1] = kout > 32 ? load8(dnext + (size_t)row * dstride + 32 + 8 * q) : zero8();
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const int col = nt * 16 + c16;  // input of layer Lw = output of layer L
      f32x4 acc = {0, 0, 0, 0};
      acc = mma(a[0], 8 * q < kout ? load8(wio + col * kout + 8 * q) : zero8(), acc);
      if (kout > 32) acc = mma(a[1], load8(wio + col * kout + 32 + 8 * q), acc);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = s0 + 4 * q + i;
        if (r < n) {
          const float h = (float)act[((size_t)L * n + r) * KP + col];
          dz[(size_t)r * KP + col] = (__bf16)(h > 0.0f ? acc[i] : 0.0f);
        }
      }
    }
  }
}

// ---------------------------------------------------- weight gradients --
// dW_L = in_L^T dz_L and db_L = sum dz_L over the block's samples, with
// in_0 = X and in_L = act_(L-1); dz_3 = dy.  Each block writes its partial
// sums (dW [64][nout] then db [nout]) to part + blockIdx.x * PSIZE(L).
__global__ __launch_bounds__(256) void k_mlp_wgrad(int L, int n, const __bf16 *__restrict__ X,
                                                   const __bf16 *__restrict__ act, const __bf16 *__restrict__ dz,
                                                   float *__restrict__ part, int chunk) {
  __shared__ __align__(16) __bf16 tin[4][32 * 64];   // [sample][in] per wave
  __shared__ __align__(16) __bf16 tdz[4][64 * 32];   // [out][sample] per wave
  __shared__ float red[64 * 64 + 64];
  const int nout = L < 3 ? H : OP, dstride = L < 3 ? KP : OP;
  const __bf16 *in = L == 0 ? X : act + (size_t)(L - 1) * n * KP;
  const int wv = threadIdx.x / 64, l = threadIdx.x % 64, c16 = l % 16, q = l / 16;
  const int NT = nout / 16;
  f32x4 acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0, 0, 0, 0};
  float bsum = 0.0f;  // lane l < nout: sum of dz[.][l]
  const int s_begin = blockIdx.x * chunk, s_end = min(n, s_begin + chunk);
  for (int s0 = s_begin + wv * 32; s0 < s_end; s0 += 4 * 32) {
    // stage 32 samples: inputs row-major, dz transposed; rows past the chunk are zero
    for (int e = l; e < 32 * 8; e += 64) {
      const int r = e / 8, k8 = (e % 8) * 8, s = s0 + r;
      *reinterpret_cast<bf16x8 *>(&tin[wv][r * 64 + k8]) = s < s_end ? load8(in + (size_t)s * KP + k8) : zero8();
    }
    for (int e = l; e < 32 * nout; e += 64) {
      const int r = e / nout, o = e % nout, s = s0 + r;
      const __bf16 v = s < s_end ? dz[(size_t)s * dstride + o] : (__bf16)0.0f;
      tdz[wv][o * 32 + r] = v;
    }
    wave_sync();
    // A[m = in][k = sample] = tin[k][m]; B[k = sample][n = out] = tdz[n][k]
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; j++) a[j] = tin[wv][(8 * q + j) * 64 + mt * 16 + c16];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        if (nt >= NT) break;
        acc[mt][nt] = mma(a, load8(&tdz[wv][(nt * 16 + c16) * 32 + 8 * q]), acc[mt][nt]);
      }
    }
    if (l < nout) {
      float s = 0.0f;
#pragma unroll
      for (int r = 0; r < 32; r++) s += (float)tdz[wv][l * 32 + r];
      bsum += s;
    }
    wave_sync();
  }
  // block reduction of the four waves' sums, in wave order
  for (int i = threadIdx.x; i < 64 * 64 + 64; i += 256) red[i] = 0.0f;
  __syncthreads();
  for (int w2 = 0; w2 < 4; w2++) {
    if (wv == w2) {
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
          if (nt >= NT) break;
#pragma unroll
          for (int i = 0; i < 4; i++) red[(mt * 16 + 4 * q + i) * nout + nt * 16 + c16] += acc[mt][nt][i];
        }
      if (l < nout) red[64 * nout + l] += bsum;
    }
    __syncthreads();
  }
  float *dst = part + (size_t)blockIdx.x * (64 * nout + nout);
  for (int i = threadIdx.x; i < 64 * nout + nout; i += 256) dst[i] = red[i];
}

// sum of nblk partial vectors of length len, in block order
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ part, int nblk, int len,
                                                    float *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * len + i];
  out[i] = s;
}

// --------------------------------------------------------------- Adam --
// master/m/v/g in the padded kernel layout (GSIZE floats); refreshes the
// BF16 weight copies in both orientations and the FP32 biases.  Padding
// entries have zero gradient and stay zero.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ master, float *__restrict__ m1,
                                                  float *__restrict__ m2, const float *__restrict__ g, Weights w,
                                                  float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= GSIZE) return;
  const float gi = g[i];
  const float mi = b1 * m1[i] + (1.0f - b1) * gi;
  const float vi = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = mi;
  m2[i] = vi;
  const float p = master[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  master[i] = p;
  write_param(w, i, p);
}

// copies the padded FP32 master to the BF16 / bias arrays (after set_params)
__global__ __launch_bounds__(256) void k_mlp_refresh(const float *__restrict__ master, Weights w) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < GSIZE) write_param(w, i, master[i]);
}

}  // namespace mlp
}  // namespace chan

// ================================================================ C-ABI ==
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/channel_closure.h"
#include "../../include/channel_hip.h"

using namespace chan;
using namespace chan::mlp;

static_assert(NPARAM == CHAN_MLP_NPARAM, "canonical parameter count");

struct chan_mlp {
  int64_t cap = 0, n = 0;
  int dev = 0;                // device the model lives on
  hipStream_t st = nullptr;
  __bf16 *X = nullptr;        // [cap][64]
  float *T = nullptr;         // [cap][6]
  __bf16 *act = nullptr;      // [3][cap][64]
  __bf16 *dz = nullptr;       // [3][cap][64]
  __bf16 *dy = nullptr;       // [cap][16]
  float *Y = nullptr;         // [cap][6]
  float *master = nullptr, *m1 = nullptr, *m2 = nullptr, *g = nullptr;  // [GSIZE]
  float *part = nullptr;      // wgrad partials
  float *lpart = nullptr;     // loss partials
  float *losses = nullptr;    // per training step
  __bf16 *woi = nullptr, *wio = nullptr;
  float *bias = nullptr;
  int64_t adam_t = 0;
  int losses_cap = 0;
};

// Staging of chan_mlp_forward_noalloc: the BF16 inputs [CHAN_MLP_NOALLOC_MAX][64]
// (8 MiB), one copy per GPU for the life of the process.  Host code never
// passes this name to a kernel: it holds a host-side stand-in, and the device
// address comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) __bf16 scratch[(size_t)CHAN_MLP_NOALLOC_MAX * KP];

namespace {

constexpr int MAXDEV = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAXDEV] = {};  // device address of `scratch` per device

// Resolves the address of `scratch` on the current device, once per device
// (called by chan_mlp_create with the model's device current).
int noalloc_scratch(int dev, void **out) {
  if (dev < 0 || dev >= MAXDEV) return set_error(CHAN_E_ARG, "device %d beyond the %d the library tracks", dev, MAXDEV);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return set_error((int)(e != hipSuccess ? e : hipErrorInvalidSymbol), "hipGetSymbolAddress of the staging array failed: %s", hipGetErrorString(e));
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return 0;
}

constexpr int MAXBLK = 1024;  // blocks of the per-sample kernels
constexpr int WBLK = 256;     // blocks of the weight-gradient kernels

#define MCHK(expr)                                                                                          \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return set_error((int)e_, "%s: %s", #expr, hipGetErrorString(e_));               \
  } while (0)

template <class T>
int mal(T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(T) + 256);
  if (e != hipSuccess) return set_error((int)e, "hipMalloc of %zu bytes failed", n * sizeof(T));
  return 0;
}

Weights weights(chan_mlp *m) { return Weights{m->woi, m->wio, m->bias}; }

__global__ void k_to_bf16(const float *__restrict__ x, __bf16 *__restrict__ xb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * KP) return;
  const int64_t r = i / KP;
  const int c = (int)(i % KP);
  xb[i] = (__bf16)(c < IN ? x[r * IN + c] : 0.0f);
}

int nblk_rows(int64_t n) { return (int)std::min<int64_t>(MAXBLK, std::max<int64_t>(1, ((n + 15) / 16 + 3) / 4)); }

// gradient of the minibatch [first, first+count) into m->g; loss partials in m->lpart
int grad_pass(chan_mlp *m, int64_t first, int count) {
  const Weights w = weights(m);
  const __bf16 *X = m->X + first * KP;
  const float *T = m->T + first * OUT;
  const int nb = nblk_rows(count);
  hipLaunchKernelGGL((k_mlp_fwd<true>), dim3(nb), dim3(256), 0, m->st, w, X, T, count, 1.0f / (6.0f * count),
                     (float *)nullptr, m->act, m->dy, m->lpart);
  MCHK(hipGetLastError());
  for (int L = 2; L >= 0; L--) {
    const __bf16 *dn = L == 2 ? m->dy : m->dz + (size_t)(L + 1) * m->cap * KP;
    hipLaunchKernelGGL(k_mlp_bwd, dim3(nb), dim3(256), 0, m->st, w, L, count, m->act, dn,
                       m->dz + (size_t)L * m->cap * KP);
    MCHK(hipGetLastError());
  }
  // weight gradients: chunks of whole 128-sample groups
  const int wb = (int)std::min<int64_t>(WBLK, (count + 127) / 128);
  const int chunk = (int)(((count + wb - 1) / wb + 127) / 128 * 128);
  for (int L = 0; L < 4; L++) {
    const __bf16 *dzL = L == 3 ? m->dy : m->dz + (size_t)L * m->cap * KP;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(wb), dim3(256), 0, m->st, L, count, X, m->act, dzL, m->part, chunk);
    MCHK(hipGetLastError());
    const int len = psize(L);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((len + 255) / 256), dim3(256), 0, m->st, m->part, wb, len, m->g + goff(L));
    MCHK(hipGetLastError());
  }
  return 0;
}

}  // namespace

hipError_t chan::mlp::forward_device(chan_mlp *m, const __bf16 *X, int64_t n, float *Y, hipStream_t st) {
  if (n < 1 || n > ((int64_t)1 << 31) / KP) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_mlp_fwd<false>), dim3(nblk_rows(n)), dim3(256), 0, st, weights(m), X, (const float *)nullptr,
                     (int)n, 0.0f, Y, (__bf16 *)nullptr, (__bf16 *)nullptr, (float *)nullptr);
  return hipGetLastError();
}

extern "C" {

int chan_mlp_create(int64_t capacity, chan_mlp **out) {
  if (!out || capacity < 1 || capacity > ((int64_t)1 << 31) / KP) return set_error(CHAN_E_ARG, "bad capacity");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(CHAN_E_NODEV, "no HIP device");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(CHAN_E_NODEV, "no current HIP device");
  void *scr = nullptr;
  int rc = noalloc_scratch(dev, &scr);  // once per device, at set-up
  if (rc) return rc;
  (void)scr;
  chan_mlp *m = new chan_mlp();
  m->cap = capacity;
  m->dev = dev;
  auto bail = [&](int code) {
    chan_mlp_destroy(m);
    return code;
  };
  if (hipStreamCreate(&m->st) != hipSuccess) return bail(set_error(CHAN_E_STATE, "hipStreamCreate failed"));
  const size_t c = (size_t)capacity;
  if ((rc = mal(&m->X, c * KP)) || (rc = mal(&m->T, c * OUT)) || (rc = mal(&m->act, 3 * c * KP)) ||
      (rc = mal(&m->dz, 3 * c * KP)) || (rc = mal(&m->dy, c * OP)) || (rc = mal(&m->Y, c * OUT)) ||
      (rc = mal(&m->master, GSIZE)) || (rc = mal(&m->m1, GSIZE)) || (rc = mal(&m->m2, GSIZE)) ||
      (rc = mal(&m->g, GSIZE)) || (rc = mal(&m->part, (size_t)WBLK * psize(0))) || (rc = mal(&m->lpart, MAXBLK)) ||
      (rc = mal(&m->woi, WSIZE)) || (rc = mal(&m->wio, WSIZE)) || (rc = mal(&m->bias, BSIZE)))
    return bail(rc);
  if (hipMemset(m->master, 0, sizeof(float) * GSIZE) != hipSuccess ||
      hipMemset(m->m1, 0, sizeof(float) * GSIZE) != hipSuccess || hipMemset(m->m2, 0, sizeof(float) * GSIZE) != hipSuccess ||
      hipMemset(m->woi, 0, sizeof(__bf16) * WSIZE) != hipSuccess || hipMemset(m->wio, 0, sizeof(__bf16) * WSIZE) != hipSuccess ||
      hipMemset(m->bias, 0, sizeof(float) * BSIZE) != hipSuccess)
    return bail(set_error(CHAN_E_STATE, "hipMemset failed"));
  *out = m;
  return 0;
}

int chan_mlp_destroy(chan_mlp *m) {
  if (!m) return 0;
  if (m->st) hipStreamSynchronize(m->st);
  void *bufs[] = {m->X, m->T, m->act, m->dz, m->dy, m->Y, m->master, m->m1, m->m2, m->g, m->part, m->lpart,
                  m->losses, m->woi, m->wio, m->bias};
  for (void *b : bufs)
    if (b) hipFree(b);
  if (m->st) hipStreamDestroy(m->st);
  delete m;
  return 0;
}

int chan_mlp_set_params(chan_mlp *m, const float *params) {
  if (!m || !params) return set_error(CHAN_E_ARG, "null argument");
  std::vector<float> pad(GSIZE, 0.0f);
  for (int c = 0; c < NPARAM; c++) pad[canon_to_padded(c)] = params[c];
  MCHK(hipMemcpyAsync(m->master, pad.data(), sizeof(float) * GSIZE, hipMemcpyHostToDevice, m->st));
  hipLaunchKernelGGL(k_mlp_refresh, dim3((GSIZE + 255) / 256), dim3(256), 0, m->st, m->master, weights(m));
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

int chan_mlp_get_params(chan_mlp *m, float *params) {
  if (!m || !params) return set_error(CHAN_E_ARG, "null argument");
  std::vector<float> pad(GSIZE);
  MCHK(hipMemcpyAsync(pad.data(), m->master, sizeof(float) * GSIZE, hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  for (int c = 0; c < NPARAM; c++) params[c] = pad[canon_to_padded(c)];
  return 0;
}

int chan_mlp_set_data(chan_mlp *m, const float *X, const float *T, int64_t n) {
  if (!m || !X || !T || n < 1 || n > m->cap) return set_error(CHAN_E_ARG, "bad argument");
  float *tmp = nullptr;
  int rc;
  if ((rc = mal(&tmp, (size_t)n * IN))) return rc;
  hipError_t e = hipMemcpyAsync(tmp, X, sizeof(float) * n * IN, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n * KP + 255) / 256)), dim3(256), 0, m->st, tmp, m->X, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(m->T, T, sizeof(float) * n * OUT, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = hipMemsetAsync(m->m1, 0, sizeof(float) * GSIZE, m->st);
  if (e == hipSuccess) e = hipMemsetAsync(m->m2, 0, sizeof(float) * GSIZE, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  hipFree(tmp);
  if (e != hipSuccess) return set_error((int)e, "set_data: %s", hipGetErrorString(e));
  m->n = n;
  m->adam_t = 0;
  return 0;
}

int chan_mlp_forward(chan_mlp *m, const float *X, int64_t n, float *Y) {
  if (!m || !X || !Y || n < 1 || n > m->cap) return set_error(CHAN_E_ARG, "bad argument");
  float *tmp = nullptr;
  int rc;
  if ((rc = mal(&tmp, (size_t)n * IN))) return rc;
  // the inputs go through the activation buffer, leaving the resident data alone
  __bf16 *xb = m->act;
  hipError_t e = hipMemcpyAsync(tmp, X, sizeof(float) * n * IN, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n * KP + 255) / 256)), dim3(256), 0, m->st, tmp, xb, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL((k_mlp_fwd<false>), dim3(nblk_rows(n)), dim3(256), 0, m->st, weights(m), xb,
                       (const float *)nullptr, (int)n, 0.0f, m->Y, (__bf16 *)nullptr, (__bf16 *)nullptr,
                       (float *)nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(Y, m->Y, sizeof(float) * n * OUT, hipMemcpyDeviceToHost, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  hipFree(tmp);
  if (e != hipSuccess) return set_error((int)e, "forward: %s", hipGetErrorString(e));
  return 0;
}

int chan_mlp_forward_noalloc(chan_mlp *m, const float *dX, int64_t n, float *dY) {
  if (!m || !dX || !dY || n < 1) return set_error(CHAN_E_ARG, "bad argument");
  if (n > CHAN_MLP_NOALLOC_MAX)
    return set_error(CHAN_E_ARG, "chan_mlp_forward_noalloc: %lld rows exceed its staging of %d rows; split the call",
                     (long long)n, CHAN_MLP_NOALLOC_MAX);
  int dev = -1;
  MCHK(hipGetDevice(&dev));
  if (dev != m->dev) return set_error(CHAN_E_STATE, "chan_mlp_forward_noalloc: device %d is current, the model lives on %d", dev, m->dev);
  __bf16 *xb = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    xb = static_cast<__bf16 *>(g_scratch[dev]);
  }
  if (!xb) return set_error(CHAN_E_STATE, "chan_mlp_forward_noalloc: no staging address for device %d", dev);
  // the same two kernels and launch shapes as chan_mlp_forward
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n * KP + 255) / 256)), dim3(256), 0, m->st, dX, xb, n);
  MCHK(hipGetLastError());
  hipLaunchKernelGGL((k_mlp_fwd<false>), dim3(nblk_rows(n)), dim3(256), 0, m->st, weights(m), xb,
                     (const float *)nullptr, (int)n, 0.0f, dY, (__bf16 *)nullptr, (__bf16 *)nullptr, (float *)nullptr);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->st));
  return 0;
}

int chan_mlp_grad(chan_mlp *m, int64_t first, int64_t count, float *grad, double *loss) {
  if (!m || !grad || !loss || first < 0 || count < 1 || first + count > m->n) return set_error(CHAN_E_ARG, "bad argument");
  int rc;
  if ((rc = grad_pass(m, first, (int)count))) return rc;
  std::vector<float> pad(GSIZE), lp(MAXBLK);
  const int nb = nblk_rows(count);
  MCHK(hipMemcpyAsync(pad.data(), m->g, sizeof(float) * GSIZE, hipMemcpyDeviceToHost, m->st));
  MCHK(hipMemcpyAsync(lp.data(), m->lpart, sizeof(float) * nb, hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  for (int c = 0; c < NPARAM; c++) grad[c] = pad[canon_to_padded(c)];
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += lp[b];
  *loss = s / (6.0 * (double)count);
  return 0;
}

int chan_mlp_train(chan_mlp *m, int32_t steps, int32_t batch, float lr, float *losses, double *ms_per_step) {
  if (!m || steps < 1 || batch < 16 || batch > m->n || !(lr > 0)) return set_error(CHAN_E_ARG, "bad argument");
  if (m->losses_cap < steps) {
    if (m->losses) hipFree(m->losses);
    m->losses = nullptr;
    int rc;
    if ((rc = mal(&m->losses, (size_t)steps))) return rc;
    m->losses_cap = steps;
  }
  hipEvent_t e0, e1;
  MCHK(hipEventCreate(&e0));
  MCHK(hipEventCreate(&e1));
  MCHK(hipEventRecord(e0, m->st));
  int64_t off = 0;
  for (int s = 0; s < steps; s++) {
    if (off + batch > m->n) off = 0;
    int rc;
    if ((rc = grad_pass(m, off, batch))) return rc;
    const int nb = nblk_rows(batch);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(1), dim3(256), 0, m->st, m->lpart, nb, 1, m->losses + s);
    MCHK(hipGetLastError());
    m->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)m->adam_t);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)m->adam_t);
    hipLaunchKernelGGL(k_mlp_adam, dim3((GSIZE + 255) / 256), dim3(256), 0, m->st, m->master, m->m1, m->m2, m->g,
                       weights(m), lr, b1, b2, 1e-8f, c1, c2);
    MCHK(hipGetLastError());
    off += batch;
  }
  MCHK(hipEventRecord(e1, m->st));
  MCHK(hipEventSynchronize(e1));
  float ms = 0;
  MCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ms_per_step) *ms_per_step = ms / steps;
  if (losses) {
    std::vector<float> l(steps);
    MCHK(hipMemcpy(l.data(), m->losses, sizeof(float) * steps, hipMemcpyDeviceToHost));
    for (int s = 0; s < steps; s++) losses[s] = l[s] / (6.0f * batch);
  }
  return 0;
}

int chan_mlp_bench_forward(chan_mlp *m, int32_t reps, double *ms) {
  if (!m || reps < 1 || !ms || m->n < 1) return set_error(CHAN_E_ARG, "bad argument");
  hipEvent_t e0, e1;
  MCHK(hipEventCreate(&e0));
  MCHK(hipEventCreate(&e1));
  const int n = (int)m->n;
  hipLaunchKernelGGL((k_mlp_fwd<false>), dim3(nblk_rows(n)), dim3(256), 0, m->st, weights(m), m->X,
                     (const float *)nullptr, n, 0.0f, m->Y, (__bf16 *)nullptr, (__bf16 *)nullptr, (float *)nullptr);
  MCHK(hipEventRecord(e0, m->st));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL((k_mlp_fwd<false>), dim3(nblk_rows(n)), dim3(256), 0, m->st, weights(m), m->X,
                       (const float *)nullptr, n, 0.0f, m->Y, (__bf16 *)nullptr, (__bf16 *)nullptr, (float *)nullptr);
  MCHK(hipEventRecord(e1, m->st));
  MCHK(hipEventSynchronize(e1));
  float t = 0;
  MCHK(hipEventElapsedTime(&t, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms = t / reps;
  return 0;
}

}  // extern "C"
