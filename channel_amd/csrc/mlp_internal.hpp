// mlp_internal.hpp -- layouts of the closure MLP shared by the kernels and
// the C-ABI (46 -> 64 -> 64 -> 64 -> 6, ReLU).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

namespace chan {
namespace mlp {

constexpr int IN = 46, H = 64, OUT = 6;
constexpr int KP = 64;  // padded inputs of every layer
constexpr int OP = 16;  // padded outputs of the last layer
// canonical FP32 parameters: W1 [46][64], b1, W2 [64][64], b2, W3 [64][64], b3, W4 [64][6], b4
constexpr int NPARAM = IN * H + H + H * H + H + H * H + H + H * OUT + OUT;
// padded kernel layout, per layer L: dW [64][nout] (in-major), then db [nout]
constexpr int nout_of(int L) { return L < 3 ? H : OP; }
constexpr int psize(int L) { return 64 * nout_of(L) + nout_of(L); }
constexpr int goff(int L) { return L == 0 ? 0 : goff(L - 1) + psize(L - 1); }
constexpr int GSIZE = goff(4);
// BF16 weights: per layer a [64][nout] block (wio) or [nout][64] block (woi)
constexpr int woff(int L) { return L == 0 ? 0 : woff(L - 1) + 64 * nout_of(L - 1); }
constexpr int WSIZE = woff(4);
constexpr int boff(int L) { return L == 0 ? 0 : boff(L - 1) + nout_of(L - 1); }
constexpr int BSIZE = boff(4);
__device__ constexpr int WOFF[5] = {woff(0), woff(1), woff(2), woff(3), woff(4)};
__device__ constexpr int BOFF[5] = {boff(0), boff(1), boff(2), boff(3), boff(4)};

struct Weights {
  __bf16 *woi;   // [L][out][64]
  __bf16 *wio;   // [L][64][out]
  float *bias;   // [L][out]
};

// padded parameter index -> BF16 weight copies / FP32 bias
__device__ __forceinline__ void write_param(const Weights &w, int i, float p) {
  int L = 0;
  while (L < 3 && i >= goff(L + 1)) L++;
  const int r = i - goff(L), no = nout_of(L);
  if (r < 64 * no) {
    const int in = r / no, out = r % no;
    const __bf16 b = (__bf16)p;
    w.woi[woff(L) + out * 64 + in] = b;
    w.wio[woff(L) + in * no + out] = b;
  } else {
    w.bias[boff(L) + (r - 64 * no)] = p;
  }
}

// host: canonical index <-> padded index
inline int canon_to_padded(int c) {
  const int din[4] = {IN, H, H, H}, dout[4] = {H, H, H, OUT};
  for (int L = 0; L < 4; L++) {
    const int nw = din[L] * dout[L];
    if (c < nw) return goff(L) + (c / dout[L]) * nout_of(L) + (c % dout[L]);
    c -= nw;
    if (c < dout[L]) return goff(L) + 64 * nout_of(L) + c;
    c -= dout[L];
  }
  return -1;
}

}  // namespace mlp
}  // namespace chan

struct chan_mlp;
namespace chan {
namespace mlp {
// inference of model m on n resident BF16 rows X[n][64] -> Y[n][6] FP32,
// enqueued on stream st (k_mlp.hip)
hipError_t forward_device(chan_mlp *m, const __bf16 *X, int64_t n, float *Y, hipStream_t st);
}  // namespace mlp
}  // namespace chan
