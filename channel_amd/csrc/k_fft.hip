// k_fft.hip -- the convolution products of one RK3 substep as three kernels
// over all y planes (convolutions, dnsdata.f90:487-602):
//
//   K1 k_zinv : V(iy,iz,ix,c) -> zero-pad z to nzd, backward c2c FFT in z
//               (IFT, ffts.f90:93-96)  -> A[plane][c][zb][ix][ZB]
//   K2 k_xpass: per physical-z row: Hermitian c2r in x (RFT, zero modes
//               nx+1..nxd), CFL maximum, the six products, r2c in x keeping
//               modes 0..nx (HFT), times `factor`       -> B[plane][q][zb][ix][ZB]
//   K3 k_zfwd : forward c2c FFT in z, keep modes -nz..nz (FFT + izd)
//                                                       -> P[q][plane][pair]
//
// The reference's pencil transposes (zTOx / xTOz, mpi_transpose.f90:50-117)
// are the layout changes between K1/K2 and K2/K3: both sides move ZB*16
// contiguous bytes per x mode.  Two real x-transforms share one complex FFT
// (u+iv on the way in, product pairs on the way out).
#include "chan_internal.hpp"

namespace chan {

// ------------------------------------------------------------------- K1 --
// the z transforms are bound by HBM traffic; the padded split exchange keeps
// their LDS footprint small
constexpr int ZXCH = fft::XCH_PAD;

// A z block is 32 (FP32) or 64 (FP64) bytes per x mode, short for memory
// transactions: the z kernels stage through LDS, where the G x modes of a
// block form one contiguous run of G*ZB values per z block, and lanes move
// whole runs (a wave touches 4 runs of 128 or 256 bytes).
template <int NZ, typename R>
constexpr bool z_staged() {
  return true;
}
// team stride of the staging area (a 32-byte shift per team keeps the
// staged writes and reads free of LDS bank conflicts)
template <int NZ>
constexpr int stage_stride() {
  return NZ + 4;
}
// The blocks stage the real and the imaginary parts in two rounds through
// one real-valued area: half the LDS of whole complex values, which would
// hold an FP64 block to three waves per SIMD and an FP32 block to six (four
// and eight with the two rounds: the FP64 step 1.5% and the FP32 z kernels
// 4-9% faster)
template <int NZ, typename R>
constexpr int zbuf_reals() {
  constexpr int G = 256 / fft::Plan<NZ>::T;
  constexpr int fftr = G * fft::lds_reals<NZ, ZXCH>();
  constexpr int stg = z_staged<NZ, R>() ? G * stage_stride<NZ>() : 0;
  return fftr > stg ? fftr : stg;
}

template <int NZ, typename R>
__global__ __launch_bounds__(256) void k_zinv(const cx<double> *__restrict__ V, cx<R> *__restrict__ A,
                                              const cx<R> *__restrict__ tw, Geom g, int plane0) {
  using PL = fft::Plan<NZ>;
  constexpr int T = PL::T, E = PL::E, G = 256 / T, SP = stage_stride<NZ>();
  __shared__ __align__(16) R lds[zbuf_reals<NZ, R>()];
  const int team = threadIdx.x / T, t = threadIdx.x % T;
  const int nxg = (g.Nx + G - 1) / G;
  int b = blockIdx.x;
  const int xg = b % nxg;
  b /= nxg;
  const int c = b % 3, pl = b / 3;
  const int ix = xg * G + team;
  const bool act = ix < g.Nx;
  const cx<double> *col = V + ((size_t)c * g.Ny + (plane0 + pl)) * g.ps + (size_t)(act ? ix : 0) * g.Nz;
  cx<R> e[E];
#pragma unroll
  for (int s = 0; s < E; s++) {
    const int z = t + T * s;
    const int iz = z <= g.nz ? z + g.nz : (z >= NZ - g.nz ? z - NZ + g.nz : -1);
    cx<double> v = {0.0, 0.0};
    if (act && iz >= 0) v = col[iz];
    e[s] = cx<R>{(R)v.x, (R)v.y};
  }
  fft::fft_slots<NZ, +1, true, ZXCH>(e, lds + team * fft::lds_reals<NZ, ZXCH>(), tw, t);
  cx<R> *dst = A + (size_t)(pl * 3 + c) * NZ * g.Nx;
  if constexpr (z_staged<NZ, R>()) {
    const int ix0 = xg * G, nact = min(G, g.Nx - ix0);
    R out[2][E];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      __syncthreads();  // the transforms' exchanges, or the previous round's reads, are done
#pragma unroll
      for (int s = 0; s < E; s++) lds[team * SP + t + T * s] = h ? e[s].y : e[s].x;
      __syncthreads();
#pragma unroll
      for (int i = 0; i < E; i++) {
        const int el = (int)threadIdx.x + 256 * i;  // G*NZ = 256*E elements per block
        const int zb = el / (G * ZB), w = el % (G * ZB), tm = w / ZB, r = w % ZB;
        out[h][i] = lds[min(tm, nact - 1) * SP + zb * ZB + r];
      }
    }
#pragma unroll
    for (int i = 0; i < E; i++) {
      const int el = (int)threadIdx.x + 256 * i;
      const int zb = el / (G * ZB), w = el % (G * ZB), tm = w / ZB;
      if (tm < nact) dst[((size_t)zb * g.Nx + ix0) * ZB + w] = cx<R>{out[0][i], out[1][i]};
    }
  } else if (act) {
#pragma unroll
    for (int s = 0; s < E; s++) {
      const int z = t + T * s;
      dst[((size_t)(z / ZB) * g.Nx + ix) * ZB + (z % ZB)] = e[s];
    }
  }
}

// ------------------------------------------------------------------- K2 --
// the x pass runs five transforms per row and is bound by their arithmetic
// and LDS exchanges: radix-12 first pass, whole-complex padded exchanges
// (measured fastest; 2 waves/SIMD either way, u, v, w stay live)
constexpr bool XBIG = true;
constexpr int XXCH = fft::XCH_CPLX | fft::XCH_PAD;

template <int NT>
__device__ __forceinline__ void block_max_atomic(double m, double *red, unsigned long long *out) {
  // a maximum that keeps a NaN
  auto mx = [](double c, double o) { return (o > c || o != o) ? o : c; };
  double r;
  if constexpr (NT % 64 == 0) {
    // within each wave by lane exchanges, then across the waves in LDS
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) m = mx(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) r = mx(r, red[w]);
  } else {
    red[threadIdx.x] = m;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = mx(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    r = red[0];
  }
  // values are >= 0 or NaN, so their bit patterns order as unsigned
  // integers (NaN above +Inf above every finite value)
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(r));
}

// x / d as the division rounds it, from r = 1/d rounded to nearest: q = x r
// and one correction by the exact remainder x - q d.  For finite x and
// normal d the corrected quotient is the correctly rounded one (Markstein's
// theorem: r within half an ulp of 1/d, q within an ulp of x/d); an
// infinite x gives NaN, which the CFL check reports as it reports Inf.
__device__ __forceinline__ double div_by(double x, double d, double r) {
  const double q = x * r;
  return fma(fma(-q, d, x), r, q);
}

// x modes per thread of the x pass: nxd >= 3(nx+1)/2 bounds Nx by M/3
template <int E>
constexpr int xmodes() {
  return (E + 2) / 3;
}

// value of the team-mate holding element M-k of a transform in slot layout:
// for t > 0 it is lane T-t, slot E-1-s; for t = 0 the thread itself, slot
// (E-s) mod E.  Teams of one wavefront at most (the 128-lane teams of the
// 1536-point rows combine their modes in fs1536_mode instead).
template <int T, int E, int MO, typename R>
__device__ __forceinline__ void mirror_pairs(const cx<R> (&e)[E], cx<R> (&m)[MO], int t, int nmode) {
  static_assert(T <= 64, "mirror_pairs works within one wavefront");
  const int lane = (int)(threadIdx.x & 63);
  const int src = lane - t + ((T - t) % T);
#pragma unroll
  for (int s = 0; s < MO; s++) {
    if (s * T < nmode) {  // uniform across the team
      const cx<R> send = e[E - 1 - s];
      const R rx = __shfl(send.x, src, 64), ry = __shfl(send.y, src, 64);
      m[s] = t ? cx<R>{rx, ry} : e[(E - s) % E];
    }
  }
}

// teams per x-pass block: one per row of a z block, or one per row of half a
// z block (TB = 2).  The halves read 32 of each 64 contiguous bytes of the
// z-block interleave; they are numbered 8 blocks apart, so that round-robin
// dispatch over the 8 XCDs puts both on one XCD and its L2.  Half blocks
// serve the 1536-point rows, whose 128-lane teams hold 29 KB of LDS each in
// FP64: two blocks then share a CU and one block's barriers overlap the
// other's work (25% less time per launch than four-row blocks at
// 511x512x511).  FP64 768-point rows gain 2.5% from half blocks; FP32 ones
// lose 7%, and keep whole z blocks.
template <int M, typename R>
constexpr int xteams() {
  return (M == 1536 || (M == 768 && sizeof(R) == 8)) ? 2 : ZB;
}
// MO_: direct-mode slots per lane, xmodes<E>() or fewer when nx + 1 <= T MO_
// (the launcher picks 3 for 768-point rows of at most 192 modes): the
// compiler then knows the middle slots of every backward transform are zero
// and drops the loads, selects and outputs that would only handle them
// FULL: nx + 1 == T MO_ exactly (nx = 191 for 768-point rows), so the mode
// bounds are compile-time constants and their selects and branches fold away
template <int M, typename R, bool PRODUCTS, int MO_ = xmodes<fft::Plan<M>::E>(), bool FULL = false>
__global__ __launch_bounds__((xteams<M, R>() * fft::Plan<M>::T), 2) void k_xpass(const cx<R> *__restrict__ A, cx<R> *__restrict__ B,
                                                                const cx<R> *__restrict__ tw, Geom g, int plane0,
                                                                int cfl_on, const double *__restrict__ dy, double dx,
                                                                double dz, R factor, unsigned long long *__restrict__ cfl_bits,
                                                                int nplanes) {
  using PL = fft::Plan<M>;
  constexpr int TB = xteams<M, R>();
  constexpr int T = PL::T, E = PL::E, NT = TB * T;
  constexpr int MO = MO_;
  // FP64 768-point rows run their transforms in interleaved pairs
  // (fs768_bwd2 / fs768_fwd2): at two waves per SIMD the single transforms
  // leave the VALU waiting on their LDS exchanges
  constexpr bool ILV = M == 768 && sizeof(R) == 8;
  static_assert(MO <= xmodes<E>() && 2 * MO <= E, "direct and mirrored slots");
  // 768-point rows use the four-step transform (fs768_bwd/fwd), 1536-point
  // rows its two-wave form (fs1536_bwd/fwd); the samples between the two
  // directions sit in a permuted layout, which the pointwise CFL and products
  // do not see
  constexpr bool FS = M == 768, FS2 = M == 1536;
  constexpr int TEAM_R = FS ? 2 * fft::FS_LDS : (FS2 ? 4 * fft::FS_LDS : fft::lds_reals<M, XXCH>());
  static_assert(!FS2 || (T == 128 && E == 12 && MO == 4), "the two-wave 1536 transform needs 128-lane teams");
  __shared__ __align__(16) R fl[TB * TEAM_R];
  __shared__ double red[NT % 64 == 0 ? NT / 64 : NT];
  const int Nx = g.Nx, nx = FULL ? fft::Plan<M>::T * MO_ - 1 : g.nx, NZB = g.nzd / ZB;
  const int tid = threadIdx.x, t = tid % T;
  int zb, pl, zi;
  if constexpr (TB < ZB) {
    // half z blocks: the halves of tile 8q + x are blocks 16q + x and
    // 16q + 8 + x, which round-robin dispatch puts on one XCD
    static_assert(TB * 2 == ZB, "half z blocks");
    const int b = blockIdx.x, half = (b >> 3) & 1, tile = ((b >> 4) << 3) | (b & 7);
    if (tile >= NZB * nplanes) return;  // the whole block: padding of the grid
    zb = tile % NZB;
    pl = tile / NZB;
    zi = TB * half + tid / T;
  } else {
    zb = blockIdx.x % NZB;
    pl = blockIdx.x / NZB;
    zi = tid / T;
  }
  const int iy = plane0 + pl - 1;
  R *fr = fl + (tid / T) * TEAM_R;
  cx<R> *frc = reinterpret_cast<cx<R> *>(fr);

  // twiddles in LDS: their reads then wait on the LDS counter, not behind
  // the row loads still in flight on the vector-memory counter.  The
  // 768-point rows need only three table entries per lane, which they read
  // from global memory before their row loads: no staging and no block
  // barrier before the loads are issued.
  constexpr int NTW = FS ? 0 : (FS2 ? fft::FS2_TW : fft::plan_tw_size(M, XBIG));
  __shared__ __align__(16) cx<R> twl[NTW > 0 ? NTW : 1];
  if constexpr (NTW > 0) {
    for (int i = tid; i < NTW; i += NT) twl[i] = tw[i];
    __syncthreads();
  }
  // the four-step transform's twiddle powers, shared by the row's five transforms
  fft::FsTw<R> fw;
  if constexpr (FS)
    fw = fft::fs_twiddles(tw, t & 63);
  else if constexpr (FS2)
    fw = fft::fs_twiddles(twl, t & 63);
  auto bwd = [&](cx<R>(&v)[E]) {
    if constexpr (FS)
      fft::fs768_bwd<R, MO <= 3>(v, frc, fw, t);
    else if constexpr (FS2)
      fft::fs1536_bwd(v, frc, fw, twl, t);
    else
      fft::fft_slots<M, +1, XBIG, XXCH>(v, fr, twl, t);
  };
  auto fwd = [&](cx<R>(&v)[E]) {
    if constexpr (FS)
      fft::fs768_fwd(v, frc, fw, t);
    else if constexpr (FS2)
      fft::fs1536_fwd(v, frc, fw, t);
    else
      fft::fft_slots<M, -1, XBIG, XXCH>(v, fr, twl, t);
  };

  // Hermitian extension of row zi to M points: slot s holds mode k = t+T*s
  // (k <= nx) or the conjugate of mode M-k (k >= M-nx); Im of mode 0 is
  // dropped as c2r does.  Teams of at most one wavefront load only the
  // direct modes (slots s < MO, clamped in-row indices, all loads in flight
  // together) and take the mirrored ones from the partner lane: mode M-k of
  // lane t > 0 sits in slot E-1-s of lane T-t, lane 0 holds it in slot E-s.
  // Wider teams load both.
  constexpr bool SHFL = T <= 64;
  auto load_direct = [&](int c, cx<R>(&r)[MO]) {
    const cx<R> *src = A + ((size_t)(pl * 3 + c) * NZB + zb) * Nx * ZB + zi;
#pragma unroll
    for (int s = 0; s < MO; s++) r[s] = T * s <= nx ? src[(size_t)min(t + T * s, nx) * ZB] : cx<R>{0, 0};
  };
  auto extend = [&](const cx<R>(&r)[MO], cx<R>(&e)[E]) {
    const int lane = (int)(threadIdx.x & 63);
    const int src = lane - t + ((T - t) % T);
#pragma unroll
    for (int s = 0; s < E; s++) {
      const int k = t + T * s;
      cx<R> val = {0, 0};
      if (s < MO) {
        val = r[s < MO ? s : 0];
        if (k > nx) val = cx<R>{0, 0};
        if (k == 0) val.y = 0;
      }
      if (s >= E - MO) {
        const int sm = E - 1 - s;
        const R px = __shfl(r[sm].x, src, 64), py = __shfl(r[sm].y, src, 64);
        cx<R> mv = cx<R>{px, py};
        if (t == 0) mv = (E - s < MO) ? r[(E - s < MO) ? E - s : 0] : cx<R>{0, 0};
        if (k >= M - nx) val = cx<R>{mv.x, -mv.y};
      }
      e[s] = val;
    }
  };
  auto load_raw = [&](int c, cx<R>(&e)[E]) {
    const cx<R> *src = A + ((size_t)(pl * 3 + c) * NZB + zb) * Nx * ZB + zi;
#pragma unroll
    for (int s = 0; s < E; s++) {
      if (T * s > nx && T * s + T - 1 < M - nx) {
        e[s] = cx<R>{0, 0};
      } else {
        const int k = t + T * s;
        const int m = k <= nx ? k : (k >= M - nx ? M - k : 0);
        e[s] = src[(size_t)m * ZB];
      }
    }
  };
  auto fix_ext = [&](cx<R>(&e)[E]) {
#pragma unroll
    for (int s = 0; s < E; s++) {
      const int k = t + T * s;
      cx<R> v = e[s];
      if (k == 0) v.y = 0;
      if (k > nx) {
        if (k >= M - nx)
          v.y = -v.y;
        else
          v = cx<R>{0, 0};
      }
      e[s] = v;
    }
  };

  R u[E], v[E], w[E];
  {
    cx<R> e[E], f[E];
    cx<R> rw[SHFL ? MO : 1];
    if constexpr (SHFL) {
      cx<R> ru[MO], rv[MO];
      load_direct(0, ru);
      load_direct(1, rv);
      load_direct(2, rw);  // in flight during the first transform
      extend(ru, e);
      extend(rv, f);
    } else {
      load_raw(0, e);
      load_raw(1, f);
      fix_ext(e);
      fix_ext(f);
    }
#pragma unroll
    for (int s = 0; s < E; s++) e[s] = cx<R>{e[s].x - f[s].y, e[s].y + f[s].x};  // U + iV
    if constexpr (ILV) {
      extend(rw, f);
      fft::fs768_bwd2<R, MO <= 3>(e, f, frc, fw, t);
    } else {
      if constexpr (!SHFL) load_raw(2, f);  // in flight during the first transform
      bwd(e);
      if constexpr (SHFL)
        extend(rw, f);
      else
        fix_ext(f);
      bwd(f);
    }
#pragma unroll
    for (int s = 0; s < E; s++) {
      u[s] = e[s].x;
      v[s] = e[s].y;
      w[s] = f[s].x;
    }
  }

  // CFL maximum over the interior planes (dnsdata.f90:552-556).  FP64 gives
  // the reference's quotients (div_by: the correctly rounded divisions from
  // three reciprocals per row); the split mode, whose samples are FP32,
  // multiplies by the reciprocals (within an FP64 rounding of the quotients).
  if (cfl_on && iy >= 1 && iy <= g.ny - 1) {
    const double dyv = dy[iy];
    const double rx = 1.0 / dx, ry = 1.0 / dyv, rz = 1.0 / dz;
    double m = 0.0;
    if constexpr (sizeof(R) == 4) {
#pragma unroll
      for (int s = 0; s < E; s++) {
        const double c = fabs((double)u[s]) * rx + fabs((double)v[s]) * ry + fabs((double)w[s]) * rz;
        m = (c > m || c != c) ? c : m;
      }
    } else {
#pragma unroll
      for (int s = 0; s < E; s++) {
        const double c = div_by(fabs((double)u[s]), dx, rx) + div_by(fabs((double)v[s]), dyv, ry) +
                         div_by(fabs((double)w[s]), dz, rz);
        m = (c > m || c != c) ? c : m;
      }
    }
    block_max_atomic<NT>(m, red, cfl_bits);
  }

  if constexpr (PRODUCTS) {
    // pairs (uu,uv), (uw,vv), (vw,ww) share one forward transform each; in
    // this order u is dead once the second pair is formed, and v and w once
    // the third is.  Product indices: 0 uu, 1 vv, 2 ww, 3 uv, 4 vw, 5 uw.
    auto form = [&](const int pp, cx<R>(&e)[E]) {
#pragma unroll
      for (int s = 0; s < E; s++) {
        R a, b;
        if (pp == 0) {
          a = u[s] * u[s];
          b = u[s] * v[s];
        } else if (pp == 1) {
          a = u[s] * w[s];
          b = v[s] * v[s];
        } else {
          a = v[s] * w[s];
          b = w[s] * w[s];
        }
        e[s] = cx<R>{a, b};
      }
    };
    // A_k = (Z_k + conj Z_{M-k})/2, B_k = (Z_k - conj Z_{M-k})/(2i), k=0..nx,
    // with the transform's normalisation factor applied here, once per mode
    auto emit = [&](const int pp, const cx<R>(&e)[E]) {
      const int qa = pp == 0 ? 0 : (pp == 1 ? 5 : 4), qb = pp == 0 ? 3 : (pp == 1 ? 1 : 2);
      cx<R> *da = B + ((size_t)(pl * 6 + qa) * NZB + zb) * Nx * ZB + zi;
      cx<R> *db = B + ((size_t)(pl * 6 + qb) * NZB + zb) * Nx * ZB + zi;
      const R hf = R(0.5) * factor;
      auto put = [&](int k, cx<R> z, cx<R> q) {
        da[(size_t)k * ZB] = cx<R>{(z.x + q.x) * hf, (z.y - q.y) * hf};
        db[(size_t)k * ZB] = cx<R>{(z.y + q.y) * hf, (q.x - z.x) * hf};
      };
      if constexpr (FS2) {
        // wave h forms the modes 256 h .. 256 h + 255 (nx < 512 when M = 1536)
        const int h = t >> 6, tl = t & 63;
#pragma unroll
        for (int s = 0; s < MO; s++) {
          const int k = 256 * h + tl + 64 * s;
          if (k <= nx) {
            cx<R> z, q;
            fft::fs1536_mode(frc, twl, k, z, q);
            put(k, z, q);
          }
        }
        fft::team_sync<T>();  // the rows are the next transform's exchange area
      } else {
        cx<R> mz[MO];
        mirror_pairs<T, E, MO, R>(e, mz, t, nx + 1);
#pragma unroll
        for (int s = 0; s < MO; s++) {
          const int k = t + T * s;
          if (k <= nx) put(k, e[s], mz[s]);
        }
      }
    };
    if constexpr (ILV) {
      // the first pair alone while u, v, w are live, then the other two
      // together
      {
        cx<R> e[E];
        form(0, e);
        fwd(e);
        emit(0, e);
      }
      cx<R> e[E], f[E];
      form(1, e);
      form(2, f);
      fft::fs768_fwd2(e, f, frc, fw, t);
      emit(1, e);
      emit(2, f);
    } else {
#pragma unroll
      for (int pp = 0; pp < 3; pp++) {
        cx<R> e[E];
        form(pp, e);
        fwd(e);
        emit(pp, e);
      }
    }
  }
}

// ------------------------------------------------------------------- K3 --
template <int NZ, typename R>
__global__ __launch_bounds__(256) void k_zfwd(const cx<R> *__restrict__ B, cx<R> *__restrict__ P,
                                              const cx<R> *__restrict__ tw, Geom g, int plane0) {
  using PL = fft::Plan<NZ>;
  constexpr int T = PL::T, E = PL::E, G = 256 / T, SP = stage_stride<NZ>();
  __shared__ __align__(16) R lds[zbuf_reals<NZ, R>()];
  const int team = threadIdx.x / T, t = threadIdx.x % T;
  const int nxg = (g.Nx + G - 1) / G;
  int b = blockIdx.x;
  const int xg = b % nxg;
  b /= nxg;
  const int q = b % 6, pl = b / 6;
  const int ix = xg * G + team;
  const bool act = ix < g.Nx;
  const cx<R> *src = B + (size_t)(pl * 6 + q) * NZ * g.Nx;
  cx<R> e[E];
  if constexpr (z_staged<NZ, R>()) {
    const int ix0 = xg * G, nact = min(G, g.Nx - ix0);
    cx<R> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) {  // all loads in flight together, clamped into the block's runs
      const int el = (int)threadIdx.x + 256 * i;
      const int zb = el / (G * ZB), w = el % (G * ZB);
      v[i] = src[((size_t)zb * g.Nx + ix0) * ZB + min(w, nact * ZB - 1)];
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
      for (int i = 0; i < E; i++) {
        const int el = (int)threadIdx.x + 256 * i;
        const int zb = el / (G * ZB), w = el % (G * ZB), tm = w / ZB, r = w % ZB;
        lds[tm * SP + zb * ZB + r] = tm < nact ? (h ? v[i].y : v[i].x) : R(0);
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < E; s++) (h ? e[s].y : e[s].x) = lds[team * SP + t + T * s];
      __syncthreads();  // the next round and the transforms reuse the staging area
    }
  } else {
#pragma unroll
    for (int s = 0; s < E; s++) {
      const int z = t + T * s;
      e[s] = act ? src[((size_t)(z / ZB) * g.Nx + ix) * ZB + (z % ZB)] : cx<R>{0, 0};
    }
  }
  fft::fft_slots<NZ, -1, true, ZXCH>(e, lds + team * fft::lds_reals<NZ, ZXCH>(), tw, t);
  if (act) {
    cx<R> *dst = P + ((size_t)q * g.Ny + (plane0 + pl)) * g.ps + (size_t)ix * g.Nz;
#pragma unroll
    for (int s = 0; s < E; s++) {
      const int k = t + T * s;
      const int iz = k <= g.nz ? k + g.nz : (k >= NZ - g.nz ? k - NZ + g.nz : -1);
      if (iz >= 0) dst[iz] = e[s];
    }
  }
}

// ------------------------------------------------------------- selftest --
template <int N, int S>
__global__ __launch_bounds__(256) void k_fft_test(cx<double> *__restrict__ data, const cx<double> *__restrict__ tw,
                                                  int batch) {
  using PL = fft::Plan<N>;
  constexpr int T = PL::T, E = PL::E, G = 256 / T;
  __shared__ double lds[G * N];
  const int team = threadIdx.x / T, t = threadIdx.x % T;
  const int row = blockIdx.x * G + team;
  const bool act = row < batch;
  cx<double> *x = data + (size_t)(act ? row : 0) * N;
  cx<double> e[E];
#pragma unroll
  for (int s = 0; s < E; s++) e[s] = act ? x[t + T * s] : cx<double>{0, 0};
  fft::fft_slots<N, S>(e, lds + team * N, tw, t);
  if (act)
#pragma unroll
    for (int s = 0; s < E; s++) x[t + T * s] = e[s];
}

// ------------------------------------------------------ four-step test --
// Rows of 768 complex FP64 values in natural order.  sign +1: slot layout ->
// fs768_bwd -> written back in natural order through fs768_index; sign -1:
// read in layout P -> fs768_fwd -> natural order.  inner > 1 first runs
// inner-1 backward/forward pairs (for timing; the data grows by 768 per pair).
__global__ __launch_bounds__(256) void k_fs768_test(const cx<double> *__restrict__ in, cx<double> *__restrict__ out,
                                                    const cx<double> *__restrict__ tw, int batch, int sign, int inner) {
  __shared__ __align__(16) cx<double> lds[4 * fft::FS_LDS];
  __shared__ cx<double> twl[fft::FS_TW];
  for (int i = threadIdx.x; i < fft::FS_TW; i += 256) twl[i] = tw[i];
  __syncthreads();
  const int team = threadIdx.x / 64, t = threadIdx.x % 64;
  const int row = blockIdx.x * 4 + team;
  const bool act = row < batch;
  const cx<double> *x = in + (size_t)(act ? row : 0) * 768;
  cx<double> *L = lds + team * fft::FS_LDS;
  const fft::FsTw<double> fw = fft::fs_twiddles(twl, t);
  cx<double> e[12];
  if (sign > 0) {
#pragma unroll
    for (int s = 0; s < 12; s++) e[s] = x[t + 64 * s];
    for (int k = 1; k < inner; k++) {
      fft::fs768_bwd(e, L, fw, t);
      fft::fs768_fwd(e, L, fw, t);
    }
    fft::fs768_bwd(e, L, fw, t);
    if (act)
#pragma unroll
      for (int s = 0; s < 12; s++) out[(size_t)row * 768 + fft::fs768_index(t, s)] = e[s];
  } else {
#pragma unroll
    for (int s = 0; s < 12; s++) e[s] = x[fft::fs768_index(t, s)];
    for (int k = 1; k < inner; k++) {
      fft::fs768_fwd(e, L, fw, t);
      fft::fs768_bwd(e, L, fw, t);
    }
    fft::fs768_fwd(e, L, fw, t);
    if (act)
#pragma unroll
      for (int s = 0; s < 12; s++) out[(size_t)row * 768 + t + 64 * s] = e[s];
  }
}

hipError_t launch_fs768_test(const cx<double> *in, cx<double> *out, const cx<double> *tw, int batch, int sign, int inner) {
  hipLaunchKernelGGL(k_fs768_test, dim3((unsigned)((batch + 3) / 4)), dim3(256), 0, 0, in, out, tw, batch, sign, inner);
  return hipGetLastError();
}

// ------------------------------------------------------------- benchmark --
// out-of-place batched transforms for timing the exchange/plan variants:
// variants 0-7: bit 0 complex exchange, bit 1 padded exchange, bit 2 radix-4/3
// plan; 8-11: complex exchange padded every 16, 8, 4 elements, XOR layout
template <int N, int S, bool BIG, int XCH>
__global__ __launch_bounds__(256) void k_fft_bench(const cx<double> *__restrict__ in, cx<double> *__restrict__ out,
                                                   const cx<double> *__restrict__ tw, int batch, int inner) {
  using PL = fft::Plan<N, BIG>;
  constexpr int T = PL::T, E = PL::E, G = 256 / T, L = fft::lds_reals<N, XCH>();
  __shared__ __align__(16) double lds[G * L];
  const int team = threadIdx.x / T, t = threadIdx.x % T;
  const int row = blockIdx.x * G + team;
  const bool act = row < batch;
  const cx<double> *x = in + (size_t)(act ? row : 0) * N;
  cx<double> e[E];
#pragma unroll
  for (int s = 0; s < E; s++) e[s] = act ? x[t + T * s] : cx<double>{0, 0};
  for (int k = 0; k < inner; k++) fft::fft_slots<N, S, BIG, XCH>(e, lds + team * L, tw, t);
  if (act)
#pragma unroll
    for (int s = 0; s < E; s++) out[(size_t)row * N + t + T * s] = e[s];
}

template <int N, int S, bool BIG, int XCH>
hipError_t launch_bench_v(const cx<double> *in, cx<double> *out, const cx<double> *tw, int batch, int inner) {
  constexpr int G = 256 / fft::Plan<N, BIG>::T;
  hipLaunchKernelGGL((k_fft_bench<N, S, BIG, XCH>), dim3((unsigned)((batch + G - 1) / G)), dim3(256), 0, 0, in, out, tw,
                     batch, inner);
  return hipGetLastError();
}

template <int N, int S>
hipError_t launch_bench_n(int variant, const cx<double> *in, cx<double> *out, const cx<double> *tw, int batch, int inner) {
  switch (variant) {
    case 0: return launch_bench_v<N, S, true, 0>(in, out, tw, batch, inner);
    case 1: return launch_bench_v<N, S, true, 1>(in, out, tw, batch, inner);
    case 2: return launch_bench_v<N, S, true, 2>(in, out, tw, batch, inner);
    case 3: return launch_bench_v<N, S, true, 3>(in, out, tw, batch, inner);
    case 4: return launch_bench_v<N, S, false, 0>(in, out, tw, batch, inner);
    case 5: return launch_bench_v<N, S, false, 1>(in, out, tw, batch, inner);
    case 6: return launch_bench_v<N, S, false, 2>(in, out, tw, batch, inner);
    case 7: return launch_bench_v<N, S, false, 3>(in, out, tw, batch, inner);
    // complex exchanges with the timing layouts
    case 8: return launch_bench_v<N, S, true, fft::XCH_CPLX | fft::XCH_PAD16>(in, out, tw, batch, inner);
    case 9: return launch_bench_v<N, S, true, fft::XCH_CPLX | fft::XCH_PAD8>(in, out, tw, batch, inner);
    case 10: return launch_bench_v<N, S, true, fft::XCH_CPLX | fft::XCH_PAD4>(in, out, tw, batch, inner);
    case 11: return launch_bench_v<N, S, true, fft::XCH_CPLX | fft::XCH_XOR>(in, out, tw, batch, inner);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fft_bench(int N, int sign, int variant, const cx<double> *in, cx<double> *out, const cx<double> *tw,
                            int batch, int inner) {
  switch (N) {
    case 384: return sign < 0 ? launch_bench_n<384, -1>(variant, in, out, tw, batch, inner) : launch_bench_n<384, 1>(variant, in, out, tw, batch, inner);
    case 768: return sign < 0 ? launch_bench_n<768, -1>(variant, in, out, tw, batch, inner) : launch_bench_n<768, 1>(variant, in, out, tw, batch, inner);
    case 1536: return sign < 0 ? launch_bench_n<1536, -1>(variant, in, out, tw, batch, inner) : launch_bench_n<1536, 1>(variant, in, out, tw, batch, inner);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------ launchers --
#define CHAN_FOR_SIZES(X) X(12) X(24) X(48) X(96) X(192) X(384) X(768) X(1536) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

template <typename R>
hipError_t launch_zinv(int NZ, hipStream_t st, const cx<double> *V, cx<R> *A, const cx<R> *tw, const Geom &g, int plane0,
                       int nplanes) {
  switch (NZ) {
#define X(N)                                                                              \
  case N: {                                                                               \
    constexpr int G = 256 / fft::Plan<N>::T;                                              \
    dim3 grid((unsigned)(((g.Nx + G - 1) / G) * 3 * nplanes));                            \
    hipLaunchKernelGGL((k_zinv<N, R>), grid, dim3(256), 0, st, V, A, tw, g, plane0);      \
    return hipGetLastError();                                                             \
  }
    CHAN_FOR_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_xpass(int M, hipStream_t st, const cx<R> *A, cx<R> *B, const cx<R> *tw, const Geom &g, int plane0,
                        int nplanes, bool products, int cfl_on, const double *dy, double dx, double dz, R factor,
                        unsigned long long *cfl_bits) {
  const unsigned ntile = (unsigned)((g.nzd / ZB) * nplanes);
  switch (M) {
#define X(N)                                                                                                          \
  case N: {                                                                                                           \
    dim3 blk(xteams<N, R>() * fft::Plan<N>::T);                                                                       \
    dim3 grid(xteams<N, R>() < ZB ? 16 * ((ntile + 7) / 8) : ntile);                                                  \
    if (products && N == 768 && g.nx + 1 == 192)                                                                      \
      hipLaunchKernelGGL((k_xpass<N, R, true, (N == 768 ? 3 : xmodes<fft::Plan<N>::E>()), true>), grid, blk, 0, st, A, \
                         B, tw, g, plane0, cfl_on, dy, dx, dz, factor, cfl_bits, nplanes);                            \
    else if (products && N == 768 && g.nx + 1 <= 192)                                                                 \
      hipLaunchKernelGGL((k_xpass<N, R, true, (N == 768 ? 3 : xmodes<fft::Plan<N>::E>())>), grid, blk, 0, st, A, B, tw, \
                         g, plane0, cfl_on, dy, dx, dz, factor, cfl_bits, nplanes);                                   \
    else if (products)                                                                                                \
      hipLaunchKernelGGL((k_xpass<N, R, true>), grid, blk, 0, st, A, B, tw, g, plane0, cfl_on, dy, dx, dz, factor,    \
                         cfl_bits, nplanes);                                                                          \
    else                                                                                                              \
      hipLaunchKernelGGL((k_xpass<N, R, false>), grid, blk, 0, st, A, B, tw, g, plane0, cfl_on, dy, dx, dz, factor,   \
                         cfl_bits, nplanes);                                                                          \
    return hipGetLastError();                                                                                         \
  }
    CHAN_FOR_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_zfwd(int NZ, hipStream_t st, const cx<R> *B, cx<R> *P, const cx<R> *tw, const Geom &g, int plane0,
                       int nplanes) {
  switch (NZ) {
#define X(N)                                                                         \
  case N: {                                                                          \
    constexpr int G = 256 / fft::Plan<N>::T;                                         \
    dim3 grid((unsigned)(((g.Nx + G - 1) / G) * 6 * nplanes));                       \
    hipLaunchKernelGGL((k_zfwd<N, R>), grid, dim3(256), 0, st, B, P, tw, g, plane0); \
    return hipGetLastError();                                                        \
  }
    CHAN_FOR_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fft_test(int N, int sign, hipStream_t st, cx<double> *data, const cx<double> *tw, int batch) {
  switch (N) {
#define X(NN)                                                                                      \
  case NN: {                                                                                       \
    constexpr int G = 256 / fft::Plan<NN>::T;                                                      \
    dim3 grid((unsigned)((batch + G - 1) / G));                                                    \
    if (sign < 0)                                                                                  \
      hipLaunchKernelGGL((k_fft_test<NN, -1>), grid, dim3(256), 0, st, data, tw, batch);           \
    else                                                                                           \
      hipLaunchKernelGGL((k_fft_test<NN, +1>), grid, dim3(256), 0, st, data, tw, batch);           \
    return hipGetLastError();                                                                      \
  }
    CHAN_FOR_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

int fft_size_supported(int N) {
  switch (N) {
#define X(NN) case NN:
    CHAN_FOR_SIZES(X)
#undef X
    return 1;
  }
  return 0;
}

template hipError_t launch_zinv<double>(int, hipStream_t, const cx<double> *, cx<double> *, const cx<double> *,
                                        const Geom &, int, int);
template hipError_t launch_xpass<double>(int, hipStream_t, const cx<double> *, cx<double> *, const cx<double> *,
                                         const Geom &, int, int, bool, int, const double *, double, double, double,
                                         unsigned long long *);
template hipError_t launch_zfwd<double>(int, hipStream_t, const cx<double> *, cx<double> *, const cx<double> *,
                                        const Geom &, int, int);
template hipError_t launch_zinv<float>(int, hipStream_t, const cx<double> *, cx<float> *, const cx<float> *,
                                       const Geom &, int, int);
template hipError_t launch_xpass<float>(int, hipStream_t, const cx<float> *, cx<float> *, const cx<float> *,
                                        const Geom &, int, int, bool, int, const double *, double, double, float,
                                        unsigned long long *);
template hipError_t launch_zfwd<float>(int, hipStream_t, const cx<float> *, cx<float> *, const cx<float> *,
                                       const Geom &, int, int);

}  // namespace chan
