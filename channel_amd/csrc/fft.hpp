// fft.hpp -- batched mixed-radix (2,3,4) Stockham FFT for CDNA4, one team of
// T threads per transform, data resident in registers between radix passes
// and exchanged through LDS.
//
// Slot convention: a team of T threads transforms N = E*T points; thread t
// holds "slot" s (0 <= s < E) = element t + T*s both on input and on output,
// so loads from and stores to global memory by consecutive threads are
// contiguous.  Sizes N = 3*2^k use E = 12, N = 2^k use E = 16; every radix
// divides E, so each pass maps whole butterflies onto each thread.
//
// Conventions match FFTW's (ffts.f90:70-75): sign -1 forward
// X_k = sum_j x_j e^{-2 pi i jk/N}, sign +1 backward, both unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace chan {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) {
  return cx<R>{a, b};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> add(cx<R> a, cx<R> b) {
  return cx<R>{a.x + b.x, a.y + b.y};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> sub(cx<R> a, cx<R> b) {
  return cx<R>{a.x - b.x, a.y - b.y};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> mul(cx<R> a, cx<R> b) {
  return cx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> scl(R s, cx<R> a) {
  return cx<R>{s * a.x, s * a.y};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> a) {
  return cx<R>{a.x, -a.y};
}

namespace fft {

// Radix plan of an N-point transform with E elements per thread.  BIG: a
// first pass of radix E (12 or 16, done in registers as 4x3 or 4x4), then
// radix-4 passes and at most one radix-2 pass.  Otherwise radix-4 passes, at
// most one radix-2 pass and a final radix-3 pass (fewer live registers, one
// more LDS exchange).  The plan is a set of constexpr functions so that the
// host builds the twiddle tables from the same definition.
constexpr bool plan_pow2(int N) { return (N & (N - 1)) == 0; }
constexpr int plan_E(int N) { return plan_pow2(N) ? 16 : 12; }
constexpr int plan_rest(int N, bool big) { return big ? N / plan_E(N) : (plan_pow2(N) ? N : N / 3); }
constexpr int plan_n4(int N, bool big) {
  int r = plan_rest(N, big), n = 0;
  while (r % 4 == 0) {
    r /= 4;
    n++;
  }
  return n;
}
constexpr bool plan_has2(int N, bool big) {
  int r = plan_rest(N, big);
  while (r % 4 == 0) r /= 4;
  return r == 2;
}
constexpr int plan_npass(int N, bool big) {
  return (big ? 1 : (plan_pow2(N) ? 0 : 1)) + plan_n4(N, big) + (plan_has2(N, big) ? 1 : 0);
}
constexpr int plan_radix(int N, bool big, int p) {
  if (big) return p == 0 ? plan_E(N) : (p <= plan_n4(N, big) ? 4 : 2);
  return p < plan_n4(N, big) ? 4 : ((p == plan_n4(N, big) && plan_has2(N, big)) ? 2 : 3);
}

// Twiddle tables: pass p >= 1 (stride NS = product of the earlier radices,
// radix RAD) reads w_N^(k*N/(NS*RAD)) for k = j mod NS, stored contiguously
// per pass so that consecutive lanes read consecutive entries (no LDS bank
// conflicts when the table sits in LDS).  The tables of the BIG plan come
// first, those of the other plan follow.
constexpr int plan_tw_offset(int N, bool big, int P) {
  int off = 0, ns = 1;
  for (int q = 0; q < P; q++) {
    if (q >= 1) off += ns;
    ns *= plan_radix(N, big, q);
  }
  return off;
}
constexpr int plan_tw_size(int N, bool big) { return plan_tw_offset(N, big, plan_npass(N, big)); }
constexpr int tw_table_size(int N) { return plan_tw_size(N, true) + plan_tw_size(N, false); }

template <int N, bool BIG = true>
struct Plan {
  static constexpr bool pow2 = plan_pow2(N);
  static constexpr int E = plan_E(N);
  static constexpr int T = N / E;
  static_assert(T >= 1 && T * E == N, "unsupported FFT length");
  static constexpr int npass() { return plan_npass(N, BIG); }
  static constexpr int radix(int p) { return plan_radix(N, BIG, p); }
  static constexpr int tw_base(int p) { return (BIG ? 0 : plan_tw_size(N, true)) + plan_tw_offset(N, BIG, p); }
};

// Host: fills w[0 .. tw_table_size(N)) with exp(-2 pi i e/N) in the layout
// above, from long-double angles.
template <typename T>
inline void build_twiddles(cx<T> *w, int N) {
  int o = 0;
  for (int big = 1; big >= 0; big--) {
    int ns = 1;
    for (int p = 0; p < plan_npass(N, big); p++) {
      const int rad = plan_radix(N, big, p);
      if (p >= 1) {
        const int tws = N / (ns * rad);
        for (int k = 0; k < ns; k++) {
          const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)(k * tws) / (long double)N;
          w[o++] = cx<T>{(T)cosl(ang), (T)sinl(ang)};
        }
      }
      ns *= rad;
    }
  }
}

// team-level LDS synchronisation: a team within one wavefront only needs the
// wave's LDS operations ordered; larger teams use the workgroup barrier.
template <int T>
__device__ __forceinline__ void team_sync() {
  if constexpr (T <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// 12 = 4 x 3: n = 3 n1 + n2, q = q1 + 4 q2;
// y_q = sum_n2 w3^(n2 q2) w12^(n2 q1) sum_n1 w4^(n1 q1) a_(3 n1 + n2).
// MID0: inputs 3..8 are known to be zero (the backward transforms of the x
// pass at nx + 1 <= 192 hold no modes there), so each DFT-4 has inputs
// (x, 0, 0, y) only; the compiler cannot drop additions of zero itself,
// which would change the sign of a zero result.
template <int RAD, int S, typename R>
__device__ __forceinline__ void dft(cx<R> *a);

template <int S, bool MID0, typename R>
__device__ __forceinline__ void dft12(cx<R> *a) {
  const R sg = S < 0 ? R(-1) : R(1);
  const R h = R(0.5), c = R(0.86602540378443864676372317075293618);
  cx<R> b[3][4];
#pragma unroll
  for (int n2 = 0; n2 < 3; n2++) {
    if constexpr (MID0) {
      const cx<R> x = a[n2], y = a[9 + n2];
      b[n2][0] = add(x, y);
      b[n2][2] = sub(x, y);
      if constexpr (S < 0) {  // y1 = x + i y, y3 = x - i y
        b[n2][1] = cx<R>{x.x - y.y, x.y + y.x};
        b[n2][3] = cx<R>{x.x + y.y, x.y - y.x};
      } else {  // y1 = x - i y, y3 = x + i y
        b[n2][1] = cx<R>{x.x + y.y, x.y - y.x};
        b[n2][3] = cx<R>{x.x - y.y, x.y + y.x};
      }
    } else {
      cx<R> q[4] = {a[n2], a[3 + n2], a[6 + n2], a[9 + n2]};
      dft<4, S>(q);
#pragma unroll
      for (int q1 = 0; q1 < 4; q1++) b[n2][q1] = q[q1];
    }
  }
  // twiddles w12^(n2 q1): (n2,q1) = (1,1) (1,2) (1,3) (2,1) (2,2) (2,3)
  b[1][1] = mul(b[1][1], cx<R>{c, sg * h});
  b[1][2] = mul(b[1][2], cx<R>{h, sg * c});
  b[1][3] = cx<R>{-sg * b[1][3].y, sg * b[1][3].x};  // * S i
  b[2][1] = mul(b[2][1], cx<R>{h, sg * c});
  b[2][2] = mul(b[2][2], cx<R>{-h, sg * c});
  b[2][3] = cx<R>{-b[2][3].x, -b[2][3].y};  // * -1
#pragma unroll
  for (int q1 = 0; q1 < 4; q1++) {
    cx<R> t[3] = {b[0][q1], b[1][q1], b[2][q1]};
    dft<3, S>(t);
#pragma unroll
    for (int q2 = 0; q2 < 3; q2++) a[q1 + 4 * q2] = t[q2];
  }
}

template <int RAD, int S, typename R>
__device__ __forceinline__ void dft(cx<R> *a) {
  if constexpr (RAD == 2) {
    cx<R> t = sub(a[0], a[1]);
    a[0] = add(a[0], a[1]);
    a[1] = t;
  } else if constexpr (RAD == 4) {
    cx<R> t0 = add(a[0], a[2]), t1 = sub(a[0], a[2]);
    cx<R> t2 = add(a[1], a[3]), t3 = sub(a[1], a[3]);
    a[0] = add(t0, t2);
    a[2] = sub(t0, t2);
    if constexpr (S < 0) {  // y1 = t1 - i t3, y3 = t1 + i t3
      a[1] = cx<R>{t1.x + t3.y, t1.y - t3.x};
      a[3] = cx<R>{t1.x - t3.y, t1.y + t3.x};
    } else {  // y1 = t1 + i t3, y3 = t1 - i t3
      a[1] = cx<R>{t1.x - t3.y, t1.y + t3.x};
      a[3] = cx<R>{t1.x + t3.y, t1.y - t3.x};
    }
  } else if constexpr (RAD == 3) {
    const R s = (S < 0 ? R(-1) : R(1)) * R(0.86602540378443864676372317075293618);
    cx<R> t1 = add(a[1], a[2]);
    cx<R> d = sub(a[1], a[2]);
    cx<R> t2 = cx<R>{a[0].x - R(0.5) * t1.x, a[0].y - R(0.5) * t1.y};
    a[0] = add(a[0], t1);
    a[1] = cx<R>{t2.x - s * d.y, t2.y + s * d.x};
    a[2] = cx<R>{t2.x + s * d.y, t2.y - s * d.x};
  } else if constexpr (RAD == 12) {
    dft12<S, false>(a);
  } else {
    static_assert(RAD == 16, "radix");
    // 16 = 4 x 4: n = 4 n1 + n2, q = q1 + 4 q2
    const R sg = S < 0 ? R(-1) : R(1);
    const R c1 = R(0.92387953251128675612818318939678829), s1 = R(0.38268343236508977172845998403039887);
    const R r2 = R(0.70710678118654752440084436210484904);
    cx<R> b[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
      cx<R> q[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
      dft<4, S>(q);
#pragma unroll
      for (int q1 = 0; q1 < 4; q1++) b[n2][q1] = q[q1];
    }
    // twiddles w16^(n2 q1) with w16 = exp(S 2 pi i/16)
    b[1][1] = mul(b[1][1], cx<R>{c1, sg * s1});
    b[1][2] = mul(b[1][2], cx<R>{r2, sg * r2});
    b[1][3] = mul(b[1][3], cx<R>{s1, sg * c1});
    b[2][1] = mul(b[2][1], cx<R>{r2, sg * r2});
    b[2][2] = cx<R>{-sg * b[2][2].y, sg * b[2][2].x};  // * S i
    b[2][3] = mul(b[2][3], cx<R>{-r2, sg * r2});
    b[3][1] = mul(b[3][1], cx<R>{s1, sg * c1});
    b[3][2] = mul(b[3][2], cx<R>{-r2, sg * r2});
    b[3][3] = mul(b[3][3], cx<R>{-c1, -sg * s1});
#pragma unroll
    for (int q1 = 0; q1 < 4; q1++) {
      cx<R> t[4] = {b[0][q1], b[1][q1], b[2][q1], b[3][q1]};
      dft<4, S>(t);
#pragma unroll
      for (int q2 = 0; q2 < 4; q2++) a[q1 + 4 * q2] = t[q2];
    }
  }
}

// LDS exchange layouts between passes.  XCH_CPLX moves whole complex
// values per access (two reals per lane, half the LDS instructions; needs 2N
// reals per team), otherwise real and imaginary parts travel in two rounds
// (N reals per team).  XCH_PAD inserts one element after every 32, which
// breaks the power-of-two strides of the Stockham scatter across LDS banks.
// The other layouts (padding after every 16, 8 or 4 elements, or an XOR of
// the low three index bits with the next three) exist for timing.
constexpr int XCH_CPLX = 1, XCH_PAD = 2, XCH_PAD16 = 8, XCH_PAD8 = 16, XCH_PAD4 = 32, XCH_XOR = 64;

template <int XCH>
__device__ __forceinline__ int xidx(int i) {
  if constexpr (XCH & XCH_PAD) return i + (i >> 5);
  if constexpr (XCH & XCH_PAD16) return i + (i >> 4);
  if constexpr (XCH & XCH_PAD8) return i + (i >> 3);
  if constexpr (XCH & XCH_PAD4) return i + (i >> 2);
  if constexpr (XCH & XCH_XOR) return i ^ ((i >> 3) & 7);
  return i;
}

template <int N, int XCH>
constexpr int xch_len() {
  return (XCH & XCH_PAD) ? N + N / 32 : (XCH & XCH_PAD16) ? N + N / 16 : (XCH & XCH_PAD8) ? N + N / 8 : (XCH & XCH_PAD4) ? N + N / 4 : N;
}

// LDS reals a team needs for an N-point transform with exchange layout XCH
template <int N, int XCH>
constexpr int lds_reals() {
  return ((XCH & XCH_CPLX) ? 2 : 1) * xch_len<N, XCH>();
}

// One Stockham pass (decimation in time, autosort) on the thread's E values
// held in "read order" v[m*RAD + r] = in[j + r*N/RAD], j = t + m*T; then the
// exchange through LDS into the read order of the next pass.
template <int N, int S, int P, int NS, bool BIG, int XCH, typename R>
__device__ __forceinline__ void run_pass(cx<R> (&v)[Plan<N>::E], R *lds, const cx<R> *__restrict__ tw, int t) {
  using PL = Plan<N, BIG>;
  constexpr int E = PL::E, T = PL::T, RAD = PL::radix(P), NB = E / RAD;
#pragma unroll
  for (int m = 0; m < NB; m++) {
    if constexpr (NS > 1) {
      // one table load per butterfly; higher powers by multiplication
      const int j = t + m * T;
      const int k = j % NS;
      cx<R> w1 = tw[PL::tw_base(P) + k];
      if constexpr (S > 0) w1.y = -w1.y;
      cx<R> wr = w1;
#pragma unroll
      for (int r = 1; r < RAD; r++) {
        v[m * RAD + r] = mul(v[m * RAD + r], wr);
        if (r + 1 < RAD) wr = mul(wr, w1);
      }
    }
    dft<RAD, S>(&v[m * RAD]);
  }
  if constexpr (P + 1 < PL::npass()) {
    constexpr int R2 = PL::radix(P + 1), NB2 = E / R2;
    if constexpr (XCH & XCH_CPLX) {
      cx<R> *c = reinterpret_cast<cx<R> *>(lds);
#pragma unroll
      for (int m = 0; m < NB; m++) {
        const int j = t + m * T;
        const int base = (j / NS) * NS * RAD + j % NS;
#pragma unroll
        for (int q = 0; q < RAD; q++) c[xidx<XCH>(base + q * NS)] = v[m * RAD + q];
      }
      team_sync<T>();
#pragma unroll
      for (int m = 0; m < NB2; m++)
#pragma unroll
        for (int r = 0; r < R2; r++) v[m * R2 + r] = c[xidx<XCH>(t + m * T + r * (N / R2))];
      team_sync<T>();
    } else {
#pragma unroll
      for (int half = 0; half < 2; half++) {
#pragma unroll
        for (int m = 0; m < NB; m++) {
          const int j = t + m * T;
          const int base = (j / NS) * NS * RAD + j % NS;
#pragma unroll
          for (int q = 0; q < RAD; q++) lds[xidx<XCH>(base + q * NS)] = half ? v[m * RAD + q].y : v[m * RAD + q].x;
        }
        team_sync<T>();
#pragma unroll
        for (int m = 0; m < NB2; m++)
#pragma unroll
          for (int r = 0; r < R2; r++) {
            const R val = lds[xidx<XCH>(t + m * T + r * (N / R2))];
            if (half)
              v[m * R2 + r].y = val;
            else
              v[m * R2 + r].x = val;
          }
        team_sync<T>();
      }
    }
    run_pass<N, S, P + 1, NS * RAD, BIG, XCH, R>(v, lds, tw, t);
  }
}

// Transform the team's N points held in slots e[s] = x[t + T*s]; on return
// e[s] = X[t + T*s].  lds: lds_reals<N, XCH>() reals private to the team.
// tw: the twiddle table of build_twiddles (tw_table_size(N) entries).
template <int N, int S, bool BIG = true, int XCH = 0, typename R>
__device__ __forceinline__ void fft_slots(cx<R> (&e)[Plan<N>::E], R *lds, const cx<R> *__restrict__ tw, int t) {
  using PL = Plan<N, BIG>;
  constexpr int E = PL::E, R1 = PL::radix(0), RL = PL::radix(PL::npass() - 1);
  cx<R> v[E];
#pragma unroll
  for (int m = 0; m < E / R1; m++)
#pragma unroll
    for (int r = 0; r < R1; r++) v[m * R1 + r] = e[m + r * (E / R1)];
  run_pass<N, S, 0, 1, BIG, XCH, R>(v, lds, tw, t);
#pragma unroll
  for (int m = 0; m < E / RL; m++)
#pragma unroll
    for (int q = 0; q < RL; q++) e[m + q * (E / RL)] = v[m * RL + q];
}

// ------------------------------------------------------ four-step 768 --
// A 768-point transform held by one wavefront, 12 values per lane, computed
// as 12 x 64: a 12-point DFT in registers, twiddles w768^(n1 t), and a
// 64-point DFT across the lanes split 4 x 4 x 4 with lane t = p + 16 g,
// p = d + 4 c.  The radix-4 over the four 16-lane rows g runs after a 4x4
// transpose of register quads by the permlane swaps (no LDS); the radix-4
// passes over c and over d each take one LDS exchange, in layouts without
// bank conflicts for b128 accesses (rows of FS_RG complex values per 16-lane
// row; c*52 + xi*4 + d, then c*12 + d*49 + eta).  Two exchanges per
// transform instead of the three of the Stockham plan.
//
// fs768_bwd (sign +1) takes the slot layout (slot s of lane t = element
// t + 64 s) and leaves the samples in the layout P of fs768_index; fs768_fwd
// (sign -1) takes layout P and returns the slot layout.  Pointwise work
// between the two (the products) needs no reordering.
//
// tw: the FS_TW-entry table of build_fs768_twiddles, forward sign:
// exp(-2 pi i t/768) for t < 64, exp(-2 pi i p/64) for p < 16,
// exp(-2 pi i d/16) for d < 4.
constexpr int FS_RG = 208, FS_LDS = 4 * FS_RG, FS_TW = 84;

template <typename T>
inline void build_fs768_twiddles(cx<T> *w) {
  const long double pi2 = 2.0L * 3.141592653589793238462643383279502884L;
  for (int t = 0; t < 64; t++) w[t] = cx<T>{(T)cosl(-pi2 * t / 768), (T)sinl(-pi2 * t / 768)};
  for (int q = 0; q < 16; q++) w[64 + q] = cx<T>{(T)cosl(-pi2 * q / 64), (T)sinl(-pi2 * q / 64)};
  for (int q = 0; q < 4; q++) w[80 + q] = cx<T>{(T)cosl(-pi2 * q / 16), (T)sinl(-pi2 * q / 16)};
}

// element index n held by lane t in slot 4j+f after fs768_bwd
__host__ __device__ __forceinline__ int fs768_index(int t, int slot) {
  const int p = t & 15, g = t >> 4, d = p & 3, c = p >> 2, j = slot >> 2, f = slot & 3;
  const int eta = 3 * d + j, i = eta >> 2, e = eta & 3;
  const int xi = 3 * c + i, q = xi >> 2, a = xi & 3;
  return (4 * q + g) + 12 * (a + 4 * (e + 4 * f));
}

template <typename R>
__device__ __forceinline__ void fs_swap(cx<R> &a, cx<R> &b, bool rows32) {
  constexpr int ND = (int)(sizeof(cx<R>) / 4);
  unsigned ua[ND], ub[ND];
  __builtin_memcpy(ua, &a, sizeof a);
  __builtin_memcpy(ub, &b, sizeof b);
#pragma unroll
  for (int k = 0; k < ND; k++) {
    if (rows32) {
      const auto r = __builtin_amdgcn_permlane32_swap(ua[k], ub[k], false, false);
      ua[k] = r[0];
      ub[k] = r[1];
    } else {
      const auto r = __builtin_amdgcn_permlane16_swap(ua[k], ub[k], false, false);
      ua[k] = r[0];
      ub[k] = r[1];
    }
  }
  __builtin_memcpy(&a, ua, sizeof a);
  __builtin_memcpy(&b, ub, sizeof b);
}

// transpose of each register quad {4q..4q+3} with the four 16-lane rows:
// afterwards register 4q+r of row g holds what register 4q+g of row r held
// (an involution)
template <typename R>
__device__ __forceinline__ void fs_xpose(cx<R> (&e)[12]) {
#pragma unroll
  for (int q = 0; q < 3; q++) {
    fs_swap(e[4 * q], e[4 * q + 2], true);
    fs_swap(e[4 * q + 1], e[4 * q + 3], true);
    fs_swap(e[4 * q], e[4 * q + 1], false);
    fs_swap(e[4 * q + 2], e[4 * q + 3], false);
  }
}

// a * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> mulc(cx<R> a, cx<R> b) {
  return cx<R>{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

// The twiddle powers of one lane, forward sign, shared by every transform of
// the lane's rows: w768^(n1 t) for n1 = 1..11, w64^(a p) and w16^(e d) for
// a, e = 1..3 (powers by a shallow product tree from the table entries).
template <typename R>
struct FsTw {
  cx<R> w768[11], w64[3], w16[3];
};

template <typename R>
__device__ __forceinline__ FsTw<R> fs_twiddles(const cx<R> *tw, int t) {
  FsTw<R> z;
  const cx<R> w = tw[t];
  const cx<R> w2 = mul(w, w), w3 = mul(w2, w), w4 = mul(w2, w2), w6 = mul(w3, w3), w8 = mul(w4, w4);
  z.w768[0] = w;
  z.w768[1] = w2;
  z.w768[2] = w3;
  z.w768[3] = w4;
  z.w768[4] = mul(w4, w);
  z.w768[5] = w6;
  z.w768[6] = mul(w6, w);
  z.w768[7] = w8;
  z.w768[8] = mul(w8, w);
  z.w768[9] = mul(w6, w4);
  z.w768[10] = mul(w8, w3);
  const cx<R> u = tw[64 + (t & 15)], u2 = mul(u, u);
  z.w64[0] = u;
  z.w64[1] = u2;
  z.w64[2] = mul(u2, u);
  const cx<R> v = tw[80 + (t & 3)], v2 = mul(v, v);
  z.w16[0] = v;
  z.w16[1] = v2;
  z.w16[2] = mul(v2, v);
  return z;
}

// v[k] *= w[k-1] (forward) or conj(w[k-1]) (backward), k = 1..n-1
template <int S, int n, typename R>
__device__ __forceinline__ void fs_apply(cx<R> *v, const cx<R> *w) {
#pragma unroll
  for (int k = 1; k < n; k++) v[k] = S < 0 ? mul(v[k], w[k - 1]) : mulc(v[k], w[k - 1]);
}

// MID0: slots 3..8 hold zeros (see dft12)
template <typename R, bool MID0 = false>
__device__ __forceinline__ void fs768_bwd(cx<R> (&e)[12], cx<R> *lds, const FsTw<R> &w, int t) {
  const int p = t & 15, g = t >> 4, d = p & 3, c = p >> 2;
  cx<R> *row = lds + g * FS_RG;
  dft12<+1, MID0>(e);
  fs_apply<+1, 12>(e, w.w768);
  fs_xpose(e);
#pragma unroll
  for (int q = 0; q < 3; q++) {
    dft<4, +1>(&e[4 * q]);
    fs_apply<+1, 4>(&e[4 * q], w.w64);
  }
#pragma unroll
  for (int xi = 0; xi < 12; xi++) row[c * 52 + xi * 4 + d] = e[xi];
  team_sync<64>();
  cx<R> v[12];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int cp = 0; cp < 4; cp++) v[4 * i + cp] = row[cp * 52 + (3 * c + i) * 4 + d];
  team_sync<64>();
#pragma unroll
  for (int i = 0; i < 3; i++) {
    dft<4, +1>(&v[4 * i]);
    fs_apply<+1, 4>(&v[4 * i], w.w16);
  }
#pragma unroll
  for (int eta = 0; eta < 12; eta++) row[c * 12 + d * 49 + eta] = v[eta];
  team_sync<64>();
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int dp = 0; dp < 4; dp++) e[4 * j + dp] = row[c * 12 + dp * 49 + 3 * d + j];
  team_sync<64>();
#pragma unroll
  for (int j = 0; j < 3; j++) dft<4, +1>(&e[4 * j]);
}

// the transpose of fs768_bwd's steps with conjugate twiddles
template <typename R>
__device__ __forceinline__ void fs768_fwd(cx<R> (&e)[12], cx<R> *lds, const FsTw<R> &w, int t) {
  const int p = t & 15, g = t >> 4, d = p & 3, c = p >> 2;
  cx<R> *row = lds + g * FS_RG;
#pragma unroll
  for (int j = 0; j < 3; j++) dft<4, -1>(&e[4 * j]);
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int dp = 0; dp < 4; dp++) row[c * 12 + dp * 49 + 3 * d + j] = e[4 * j + dp];
  team_sync<64>();
  cx<R> v[12];
#pragma unroll
  for (int eta = 0; eta < 12; eta++) v[eta] = row[c * 12 + d * 49 + eta];
  team_sync<64>();
#pragma unroll
  for (int i = 0; i < 3; i++) {
    fs_apply<-1, 4>(&v[4 * i], w.w16);
    dft<4, -1>(&v[4 * i]);
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int cp = 0; cp < 4; cp++) row[cp * 52 + (3 * c + i) * 4 + d] = v[4 * i + cp];
  team_sync<64>();
#pragma unroll
  for (int xi = 0; xi < 12; xi++) e[xi] = row[c * 52 + xi * 4 + d];
  team_sync<64>();
#pragma unroll
  for (int q = 0; q < 3; q++) {
    fs_apply<-1, 4>(&e[4 * q], w.w64);
    dft<4, -1>(&e[4 * q]);
  }
  fs_xpose(e);
  fs_apply<-1, 12>(e, w.w768);
  dft<12, -1>(e);
}

// Two independent transforms of one wave in an interleaved order: each LDS
// exchange of one transform is issued, then the other transform's next
// register stage runs while the exchange's reads are in flight.  One exchange
// area serves both, as the wave's LDS operations execute in order (a read of
// one exchange is done before the next exchange's writes land).  The same
// arithmetic as fs768_bwd / fs768_fwd on each array, so the results are
// bitwise those of two calls.
template <typename R, bool MID0 = false>
__device__ __forceinline__ void fs768_bwd2(cx<R> (&e)[12], cx<R> (&f)[12], cx<R> *lds, const FsTw<R> &w, int t) {
  const int p = t & 15, g = t >> 4, d = p & 3, c = p >> 2;
  cx<R> *row = lds + g * FS_RG;
  auto stage_a = [&](cx<R>(&x)[12]) {
    dft12<+1, MID0>(x);
    fs_apply<+1, 12>(x, w.w768);
    fs_xpose(x);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      dft<4, +1>(&x[4 * q]);
      fs_apply<+1, 4>(&x[4 * q], w.w64);
    }
  };
  auto xch1 = [&](const cx<R>(&x)[12], cx<R>(&v)[12]) {
#pragma unroll
    for (int xi = 0; xi < 12; xi++) row[c * 52 + xi * 4 + d] = x[xi];
    team_sync<64>();
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int cp = 0; cp < 4; cp++) v[4 * i + cp] = row[cp * 52 + (3 * c + i) * 4 + d];
    team_sync<64>();
  };
  auto stage_b = [&](cx<R>(&v)[12]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      dft<4, +1>(&v[4 * i]);
      fs_apply<+1, 4>(&v[4 * i], w.w16);
    }
  };
  auto xch2 = [&](const cx<R>(&v)[12], cx<R>(&x)[12]) {
#pragma unroll
    for (int eta = 0; eta < 12; eta++) row[c * 12 + d * 49 + eta] = v[eta];
    team_sync<64>();
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int dp = 0; dp < 4; dp++) x[4 * j + dp] = row[c * 12 + dp * 49 + 3 * d + j];
    team_sync<64>();
  };
  auto stage_c = [&](cx<R>(&x)[12]) {
#pragma unroll
    for (int j = 0; j < 3; j++) dft<4, +1>(&x[4 * j]);
  };
  cx<R> ve[12], vf[12];
  stage_a(e);
  xch1(e, ve);
  stage_a(f);  // while e's reads are in flight
  xch1(f, vf);
  stage_b(ve);
  xch2(ve, e);
  stage_b(vf);
  xch2(vf, f);
  stage_c(e);
  stage_c(f);
}

template <typename R>
__device__ __forceinline__ void fs768_fwd2(cx<R> (&e)[12], cx<R> (&f)[12], cx<R> *lds, const FsTw<R> &w, int t) {
  const int p = t & 15, g = t >> 4, d = p & 3, c = p >> 2;
  cx<R> *row = lds + g * FS_RG;
  auto stage_c = [&](cx<R>(&x)[12]) {
#pragma unroll
    for (int j = 0; j < 3; j++) dft<4, -1>(&x[4 * j]);
  };
  auto xch2 = [&](const cx<R>(&x)[12], cx<R>(&v)[12]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int dp = 0; dp < 4; dp++) row[c * 12 + dp * 49 + 3 * d + j] = x[4 * j + dp];
    team_sync<64>();
#pragma unroll
    for (int eta = 0; eta < 12; eta++) v[eta] = row[c * 12 + d * 49 + eta];
    team_sync<64>();
  };
  auto stage_b = [&](cx<R>(&v)[12]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      fs_apply<-1, 4>(&v[4 * i], w.w16);
      dft<4, -1>(&v[4 * i]);
    }
  };
  auto xch1 = [&](const cx<R>(&v)[12], cx<R>(&x)[12]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int cp = 0; cp < 4; cp++) row[cp * 52 + (3 * c + i) * 4 + d] = v[4 * i + cp];
    team_sync<64>();
#pragma unroll
    for (int xi = 0; xi < 12; xi++) x[xi] = row[c * 52 + xi * 4 + d];
    team_sync<64>();
  };
  auto stage_a = [&](cx<R>(&x)[12]) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
      fs_apply<-1, 4>(&x[4 * q], w.w64);
      dft<4, -1>(&x[4 * q]);
    }
    fs_xpose(x);
    fs_apply<-1, 12>(x, w.w768);
    dft<12, -1>(x);
  };
  cx<R> ve[12], vf[12];
  stage_c(e);
  xch2(e, ve);
  stage_c(f);  // while e's reads are in flight
  xch2(f, vf);
  stage_b(ve);
  xch1(ve, e);
  stage_b(vf);
  xch1(vf, f);
  stage_a(e);
  stage_a(f);
}

// ---------------------------------------------------- two-wave 1536 --
// A 1536-point transform held by a 128-lane team of two wavefronts, 12 values
// per lane: one radix-2 step between the waves, then the four-step 768-point
// transform in each wave.  area: the team's LDS, two rows of FS_LDS complex
// values; wave h = t / 64 runs its fs768 passes in row h.
//
// fs1536_bwd (sign +1) takes the slot layout (slot s of lane t = element
// t + 128 s).  Splitting the output into even and odd bins, the pairs
// (n, n + 768) sit in slots s and s + 6 of one lane:
//   a_n = x_n + x_{n+768},  b_n = (x_n - x_{n+768}) w1536^n,
// and X_{2m} / X_{2m+1} are the 768-point transforms of a / b.  Each wave
// sends the six values that the other one needs and takes six, so that wave
// 0 holds a and wave 1 holds b in the fs768 slot layout (element n' =
// t%64 + 64 s'); afterwards wave h holds X_{2m+h} for m in the layout P of
// fs768_index.
//
// fs1536_fwd (sign -1) is the transpose: wave h's fs768_fwd gives the
// transform E (h = 0) or O (h = 1) of the samples 2m + h, which it stores in
// its row in natural order; fs1536_mode then forms Z_k = E_k + w1536^k O_k and
// Z_{1536-k} = E_{768-k} + conj(w1536^k) O_{768-k} (w1536^768 = -1).  A
// team_sync after the last fs1536_mode frees the rows.
//
// tw: the FS2_TW-entry table of build_fs1536_twiddles: build_fs768_twiddles'
// FS_TW entries, then exp(-2 pi i t/1536) for t < 64 and exp(-2 pi i j/24)
// for j < 12.
constexpr int FS2_TW = FS_TW + 64 + 12;

template <typename T>
inline void build_fs1536_twiddles(cx<T> *w) {
  build_fs768_twiddles(w);
  const long double pi2 = 2.0L * 3.141592653589793238462643383279502884L;
  for (int t = 0; t < 64; t++) w[FS_TW + t] = cx<T>{(T)cosl(-pi2 * t / 1536), (T)sinl(-pi2 * t / 1536)};
  for (int j = 0; j < 12; j++) w[FS_TW + 64 + j] = cx<T>{(T)cosl(-pi2 * j / 24), (T)sinl(-pi2 * j / 24)};
}

// w1536^(t + 64 j), forward sign, t < 64, j < 12
template <typename R>
__device__ __forceinline__ cx<R> fs2_twiddle(const cx<R> *tw, int t, int j) {
  return mul(tw[FS_TW + t], tw[FS_TW + 64 + j]);
}

// c ? x : y per component (a select of whole complex values can become a
// two-entry private array indexed by c)
template <typename R>
__device__ __forceinline__ cx<R> pick(bool c, cx<R> x, cx<R> y) {
  return cx<R>{c ? x.x : y.x, c ? x.y : y.y};
}

template <typename R>
__device__ __forceinline__ void fs1536_bwd(cx<R> (&e)[12], cx<R> *area, const FsTw<R> &w, const cx<R> *tw, int t) {
  const int h = t >> 6, tl = t & 63;
  cx<R> *own = area + h * FS_LDS;
  const cx<R> *other = area + (1 - h) * FS_LDS;
  cx<R> keep[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const cx<R> a = add(e[s], e[s + 6]);
    const cx<R> b = mulc(sub(e[s], e[s + 6]), fs2_twiddle(tw, tl, h + 2 * s));  // n = tl + 64 (h + 2 s)
    keep[s] = pick(h != 0, b, a);
    own[s * 64 + tl] = pick(h != 0, a, b);
  }
  team_sync<128>();
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const cx<R> r = other[s * 64 + tl];
    e[2 * s] = pick(h != 0, r, keep[s]);
    e[2 * s + 1] = pick(h != 0, keep[s], r);
  }
  team_sync<128>();
  fs768_bwd(e, own, w, tl);
}

template <typename R>
__device__ __forceinline__ void fs1536_fwd(cx<R> (&e)[12], cx<R> *area, const FsTw<R> &w, int t) {
  const int h = t >> 6, tl = t & 63;
  cx<R> *own = area + h * FS_LDS;
  fs768_fwd(e, own, w, tl);
#pragma unroll
  for (int s = 0; s < 12; s++) own[tl + 64 * s] = e[s];
  team_sync<128>();
}

// Z_k and Z_{1536-k} for 0 <= k < 768 from the rows that fs1536_fwd filled
// (k = 0 gives Z_0 for both)
template <typename R>
__device__ __forceinline__ void fs1536_mode(const cx<R> *area, const cx<R> *tw, int k, cx<R> &zk, cx<R> &zm) {
  const cx<R> *E = area, *O = area + FS_LDS;
  const cx<R> wk = fs2_twiddle(tw, k & 63, k >> 6);
  zk = add(E[k], mul(O[k], wk));
  const int km = k ? 768 - k : 0;
  const cx<R> zr = add(E[km], mulc(O[km], wk));
  zm = pick(k != 0, zr, zk);
}

}  // namespace fft
}  // namespace chan
