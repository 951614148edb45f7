// k_solve.hip -- K5: the implicit viscous solve of one RK3 substep
// (linsolve, linsolve_blocking.inc:3-107) as three kernels:
//
//   k_solve_ve : one lane per (pair, system, real/imaginary part); system 0 =
//                eta, 1 = v.  The pairs (kx, kz) and (kx, -kz) have the same
//                k^2, hence the same matrix: their four lanes (sign x part)
//                sit side by side and share one row of stored lower factors.
//                Builds the pentadiagonal rows lambda*D - nu*L
//                of etamat / D2vmat, folds the wall conditions into rows 1,2
//                and ny-2,ny-1 (applybc_0 / applybc_n, dnsdata.f90:458-472),
//                factorises bottom-up with reciprocal pivots fused with the
//                backward sweep (LU5decompStep + LeftLU5divStep1,
//                rbparmat_blocking.f90:20-77), then runs the forward sweep
//                (LeftLU5divStep2, :82-100) and recovers wall and ghost values
//                (linsolve_blocking.inc:50-61).  The matrices are real, so the
//                real and imaginary parts of a right-hand side are independent
//                recurrences on adjacent lanes.  All four lanes of a group
//                store the same lower factors, so each reads back only what
//                it wrote itself.
//                The spare pair slot npair (plane padding) solves the
//                unit-forcing response ucor: the (0,0) eta matrix with
//                right-hand side 1 and homogeneous walls (:65-76).
//   k_mean     : the (0,0) mode: u/w split, flow rates by the three-point
//                integral (yintegr, dnsdata.f90:312-324), constant-flow
//                correction or CPI pressure gradient (linsolve_blocking.inc:62-97).
//   k_uw       : every other pair, one lane per real/imaginary part: dv/dy by
//                the compact first derivative (COMPLEXderiv + LeftLU5divStep2
//                with D0, dnsdata.f90:337-373) and u = (i kx v' - i kz eta)/k^2,
//                w = (i kz v' + i kx eta)/k^2 (linsolve_blocking.inc:99-103).
//
// The factorisation order (bottom-up UL, no pivoting) is the reference's, so
// FP64 results agree to rounding.
//
// Every sweep loads PF rows ahead into a register ring.  The ring loads are
// unconditional (row indices clamped into the column), so no ring register
// is ever the merge of a loaded and an old value: such a merge makes the
// compiler wait for the load at once and serialises the sweep on memory
// latency.  Full groups of PF rows run in the main loop, the last partial
// group from the ring alone.
#include "chan_internal.hpp"

namespace chan {

constexpr int PF = 4;  // rows of software prefetch in the sequential sweeps

// Block b of a grid of nb = 8m blocks solves column block (b % 8) m + b / 8:
// round-robin dispatch then gives each XCD a contiguous run of column blocks,
// so the cache lines that neighbouring blocks share (a block's pair range
// starts and ends inside a line) stay in one L2.  2% on k_solve_ve at
// 191x384x189; the extra blocks of the rounded grid exit at once.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return (b & 7) * (nb >> 3) + (b >> 3);
}
static unsigned xcd_grid(size_t lanes) {
  return (unsigned)(((lanes + 255) / 256 + 7) / 8 * 8);
}

// ------------------------------------------------------------ k_solve_ve --
__global__ __launch_bounds__(256) void k_solve_ve(SolveArgs a, const double *__restrict__ der) {
  // der is a separate restrict argument: the compiler then proves that the
  // kernel's stores never reach it and reads its wave-uniform rows through
  // the scalar cache, off the vector-memory counter that the sweeps' own
  // prefetches wait on.
  const Geom g = a.g;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  // lane = 4*group + 2*sign + part; group (ix, |iz|) for iz >= 0, part 0 real,
  // 1 imaginary; the group after the last one solves ucor in pair slot npair
  const int grp = gid >> 2, sgn = (gid >> 1) & 1, part = gid & 1;
  const int ngrp = g.Nx * (g.nz + 1);
  if (grp > ngrp) return;
  const bool spare = grp == ngrp;
  const int ix = spare ? 0 : grp / (g.nz + 1), iza = spare ? 0 : grp % (g.nz + 1);
  if (sgn && (spare || iza == 0)) return;  // kz = 0 has no mirror
  const int iz = sgn ? -iza : iza;
  const int p = spare ? g.npair : ix * g.Nz + g.nz + iz;
  const int sys = blockIdx.y;  // 0: eta (component 1), 1: v (component 2)
  const int ny = g.ny;
  const size_t gs = ((size_t)ngrp + 1 + 63) / 64 * 64;  // group stride of L
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const bool mean = !spare && ix == 0 && iz == 0;
  const size_t ps = g.ps, rs = 2 * ps;  // row stride in doubles
  double *col = reinterpret_cast<double *>(a.V + (size_t)sys * g.Ny * ps + p) + part;  // row iy at col[(iy+1)*rs]
  const double *rc = reinterpret_cast<const double *>(a.rhs + (size_t)sys * (ny - 1) * ps + p) + part;
  double2 *Lc = a.L + (size_t)sys * (ny - 1) * gs + grp;  // row iy at Lc[(iy-1)*gs]
  const double *bc0 = a.bc[sys ? BC_V0 : BC_ETA0], *bc0m1 = a.bc[sys ? BC_V0M1 : BC_ETA0M1];
  const double *bcn = a.bc[sys ? BC_VN : BC_ETAN], *bcnp1 = a.bc[sys ? BC_VNP1 : BC_ETANP1];
  // wall value of eta: (Re u - Im w) + i (Im u + Re w) with u = u0/uN, w = 0
  // in the (0,0) mode, zero otherwise; v and dv/dy vanish at the walls.
  const double eta0 = (mean && sys == 0 && part == 0) ? a.u0 : 0.0;
  const double etan = (mean && sys == 0 && part == 0) ? a.uN : 0.0;
  const double bspare = (sys == 0 && part == 0) ? 1.0 : 0.0;  // ucor right-hand side

  // ---- bottom-up: rows, wall folding, factorisation, backward sweep
  double l1m2 = 0.0, l1m1 = 0.0, l2m2 = 0.0, l2m1 = 0.0;  // lower factors of rows iy+1, iy+2
  double x1 = 0.0, x2 = 0.0;
  auto up_row = [&](const int iy, double b) {
    if (spare) b = bspare;
    const double *d = der + (size_t)(iy - 1) * 20;
    double A[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const double d0 = d[j], d2 = d[10 + j], d4 = d[15 + j];
      if (sys) {
        const double os = a.ni * (d4 - 2.0 * k2 * d2 + k2 * k2 * d0);
        A[j] = a.lambda * (d2 - k2 * d0) - os;
      } else {
        const double sq = a.ni * (d2 - k2 * d0);
        A[j] = a.lambda * d0 - sq;
      }
    }
    if (iy == 1) {  // applybc_0 row 1: ghost then wall column
      const double e = A[0];
      for (int j = 1; j <= 4; j++) A[j] = A[j] - e * bc0m1[j] / bc0m1[0];
      const double f = A[1];
      for (int j = 2; j <= 4; j++) A[j] = A[j] - f * bc0[j] / bc0[1];
      if (sys == 0) b = b - A[1] * eta0 / bc0[1];
    } else if (iy == 2) {
      if (sys == 0) b = b - A[0] * eta0 / bc0[1];
      const double e = A[0];
      for (int j = 1; j <= 3; j++) A[j] = A[j] - e * bc0[j + 1] / bc0[1];
    }
    if (iy == ny - 1) {  // applybc_n row ny-1
      const double e = A[4];
      for (int j = 0; j <= 3; j++) A[j] = A[j] - e * bcnp1[j] / bcnp1[4];
      const double f = A[3];
      for (int j = 0; j <= 2; j++) A[j] = A[j] - f * bcn[j] / bcn[3];
      if (sys == 0) b = b - A[3] * etan / bcn[3];
      A[3] = 0.0;
      A[4] = 0.0;
    } else if (iy == ny - 2) {
      if (sys == 0) b = b - A[4] * etan / bcn[3];
      const double e = A[4];
      for (int j = 1; j <= 3; j++) A[j] = A[j] - e * bcn[j - 1] / bcn[3];
      A[4] = 0.0;
    }
    // factorise row iy against the already factorised rows iy+1, iy+2
    const double p2 = A[4];
    A[3] = A[3] - p2 * l2m1;
    A[2] = A[2] - p2 * l2m2;
    const double p1 = A[3];
    A[2] = A[2] - p1 * l1m1;
    A[1] = A[1] - p1 * l1m2;
    const double r = 1.0 / A[2];
    A[0] = A[0] * r;
    A[1] = A[1] * r;
    // backward sweep
    const double xi = (b - (p1 * x1 + p2 * x2)) * r;
    col[(size_t)(iy + 1) * rs] = xi;
    Lc[(size_t)(iy - 1) * gs] = make_double2(iy <= 2 ? 0.0 : A[0], iy == 1 ? 0.0 : A[1]);
    l2m2 = l1m2;
    l2m1 = l1m1;
    l1m2 = A[0];
    l1m1 = A[1];
    x2 = x1;
    x1 = xi;
  };
  {
    double ring[PF];  // ring[u] holds the right-hand side of row iy-u
#pragma unroll
    for (int u = 0; u < PF; u++) ring[u] = rc[(size_t)max(ny - 2 - u, 0) * rs];
    int iy = ny - 1;
    for (; iy - (PF - 1) >= 1; iy -= PF) {
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int r = iy - u;
        const double b = ring[u];
        ring[u] = rc[(size_t)max(r - PF - 1, 0) * rs];
        up_row(r, b);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
      if (iy - u >= 1) up_row(iy - u, ring[u]);
  }

  // ---- top-down: forward sweep, keep rows 1..3 and ny-3..ny-1
  double b1 = 0.0, b2 = 0.0, b3 = 0.0;
  double bot[3] = {0.0, 0.0, 0.0};
  auto down_row = [&](const int iy, const double x, const double2 L) {
    const double bi = x - (L.x * b2 + L.y * b1);
    col[(size_t)(iy + 1) * rs] = bi;
    if (iy <= 3) bot[iy - 1] = bi;
    b3 = b2;
    b2 = b1;
    b1 = bi;
  };
  {
    double xr[PF];
    double2 lr[PF];
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int r = min(1 + u, ny - 1);
      xr[u] = col[(size_t)(r + 1) * rs];
      lr[u] = Lc[(size_t)(r - 1) * gs];
    }
    int iy = 1;
    for (; iy + (PF - 1) <= ny - 1; iy += PF) {
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int r = iy + u;
        const double x = xr[u];
        const double2 L = lr[u];
        const int nr = min(r + PF, ny - 1);
        xr[u] = col[(size_t)(nr + 1) * rs];
        lr[u] = Lc[(size_t)(nr - 1) * gs];
        down_row(r, x, L);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
      if (iy + u <= ny - 1) down_row(iy + u, xr[u], lr[u]);
  }
  // wall and ghost rows
  {
    const double s0 = bc0[2] * bot[0] + bc0[3] * bot[1] + bc0[4] * bot[2];
    const double w0 = (eta0 - s0) / bc0[1];
    const double s1 = bc0m1[1] * w0 + bc0m1[2] * bot[0] + bc0m1[3] * bot[1] + bc0m1[4] * bot[2];
    const double wm1 = sys ? (0.0 - s1) / bc0m1[0] : -(s1 / bc0m1[0]);
    col[rs] = w0;
    col[0] = wm1;
  }
  {
    const double s0 = bcn[0] * b3 + bcn[1] * b2 + bcn[2] * b1;  // rows ny-3, ny-2, ny-1
    const double wn = (etan - s0) / bcn[3];
    const double s1 = bcnp1[0] * b3 + bcnp1[1] * b2 + bcnp1[2] * b1 + bcnp1[3] * wn;
    const double wp1 = sys ? (0.0 - s1) / bcnp1[4] : -(s1 / bcnp1[4]);
    col[(size_t)(ny + 1) * rs] = wn;
    col[(size_t)(ny + 2) * rs] = wp1;
  }
}

// ---------------------------------------------------------------- k_mean --
// yintegr (dnsdata.f90:312-324): the three-point rule on the odd rows; f and y
// are indexed by row + 1.  Term t covers rows 2t+1 .. 2t+3.
__device__ __forceinline__ double yintegr_term(const double *y, const double *f, int iy) {
  const double yp1 = y[iy + 2] - y[iy + 1], ym1 = y[iy] - y[iy + 1];
  const double a1 = -(1.0 / 3.0 * ym1) + 1.0 / 6.0 * yp1 + 1.0 / 6.0 * yp1 * yp1 / ym1;
  const double a3 = +(1.0 / 3.0 * yp1) - 1.0 / 6.0 * ym1 - 1.0 / 6.0 * ym1 * ym1 / yp1;
  const double a2 = yp1 - ym1 - a1 - a3;
  return a1 * f[iy] + a2 * f[iy + 1] + a3 * f[iy + 2];
}

// Sums yintegr of nf columns (stride Ny in LDS) with the block; red holds
// nf*256 doubles.  Every thread of the block calls it.
template <int NF>
__device__ void yintegr_block(const double *y, const double *f, int Ny, int ny, double *red, double *out) {
  double s[NF];
#pragma unroll
  for (int k = 0; k < NF; k++) s[k] = 0.0;
  for (int iy = 1 + 2 * (int)threadIdx.x; iy <= ny - 1; iy += 2 * (int)blockDim.x)
#pragma unroll
    for (int k = 0; k < NF; k++) s[k] += yintegr_term(y, f + k * Ny, iy);
#pragma unroll
  for (int k = 0; k < NF; k++) red[k * 256 + threadIdx.x] = s[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < NF; k++) red[k * 256 + threadIdx.x] += red[k * 256 + threadIdx.x + h];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NF; k++) out[k] = red[k * 256];
}

__global__ __launch_bounds__(256) void k_mean(MeanArgs a) {
  extern __shared__ double sm[];
  const Geom g = a.g;
  const int Ny = g.Ny, ny = g.ny;
  const size_t ps = g.ps, pstride = (size_t)Ny * ps;
  const int p0 = g.nz;  // pair (ix=0, iz=0)
  double *ur = sm, *wr = sm + Ny, *uc = sm + 2 * Ny, *yy = sm + 3 * Ny, *red = sm + 4 * Ny;
  __shared__ double cv[4];  // flag x, flag z, corrpx, corrpz
  for (int i = threadIdx.x; i < Ny; i += blockDim.x) {
    const cx<double> e = a.V[(size_t)i * ps + p0];  // eta slot: Re u + i Re w
    ur[i] = e.x;
    wr[i] = e.y;
    uc[i] = a.V[(size_t)i * ps + g.npair].x;  // ucor from the spare slot
    yy[i] = a.y[i];
  }
  __syncthreads();
  double fr[3];
  yintegr_block<3>(yy, sm, Ny, ny, red, fr);
  if (threadIdx.x == 0) {
    DevScalars *sc = a.sc;
    sc->fr[0] = fr[0];
    sc->fr[1] = fr[1];
    sc->fr[2] = fr[2];
    cv[0] = 0.0;
    cv[1] = 0.0;
    if (fabs(a.meanflowx) > 1.0e-7 && !a.cpi) {
      sc->corrpx = (a.meanflowx - fr[0]) / fr[2];
      cv[0] = 1.0;
    }
    if (fabs(a.meanflowz) > 1.0e-7 && !a.cpi) {
      sc->corrpz = (a.meanflowz - fr[1]) / fr[2];
      cv[1] = 1.0;
    }
    cv[2] = sc->corrpx;
    cv[3] = sc->corrpz;
    if (a.cpi) {
      if (a.cpi_type == 0)
        sc->meanpx = (1 - a.gamma) * 6 * a.ni / fr[0];
      else
        sc->meanpx = (1.5 / a.gamma) * fr[0] * a.ni;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < Ny; i += blockDim.x) {
    double u = ur[i], w = wr[i];
    if (cv[0] != 0.0) u = u + cv[2] * uc[i];
    if (cv[1] != 0.0) w = w + cv[3] * uc[i];
    a.V[(size_t)i * ps + p0] = cx<double>{u, 0.0};
    a.V[2 * pstride + (size_t)i * ps + p0] = cx<double>{w, 0.0};
  }
}

// ------------------------------------------------------------------ k_uw --
// Blocks of two waves: the kernel has few waves (one lane per pair and part)
// and each runs a long sequential sweep, so smaller blocks spread them more
// evenly over the CUs (5% faster than 256-thread blocks at 191x384x189).
constexpr int UW_BLOCK = 128;
__global__ __launch_bounds__(UW_BLOCK) void k_uw(UwArgs a, const double *__restrict__ der, const double *__restrict__ d0fac) {
  const Geom g = a.g;
  const int gid = blockIdx.x * UW_BLOCK + threadIdx.x;
  const int p = gid >> 1, part = gid & 1;  // part 0 real, 1 imaginary
  if (p >= g.npair) return;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  if (ix == 0 && iz == 0) return;  // both lanes of the (0,0) pair leave together
  const int ny = g.ny;
  const double al = g.alfa0 * ix, be = g.beta0 * iz;
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const size_t ps = g.ps, rs = 2 * ps, pstride = (size_t)g.Ny * ps;
  double *c1 = reinterpret_cast<double *>(a.V + p) + part;
  double *c2 = reinterpret_cast<double *>(a.V + pstride + p) + part;
  double *c3 = reinterpret_cast<double *>(a.V + 2 * pstride + p) + part;
#define ROW(c, iy) (c)[(size_t)((iy) + 1) * rs]
  // one-sided wall and ghost derivatives of this part of v
  double vb[5], vt[5];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    vb[j] = ROW(c2, -1 + j);
    vt[j] = ROW(c2, ny - 3 + j);
  }
  auto dot5 = [](const double *w, const double *v) {
    double s = w[0] * v[0];
#pragma unroll
    for (int j = 1; j < 5; j++) s = s + w[j] * v[j];
    return s;
  };
  const double f10 = dot5(a.bnd[B_D140], vb), f1m1 = dot5(a.bnd[B_D14M1], vb);
  const double f1n = dot5(a.bnd[B_D14N], vt), f1np1 = dot5(a.bnd[B_D14NP1], vt);

  // bottom-up: interior D1 v with boundary fix-ups and the D0 backward sweep;
  // window w[j] = v(iy-2+j)
  double w[5];
#pragma unroll
  for (int j = 0; j < 5; j++) w[j] = vt[j];
  double x1 = f1n, x2 = f1np1;
  auto up_row = [&](const int iy, const double vnext) {  // vnext = v(iy-3)
    const double *d = der + (size_t)(iy - 1) * 20;
    double f = d[5] * w[0];
#pragma unroll
    for (int j = 1; j < 5; j++) f = f + d[5 + j] * w[j];
    if (iy == ny - 1) f = f - (d[3] * f1n + d[4] * f1np1);
    if (iy == ny - 2) f = f - d[4] * f1n;
    if (iy == 2) f = f - d[0] * f10;
    if (iy == 1) f = f - (d[1] * f10 + d[0] * f1m1);
    const double *D = d0fac + (size_t)(iy - 1) * 5;
    const double xi = (f - (D[3] * x1 + D[4] * x2)) * D[2];
    ROW(c3, iy) = xi;
    x2 = x1;
    x1 = xi;
#pragma unroll
    for (int j = 4; j > 0; j--) w[j] = w[j - 1];
    w[0] = vnext;
  };
  {
    double ring[PF];  // ring[u] holds v(iy-u-3)
#pragma unroll
    for (int u = 0; u < PF; u++) ring[u] = ROW(c2, max(ny - 4 - u, -1));
    int iy = ny - 1;
    for (; iy - (PF - 1) >= 1; iy -= PF) {
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int r = iy - u;
        const double v = ring[u];
        ring[u] = ROW(c2, max(r - 3 - PF, -1));
        up_row(r, v);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
      if (iy - u >= 1) up_row(iy - u, ring[u]);
  }

  // top-down: D0 forward sweep, then u and w on every row; the other part of
  // v' and eta comes from the neighbouring lane
  double b1 = f10, b2 = f1m1;
  auto down_row = [&](const int iy, const double x, const double eta) {
    double vp;
    if (iy == -1)
      vp = f1m1;
    else if (iy == 0)
      vp = f10;
    else if (iy <= ny - 1) {
      const double *D = d0fac + (size_t)(iy - 1) * 5;
      vp = x - (D[0] * b2 + D[1] * b1);
      b2 = b1;
      b1 = vp;
    } else
      vp = iy == ny ? f1n : f1np1;
    const double vo = __shfl_xor(vp, 1, 64), eo = __shfl_xor(eta, 1, 64);
    // u = (i al v' - i be eta)/k2, w = (i be v' + i al eta)/k2, i*a*X = (-a Xi, a Xr)
    double uu, ww;
    if (part == 0) {
      uu = (-(al * vo)) - (-(be * eo));
      ww = (-(be * vo)) + (-(al * eo));
    } else {
      uu = (al * vo) - (be * eo);
      ww = (be * vo) + (al * eo);
    }
    ROW(c1, iy) = uu / k2;
    ROW(c3, iy) = ww / k2;
  };
  {
    double xr[PF], er[PF];  // D1 v / D0 backward result and eta of row iy+u
#pragma unroll
    for (int u = 0; u < PF; u++) {
      xr[u] = ROW(c3, -1 + u);
      er[u] = ROW(c1, -1 + u);
    }
    int iy = -1;
    for (; iy + (PF - 1) <= ny + 1; iy += PF) {
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int r = iy + u;
        const double x = xr[u], e = er[u];
        const int nr = min(r + PF, ny + 1);
        xr[u] = ROW(c3, nr);
        er[u] = ROW(c1, nr);
        down_row(r, x, e);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
      if (iy + u <= ny + 1) down_row(iy + u, xr[u], er[u]);
  }
#undef ROW
}

// ------------------------------------------------------- initial flow rate --
__global__ __launch_bounds__(256) void k_flowrate(MeanArgs a) {
  extern __shared__ double sm[];
  const Geom g = a.g;
  const int Ny = g.Ny, ny = g.ny;
  const size_t ps = g.ps, pstride = (size_t)Ny * ps;
  double *ur = sm, *wr = sm + Ny, *yy = sm + 2 * Ny, *red = sm + 3 * Ny;
  for (int i = threadIdx.x; i < Ny; i += blockDim.x) {
    ur[i] = a.V[(size_t)i * ps + g.nz].x;
    wr[i] = a.V[2 * pstride + (size_t)i * ps + g.nz].x;
    yy[i] = a.y[i];
  }
  __syncthreads();
  double fr[2];
  yintegr_block<2>(yy, sm, Ny, ny, red, fr);
  if (threadIdx.x == 0) {
    DevScalars *sc = a.sc;
    sc->fr[0] = fr[0];
    sc->fr[1] = fr[1];
    sc->fr[2] = 0.0;
    if (a.cpi) {
      if (a.cpi_type == 0)
        sc->meanpx = (1 - a.gamma) * 6 * a.ni / sc->fr[0];
      else
        sc->meanpx = (1.5 / a.gamma) * sc->fr[0] * a.ni;
    }
  }
}

// ------------------------------------------------------------ launchers --
hipError_t launch_solve(hipStream_t st, const SolveArgs &a) {
  const size_t lanes = (size_t)4 * (a.g.Nx * (a.g.nz + 1) + 1);
  dim3 grid(xcd_grid(lanes), 2);
  hipLaunchKernelGGL(k_solve_ve, grid, dim3(256), 0, st, a, a.der);
  return hipGetLastError();
}
hipError_t launch_mean(hipStream_t st, const MeanArgs &a) {
  hipLaunchKernelGGL(k_mean, dim3(1), dim3(256), sizeof(double) * (4 * a.g.Ny + 3 * 256), st, a);
  return hipGetLastError();
}
hipError_t launch_uw(hipStream_t st, const UwArgs &a) {
  hipLaunchKernelGGL(k_uw, dim3((unsigned)((2 * a.g.npair + UW_BLOCK - 1) / UW_BLOCK)), dim3(UW_BLOCK), 0, st, a, a.der,
                     a.d0fac);
  return hipGetLastError();
}
hipError_t launch_flowrate(hipStream_t st, const MeanArgs &a) {
  hipLaunchKernelGGL(k_flowrate, dim3(1), dim3(256), sizeof(double) * (3 * a.g.Ny + 2 * 256), st, a);
  return hipGetLastError();
}

}  // namespace chan
