// k_solve.hip -- K5: the implicit viscous solve of one RK3 substep
// (linsolve, linsolve_blocking.inc:3-107) as three kernels:
//
//   k_solve_ve : one lane per (pair, system, real/imaginary part); system 0 =
//                eta, 1 = v.  Builds the pentadiagonal rows lambda*D - nu*L
//                of etamat / D2vmat, folds the wall conditions into rows 1,2
//                and ny-2,ny-1 (applybc_0 / applybc_n, dnsdata.f90:458-472),
//                factorises bottom-up with reciprocal pivots fused with the
//                backward sweep (LU5decompStep + LeftLU5divStep1,
//                rbparmat_blocking.f90:20-77), then runs the forward sweep
//                (LeftLU5divStep2, :82-100) and recovers wall and ghost values
//                (linsolve_blocking.inc:50-61).  The matrices are real, so the
//                real and imaginary parts of a right-hand side are independent
//                recurrences on adjacent lanes.  Both lanes store the same
//                lower factors, so each reads back only what it wrote.  The
//                real lane of the (0,0) eta system also solves the
//                unit-forcing response ucor (:65-68).
//   k_mean     : the (0,0) mode: u/w split, ucor ghosts, flow rates by the
//                three-point integral (yintegr, dnsdata.f90:312-324),
//                constant-flow correction or CPI pressure gradient
//                (linsolve_blocking.inc:62-97).
//   k_uw       : every other pair, one lane per real/imaginary part: dv/dy by
//                the compact first derivative (COMPLEXderiv + LeftLU5divStep2
//                with D0, dnsdata.f90:337-373) and u = (i kx v' - i kz eta)/k^2,
//                w = (i kz v' + i kx eta)/k^2 (linsolve_blocking.inc:99-103).
//
// The factorisation order (bottom-up UL, no pivoting) is the reference's, so
// FP64 results agree to rounding.  The row loops load PF rows ahead.
#include "chan_internal.hpp"

namespace chan {

constexpr int PF = 4;  // rows of software prefetch in the sequential sweeps

// ------------------------------------------------------------ k_solve_ve --
__global__ __launch_bounds__(256) void k_solve_ve(SolveArgs a) {
  const Geom g = a.g;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int p = gid >> 1, part = gid & 1;  // part 0 real, 1 imaginary
  if (p >= g.npair) return;
  const int sys = blockIdx.y;  // 0: eta (component 1), 1: v (component 2)
  const int ny = g.ny;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const bool mean = (ix == 0 && iz == 0);
  const bool do_ucor = mean && sys == 0 && part == 0;
  const size_t ps = g.ps, rs = 2 * ps;  // row stride in doubles
  double *col = reinterpret_cast<double *>(a.V + (size_t)sys * g.Ny * ps + p) + part;  // row iy at col[(iy+1)*rs]
  const double *rc = reinterpret_cast<const double *>(a.rhs + (size_t)sys * (ny - 1) * ps + p) + part;
  double2 *Lc = a.L + (size_t)sys * (ny - 1) * ps + p;
  const double *bc0 = a.bc[sys ? BC_V0 : BC_ETA0], *bc0m1 = a.bc[sys ? BC_V0M1 : BC_ETA0M1];
  const double *bcn = a.bc[sys ? BC_VN : BC_ETAN], *bcnp1 = a.bc[sys ? BC_VNP1 : BC_ETANP1];
  // wall value of eta: (Re u - Im w) + i (Im u + Re w) with u = u0/uN, w = 0
  // in the (0,0) mode, zero otherwise; v and dv/dy vanish at the walls.
  const double eta0 = (mean && sys == 0 && part == 0) ? a.u0 : 0.0;
  const double etan = (mean && sys == 0 && part == 0) ? a.uN : 0.0;

  // ---- bottom-up: rows, wall folding, factorisation, backward sweep
  double l1m2 = 0.0, l1m1 = 0.0, l2m2 = 0.0, l2m1 = 0.0;  // lower factors of rows iy+1, iy+2
  double x1 = 0.0, x2 = 0.0, u1 = 0.0, u2 = 0.0;
  double ring[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int iy = ny - 1 - u;
    ring[u] = iy >= 1 ? rc[(size_t)(iy - 1) * rs] : 0.0;
  }
  for (int base = ny - 1; base >= 1; base -= PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int iy = base - u;
      if (iy < 1) break;
      double b = ring[u];
      if (iy - PF >= 1) ring[u] = rc[(size_t)(iy - PF - 1) * rs];
      const double *d = a.der + (size_t)(iy - 1) * 20;
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
      Lc[(size_t)(iy - 1) * ps] = make_double2(iy <= 2 ? 0.0 : A[0], iy == 1 ? 0.0 : A[1]);
      if (do_ucor) {
        const double ui = (1.0 - (p1 * u1 + p2 * u2)) * r;
        a.ucor[iy + 1] = cx<double>{ui, 0.0};
        u2 = u1;
        u1 = ui;
      }
      l2m2 = l1m2;
      l2m1 = l1m1;
      l1m2 = A[0];
      l1m1 = A[1];
      x2 = x1;
      x1 = xi;
    }
  }

  // ---- top-down: forward sweep, keep rows 1..3 and ny-3..ny-1
  double b1 = 0.0, b2 = 0.0, b3 = 0.0, ub1 = 0.0, ub2 = 0.0;
  double bot[3] = {0.0, 0.0, 0.0};
  double xr[PF];
  double2 lr[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int iy = 1 + u;
    xr[u] = iy <= ny - 1 ? col[(size_t)(iy + 1) * rs] : 0.0;
    lr[u] = iy <= ny - 1 ? Lc[(size_t)(iy - 1) * ps] : make_double2(0.0, 0.0);
  }
  for (int base = 1; base <= ny - 1; base += PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int iy = base + u;
      if (iy > ny - 1) break;
      const double x = xr[u];
      const double2 L = lr[u];
      if (iy + PF <= ny - 1) {
        xr[u] = col[(size_t)(iy + PF + 1) * rs];
        lr[u] = Lc[(size_t)(iy + PF - 1) * ps];
      }
      const double bi = x - (L.x * b2 + L.y * b1);
      col[(size_t)(iy + 1) * rs] = bi;
      if (iy <= 3) bot[iy - 1] = bi;
      b3 = b2;
      b2 = b1;
      b1 = bi;
      if (do_ucor) {
        const double ux = a.ucor[iy + 1].x;
        const double ui = ux - (L.x * ub2 + L.y * ub1);
        a.ucor[iy + 1] = cx<double>{ui, 0.0};
        ub2 = ub1;
        ub1 = ui;
      }
    }
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
__device__ double yintegr_dev(const double *y, const double *f, int ny) {
  // f[iy+1], y[iy+1]
  double II = 0.0;
  for (int iy = 1; iy <= ny - 1; iy += 2) {
    const double yp1 = y[iy + 2] - y[iy + 1], ym1 = y[iy] - y[iy + 1];
    const double a1 = -(1.0 / 3.0 * ym1) + 1.0 / 6.0 * yp1 + 1.0 / 6.0 * yp1 * yp1 / ym1;
    const double a3 = +(1.0 / 3.0 * yp1) - 1.0 / 6.0 * ym1 - 1.0 / 6.0 * ym1 * ym1 / yp1;
    const double a2 = yp1 - ym1 - a1 - a3;
    II = II + a1 * f[iy] + a2 * f[iy + 1] + a3 * f[iy + 2];
  }
  return II;
}

__global__ __launch_bounds__(256) void k_mean(MeanArgs a) {
  extern __shared__ double sm[];
  const Geom g = a.g;
  const int Ny = g.Ny, ny = g.ny;
  const size_t ps = g.ps, pstride = (size_t)Ny * ps;
  const int p0 = g.nz;  // pair (ix=0, iz=0)
  double *ur = sm, *wr = sm + Ny, *uc = sm + 2 * Ny, *yy = sm + 3 * Ny;
  __shared__ double cv[4];  // flag x, flag z, corrpx, corrpz
  for (int i = threadIdx.x; i < Ny; i += blockDim.x) {
    const cx<double> e = a.V[(size_t)i * ps + p0];  // eta slot: Re u + i Re w
    ur[i] = e.x;
    wr[i] = e.y;
    uc[i] = (i >= 2 && i <= ny) ? a.ucor[i].x : 0.0;
    yy[i] = a.y[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double *e = a.bc[BC_ETA0], *f = a.bc[BC_ETA0M1], *en = a.bc[BC_ETAN], *fn = a.bc[BC_ETANP1];
    // ucor wall and ghost values (linsolve_blocking.inc:69-76)
    uc[1] = -((uc[2] * e[2] + uc[3] * e[3] + uc[4] * e[4]) / e[1]);
    uc[0] = -((uc[1] * f[1] + uc[2] * f[2] + uc[3] * f[3] + uc[4] * f[4]) / f[0]);
    uc[ny + 1] = -((uc[ny - 2] * en[0] + uc[ny - 1] * en[1] + uc[ny] * en[2]) / en[3]);
    uc[ny + 2] = -((uc[ny - 2] * fn[0] + uc[ny - 1] * fn[1] + uc[ny] * fn[2] + uc[ny + 1] * fn[3]) / fn[4]);
    DevScalars *sc = a.sc;
    const double fr1 = yintegr_dev(yy, ur, ny), fr2 = yintegr_dev(yy, wr, ny), fr3 = yintegr_dev(yy, uc, ny);
    sc->fr[0] = fr1;
    sc->fr[1] = fr2;
    sc->fr[2] = fr3;
    cv[0] = 0.0;
    cv[1] = 0.0;
    if (fabs(a.meanflowx) > 1.0e-7 && !a.cpi) {
      sc->corrpx = (a.meanflowx - fr1) / fr3;
      cv[0] = 1.0;
    }
    if (fabs(a.meanflowz) > 1.0e-7 && !a.cpi) {
      sc->corrpz = (a.meanflowz - fr2) / fr3;
      cv[1] = 1.0;
    }
    cv[2] = sc->corrpx;
    cv[3] = sc->corrpz;
    if (a.cpi) {
      if (a.cpi_type == 0)
        sc->meanpx = (1 - a.gamma) * 6 * a.ni / fr1;
      else
        sc->meanpx = (1.5 / a.gamma) * fr1 * a.ni;
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
__global__ __launch_bounds__(256) void k_uw(UwArgs a) {
  const Geom g = a.g;
  const int gid = blockIdx.x * 256 + threadIdx.x;
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
  // window w[j] = v(iy-2+j), ring holds v(iy-3-u) loaded PF rows early
  double w[5];
#pragma unroll
  for (int j = 0; j < 5; j++) w[j] = vt[j];
  double ring[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int r = ny - 1 - 3 - u;
    ring[u] = r >= -1 ? ROW(c2, r) : 0.0;
  }
  double x1 = f1n, x2 = f1np1;
  for (int base = ny - 1; base >= 1; base -= PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int iy = base - u;
      if (iy < 1) break;
      const double nextv = ring[u];  // v(iy-3)
      if (iy - 3 - PF >= -1) ring[u] = ROW(c2, iy - 3 - PF);
      const double *d = a.der + (size_t)(iy - 1) * 20;
      double f = d[5] * w[0];
#pragma unroll
      for (int j = 1; j < 5; j++) f = f + d[5 + j] * w[j];
      if (iy == ny - 1) f = f - (d[3] * f1n + d[4] * f1np1);
      if (iy == ny - 2) f = f - d[4] * f1n;
      if (iy == 2) f = f - d[0] * f10;
      if (iy == 1) f = f - (d[1] * f10 + d[0] * f1m1);
      const double *D = a.d0fac + (size_t)(iy - 1) * 5;
      const double xi = (f - (D[3] * x1 + D[4] * x2)) * D[2];
      ROW(c3, iy) = xi;
      x2 = x1;
      x1 = xi;
#pragma unroll
      for (int j = 4; j > 0; j--) w[j] = w[j - 1];
      w[0] = nextv;
    }
  }
  // top-down: D0 forward sweep, then u and w on every row; the other part of
  // v' and eta comes from the neighbouring lane
  double b1 = f10, b2 = f1m1;
  double xr[PF], er[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) {
    const int iy = -1 + u;
    xr[u] = (iy >= 1 && iy <= ny - 1) ? ROW(c3, iy) : 0.0;
    er[u] = ROW(c1, iy);
  }
  for (int base = -1; base <= ny + 1; base += PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int iy = base + u;
      const bool live = iy <= ny + 1;  // uniform across the wave
      double vp = 0.0, eta = 0.0;
      if (live) {
        const double x = xr[u];
        eta = er[u];
        const int nxt = iy + PF;
        if (nxt <= ny + 1) {
          xr[u] = (nxt >= 1 && nxt <= ny - 1) ? ROW(c3, nxt) : 0.0;
          er[u] = ROW(c1, nxt);
        }
        if (iy == -1)
          vp = f1m1;
        else if (iy == 0)
          vp = f10;
        else if (iy <= ny - 1) {
          const double *D = a.d0fac + (size_t)(iy - 1) * 5;
          vp = x - (D[0] * b2 + D[1] * b1);
          b2 = b1;
          b1 = vp;
        } else
          vp = iy == ny ? f1n : f1np1;
      }
      const double vo = __shfl_xor(vp, 1, 64), eo = __shfl_xor(eta, 1, 64);
      if (live) {
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
      }
    }
  }
#undef ROW
}

// ------------------------------------------------------- initial flow rate --
__global__ __launch_bounds__(256) void k_flowrate(MeanArgs a) {
  extern __shared__ double sm[];
  const Geom g = a.g;
  const int Ny = g.Ny, ny = g.ny;
  const size_t ps = g.ps, pstride = (size_t)Ny * ps;
  double *ur = sm, *wr = sm + Ny, *yy = sm + 2 * Ny;
  for (int i = threadIdx.x; i < Ny; i += blockDim.x) {
    ur[i] = a.V[(size_t)i * ps + g.nz].x;
    wr[i] = a.V[2 * pstride + (size_t)i * ps + g.nz].x;
    yy[i] = a.y[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    DevScalars *sc = a.sc;
    sc->fr[0] = yintegr_dev(yy, ur, ny);
    sc->fr[1] = yintegr_dev(yy, wr, ny);
    sc->fr[2] = 0.0;
    if (a.cpi) {
      if (a.cpi_type == 0)
        sc->meanpx = (1 - a.gamma) * 6 * a.ni / sc->fr[0];
      else
        sc->meanpx = (1.5 / a.gamma) * sc->fr[0] * a.ni;
    }
  }
}

// ------------------------------------------------------------ CFL reduce --
__global__ __launch_bounds__(256) void k_cfl_reduce(const double *__restrict__ part, int n, DevScalars *sc) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double c = part[i];
    m = (c > m || c != c) ? c : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = red[threadIdx.x + s], c = red[threadIdx.x];
      red[threadIdx.x] = (o > c || o != o) ? o : c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cur = __longlong_as_double((long long)sc->cfl_bits);
    const double v = red[0];
    if (v > cur || v != v) sc->cfl_bits = (unsigned long long)__double_as_longlong(v);
  }
}

// ------------------------------------------------------------ launchers --
hipError_t launch_solve(hipStream_t st, const SolveArgs &a) {
  dim3 grid((unsigned)((2 * a.g.npair + 255) / 256), 2);
  hipLaunchKernelGGL(k_solve_ve, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_mean(hipStream_t st, const MeanArgs &a) {
  hipLaunchKernelGGL(k_mean, dim3(1), dim3(256), sizeof(double) * 4 * a.g.Ny, st, a);
  return hipGetLastError();
}
hipError_t launch_uw(hipStream_t st, const UwArgs &a) {
  hipLaunchKernelGGL(k_uw, dim3((unsigned)((2 * a.g.npair + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_flowrate(hipStream_t st, const MeanArgs &a) {
  hipLaunchKernelGGL(k_flowrate, dim3(1), dim3(256), sizeof(double) * 3 * a.g.Ny, st, a);
  return hipGetLastError();
}
hipError_t launch_cfl_reduce(hipStream_t st, const double *part, int n, DevScalars *sc) {
  hipLaunchKernelGGL(k_cfl_reduce, dim3(1), dim3(256), 0, st, part, n, sc);
  return hipGetLastError();
}

}  // namespace chan
