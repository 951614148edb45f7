// k_solve.hip -- K5: the implicit viscous solve of one RK3 substep
// (linsolve, linsolve_blocking.inc:3-107) as three kernels:
//
//   k_solve_ve : one thread per (pair, system), system 0 = eta, 1 = v.  Builds
//                the pentadiagonal rows lambda*D - nu*L of etamat / D2vmat,
//                folds the wall conditions into rows 1,2 and ny-2,ny-1
//                (applybc_0 / applybc_n, dnsdata.f90:458-472), factorises
//                bottom-up with reciprocal pivots fused with the backward
//                sweep (LU5decompStep + LeftLU5divStep1,
//                rbparmat_blocking.f90:20-77), then runs the forward sweep
//                (LeftLU5divStep2, :82-100) and recovers wall and ghost
//                values (linsolve_blocking.inc:50-61).  The (0,0) eta thread
//                also solves the unit-forcing response ucor (:65-68).
//   k_mean     : the (0,0) mode: u/w split, ucor ghosts, flow rates by the
//                three-point integral (yintegr, dnsdata.f90:312-324),
//                constant-flow correction or CPI pressure gradient
//                (linsolve_blocking.inc:62-97).
//   k_uw       : every other pair: dv/dy by the compact first derivative
//                (COMPLEXderiv + LeftLU5divStep2 with D0, dnsdata.f90:337-373)
//                and u = (i kx v' - i kz eta)/k^2, w = (i kz v' + i kx eta)/k^2.
//
// The factorisation order (bottom-up UL, no pivoting) is the reference's, so
// FP64 results agree to rounding.
#include "chan_internal.hpp"

namespace chan {

__device__ __forceinline__ cx<double> imul2(double a, cx<double> X) { return cx<double>{-(a * X.y), a * X.x}; }

// ------------------------------------------------------------ k_solve_ve --
__global__ __launch_bounds__(256) void k_solve_ve(SolveArgs a) {
  const Geom g = a.g;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npair) return;
  const int sys = blockIdx.y;  // 0: eta (component 1), 1: v (component 2)
  const int ny = g.ny;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const bool mean = (ix == 0 && iz == 0);
  const bool do_ucor = mean && sys == 0;
  const size_t ps = g.ps;
  cx<double> *col = a.V + (size_t)sys * g.Ny * ps + p;  // row iy at col[(iy+1)*ps]
  const cx<double> *rc = a.rhs + (size_t)sys * (ny - 1) * ps + p;
  double2 *Lc = a.L + (size_t)sys * (ny - 1) * ps + p;
  const double *bc0 = a.bc[sys ? BC_V0 : BC_ETA0], *bc0m1 = a.bc[sys ? BC_V0M1 : BC_ETA0M1];
  const double *bcn = a.bc[sys ? BC_VN : BC_ETAN], *bcnp1 = a.bc[sys ? BC_VNP1 : BC_ETANP1];
  // wall value of eta: (Re u - Im w) + i (Im u + Re w) with u = u0/uN, w = 0
  // in the (0,0) mode, zero otherwise; v and dv/dy vanish at the walls.
  const cx<double> eta0 = (mean && sys == 0) ? cx<double>{a.u0, 0.0} : cx<double>{0.0, 0.0};
  const cx<double> etan = (mean && sys == 0) ? cx<double>{a.uN, 0.0} : cx<double>{0.0, 0.0};

  // ---- bottom-up: rows, wall folding, factorisation, backward sweep
  double l1m2 = 0.0, l1m1 = 0.0, l2m2 = 0.0, l2m1 = 0.0;  // lower factors of rows iy+1, iy+2
  cx<double> x1 = {0.0, 0.0}, x2 = {0.0, 0.0}, u1 = {0.0, 0.0}, u2 = {0.0, 0.0};
  for (int iy = ny - 1; iy >= 1; iy--) {
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
    cx<double> b = rc[(size_t)(iy - 1) * ps];
    if (iy == 1) {  // applybc_0 rows 1: ghost then wall column
      const double e = A[0];
      for (int j = 1; j <= 4; j++) A[j] = A[j] - e * bc0m1[j] / bc0m1[0];
      const double f = A[1];
      for (int j = 2; j <= 4; j++) A[j] = A[j] - f * bc0[j] / bc0[1];
      if (sys == 0) b = sub(b, cx<double>{A[1] * eta0.x / bc0[1], A[1] * eta0.y / bc0[1]});
    } else if (iy == 2) {
      if (sys == 0) b = sub(b, cx<double>{A[0] * eta0.x / bc0[1], A[0] * eta0.y / bc0[1]});
      const double e = A[0];
      for (int j = 1; j <= 3; j++) A[j] = A[j] - e * bc0[j + 1] / bc0[1];
    }
    if (iy == ny - 1) {  // applybc_n rows ny-1
      const double e = A[4];
      for (int j = 0; j <= 3; j++) A[j] = A[j] - e * bcnp1[j] / bcnp1[4];
      const double f = A[3];
      for (int j = 0; j <= 2; j++) A[j] = A[j] - f * bcn[j] / bcn[3];
      if (sys == 0) b = sub(b, cx<double>{A[3] * etan.x / bcn[3], A[3] * etan.y / bcn[3]});
      A[3] = 0.0;
      A[4] = 0.0;
    } else if (iy == ny - 2) {
      if (sys == 0) b = sub(b, cx<double>{A[4] * etan.x / bcn[3], A[4] * etan.y / bcn[3]});
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
    const cx<double> s = add(scl(p1, x1), scl(p2, x2));
    const cx<double> xi = scl(r, sub(b, s));
    col[(size_t)(iy + 1) * ps] = xi;
    Lc[(size_t)(iy - 1) * ps] = make_double2(iy == 1 ? 0.0 : (iy == 2 ? 0.0 : A[0]), iy == 1 ? 0.0 : A[1]);
    if (do_ucor) {
      const cx<double> su = add(scl(p1, u1), scl(p2, u2));
      const cx<double> ui = scl(r, sub(cx<double>{1.0, 0.0}, su));
      a.ucor[iy + 1] = ui;
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

  // ---- top-down: forward sweep, keep rows 1..3 and ny-3..ny-1
  cx<double> b1 = {0.0, 0.0}, b2 = {0.0, 0.0}, b3 = {0.0, 0.0}, ub1 = {0.0, 0.0}, ub2 = {0.0, 0.0};
  cx<double> bot[3];
  for (int iy = 1; iy <= ny - 1; iy++) {
    const double2 L = Lc[(size_t)(iy - 1) * ps];
    const cx<double> x = col[(size_t)(iy + 1) * ps];
    const cx<double> bi = sub(x, add(scl(L.x, b2), scl(L.y, b1)));
    col[(size_t)(iy + 1) * ps] = bi;
    if (iy <= 3) bot[iy - 1] = bi;
    b3 = b2;
    b2 = b1;
    b1 = bi;
    if (do_ucor) {
      const cx<double> ux = a.ucor[iy + 1];
      const cx<double> ui = sub(ux, add(scl(L.x, ub2), scl(L.y, ub1)));
      a.ucor[iy + 1] = ui;
      ub2 = ub1;
      ub1 = ui;
    }
  }
  // wall and ghost rows
  const cx<double> t3 = b3, t2 = b2, t1 = b1;  // rows ny-3, ny-2, ny-1
  {
    const cx<double> s0 = add(add(scl(bc0[2], bot[0]), scl(bc0[3], bot[1])), scl(bc0[4], bot[2]));
    const cx<double> v0 = sub(eta0, s0);
    const cx<double> w0 = cx<double>{v0.x / bc0[1], v0.y / bc0[1]};
    const cx<double> s1 = add(add(add(scl(bc0m1[1], w0), scl(bc0m1[2], bot[0])), scl(bc0m1[3], bot[1])), scl(bc0m1[4], bot[2]));
    cx<double> wm1;
    if (sys)
      wm1 = cx<double>{(0.0 - s1.x) / bc0m1[0], (0.0 - s1.y) / bc0m1[0]};
    else
      wm1 = cx<double>{-(s1.x / bc0m1[0]), -(s1.y / bc0m1[0])};
    col[ps] = w0;
    col[0] = wm1;
  }
  {
    const cx<double> s0 = add(add(scl(bcn[0], t3), scl(bcn[1], t2)), scl(bcn[2], t1));
    const cx<double> v0 = sub(etan, s0);
    const cx<double> wn = cx<double>{v0.x / bcn[3], v0.y / bcn[3]};
    const cx<double> s1 = add(add(add(scl(bcnp1[0], t3), scl(bcnp1[1], t2)), scl(bcnp1[2], t1)), scl(bcnp1[3], wn));
    cx<double> wp1;
    if (sys)
      wp1 = cx<double>{(0.0 - s1.x) / bcnp1[4], (0.0 - s1.y) / bcnp1[4]};
    else
      wp1 = cx<double>{-(s1.x / bcnp1[4]), -(s1.y / bcnp1[4])};
    col[(size_t)(ny + 1) * ps] = wn;
    col[(size_t)(ny + 2) * ps] = wp1;
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
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npair) return;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  if (ix == 0 && iz == 0) return;
  const int ny = g.ny;
  const double al = g.alfa0 * ix, be = g.beta0 * iz;
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const size_t ps = g.ps, pstride = (size_t)g.Ny * ps;
  cx<double> *c1 = a.V + p, *c2 = a.V + pstride + p, *c3 = a.V + 2 * pstride + p;
#define ROW(c, iy) (c)[(size_t)((iy) + 1) * ps]
  // one-sided wall and ghost derivatives
  cx<double> vb[5], vt[5];
#pragma unroll
  for (int j = 0; j < 5; j++) {
    vb[j] = ROW(c2, -1 + j);
    vt[j] = ROW(c2, ny - 3 + j);
  }
  auto dot5 = [](const double *w, const cx<double> *v) {
    cx<double> s = scl(w[0], v[0]);
#pragma unroll
    for (int j = 1; j < 5; j++) s = add(s, scl(w[j], v[j]));
    return s;
  };
  const cx<double> f10 = dot5(a.bnd[B_D140], vb), f1m1 = dot5(a.bnd[B_D14M1], vb);
  const cx<double> f1n = dot5(a.bnd[B_D14N], vt), f1np1 = dot5(a.bnd[B_D14NP1], vt);
  // bottom-up: interior D1 v with boundary fix-ups and the D0 backward sweep
  cx<double> w[5];
#pragma unroll
  for (int j = 0; j < 5; j++) w[j] = vt[j];  // rows ny-3..ny+1 = window of iy = ny-1
  cx<double> x1 = f1n, x2 = f1np1;
  for (int iy = ny - 1; iy >= 1; iy--) {
    const double *d = a.der + (size_t)(iy - 1) * 20;
    cx<double> f = scl(d[5], w[0]);
#pragma unroll
    for (int j = 1; j < 5; j++) f = add(f, scl(d[5 + j], w[j]));
    if (iy == ny - 1) f = sub(f, add(scl(d[3], f1n), scl(d[4], f1np1)));
    if (iy == ny - 2) f = sub(f, scl(d[4], f1n));
    if (iy == 2) f = sub(f, scl(d[0], f10));
    if (iy == 1) f = sub(f, add(scl(d[1], f10), scl(d[0], f1m1)));
    const double *D = a.d0fac + (size_t)(iy - 1) * 5;
    const cx<double> xi = scl(D[2], sub(f, add(scl(D[3], x1), scl(D[4], x2))));
    ROW(c3, iy) = xi;
    x2 = x1;
    x1 = xi;
    if (iy > 1) {
#pragma unroll
      for (int j = 4; j > 0; j--) w[j] = w[j - 1];
      w[0] = ROW(c2, iy - 3);
    }
  }
  // top-down: D0 forward sweep, then u and w on every row
  cx<double> b1 = f10, b2 = f1m1;
  for (int iy = -1; iy <= ny + 1; iy++) {
    cx<double> vp;
    if (iy == -1)
      vp = f1m1;
    else if (iy == 0)
      vp = f10;
    else if (iy <= ny - 1) {
      const double *D = a.d0fac + (size_t)(iy - 1) * 5;
      vp = sub(ROW(c3, iy), add(scl(D[0], b2), scl(D[1], b1)));
      b2 = b1;
      b1 = vp;
    } else
      vp = iy == ny ? f1n : f1np1;
    const cx<double> eta = ROW(c1, iy);
    const cx<double> tu = sub(imul2(al, vp), imul2(be, eta));
    const cx<double> tw = add(imul2(be, vp), imul2(al, eta));
    ROW(c1, iy) = cx<double>{tu.x / k2, tu.y / k2};
    ROW(c3, iy) = cx<double>{tw.x / k2, tw.y / k2};
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
  dim3 grid((unsigned)((a.g.npair + 255) / 256), 2);
  hipLaunchKernelGGL(k_solve_ve, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_mean(hipStream_t st, const MeanArgs &a) {
  hipLaunchKernelGGL(k_mean, dim3(1), dim3(256), sizeof(double) * 4 * a.g.Ny, st, a);
  return hipGetLastError();
}
hipError_t launch_uw(hipStream_t st, const UwArgs &a) {
  hipLaunchKernelGGL(k_uw, dim3((unsigned)((a.g.npair + 255) / 256)), dim3(256), 0, st, a);
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
