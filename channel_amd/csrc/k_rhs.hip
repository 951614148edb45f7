// k_rhs.hip -- K4: right-hand sides of the v-eta formulation for every
// interior row and wavenumber pair (buildrhs, dnsdata.f90:611-673).
//
// One thread per (pair, chunk of rows) walks y with a five-row window of the
// six products and three velocity components held in registers, so each
// plane is read once.  For each row it forms rhsu/rhsv/rhsw from the compact
// d0/d1/d2 stencils of the products (the DD macro, :609-610), the explicit
// terms of the D2v and eta equations (:637-661), the implicit operators
// OS/SQ (:476-477) and the CN-RK3 time scheme (:486), writes the new rhs and
// replaces the stored explicit term.
#include "chan_internal.hpp"

namespace chan {

__device__ __forceinline__ cx<double> imul(double a, cx<double> X) { return cx<double>{-(a * X.y), a * X.x}; }
__device__ __forceinline__ cx<double> neg(cx<double> X) { return cx<double>{-X.x, -X.y}; }

template <typename R>
__global__ __launch_bounds__(256) void k_rhs(RhsArgs<R> a, const double *__restrict__ der,
                                             const DevScalars *__restrict__ sc) {
  // der and sc are separate restrict arguments so that their wave-uniform
  // values come through the scalar cache instead of vector registers.
  const Geom g = a.g;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npair) return;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  const double al = g.alfa0 * ix, be = g.beta0 * iz;
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const int y0 = 1 + (int)blockIdx.y * a.chunk;
  if (y0 > g.ny - 1) return;
  const int y1 = min(g.ny - 1, y0 + a.chunk - 1);
  const size_t ps = g.ps, pstride = (size_t)g.Ny * ps;
  const bool mean = (ix == 0 && iz == 0);
  const double meanpx = sc->meanpx, meanpz = sc->meanpz;

  // window slot j holds row iy-2+j; plane index of row r is r+1.  The stored
  // explicit terms of row iy are loaded one iteration early into nxo,
  // unconditionally (clamped row), so the loads overlap the previous row.
  const size_t rs = (size_t)(g.ny - 1) * ps;  // component stride of old/rhs
  cx<double> wp[6][5], wv[3][5], nxo[2];
  auto load_row = [&](int r, int j) {
    const size_t off = (size_t)(r + 1) * ps + p;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const cx<R> v = a.P[q * pstride + off];
      wp[q][j] = cx<double>{(double)v.x, (double)v.y};
    }
#pragma unroll
    for (int c = 0; c < 3; c++) wv[c][j] = a.V[c * pstride + off];
  };
#pragma unroll
  for (int j = 0; j < 4; j++) load_row(y0 - 2 + j, j);
  auto ldold = [&](size_t i) {
    const cx<R> o = a.old[i];
    return cx<double>{(double)o.x, (double)o.y};
  };
  nxo[0] = ldold((size_t)(y0 - 1) * ps + p);
  nxo[1] = ldold(rs + (size_t)(y0 - 1) * ps + p);

  for (int iy = y0; iy <= y1; iy++) {
    load_row(iy + 2, 4);
    const cx<double> old_eta = nxo[0], old_v = nxo[1];
    {
      const int on = min(iy, g.ny - 2);
      nxo[0] = ldold((size_t)on * ps + p);
      nxo[1] = ldold(rs + (size_t)on * ps + p);
    }
    const double *d = der + (size_t)(iy - 1) * 20;  // d0[5] d1[5] d2[5] d4[5]
    auto DD = [&](int f, int q) {
      cx<double> s = scl(d[f * 5 + 0], wp[q][0]);
#pragma unroll
      for (int j = 1; j < 5; j++) s = add(s, scl(d[f * 5 + j], wp[q][j]));
      return s;
    };
    // products: 0 uu, 1 vv, 2 ww, 3 uv, 4 vw, 5 uw
    const cx<double> DD0_6 = DD(0, 5), DD1_6 = DD(1, 5);
    const cx<double> rhsu = sub(sub(neg(imul(al, DD(0, 0))), DD(1, 3)), imul(be, DD0_6));
    const cx<double> rhsv = sub(sub(neg(imul(al, DD(0, 3))), DD(1, 1)), imul(be, DD(0, 4)));
    const cx<double> rhsw = sub(sub(neg(imul(al, DD0_6)), DD(1, 4)), imul(be, DD(0, 2)));
    const cx<double> expl =
        sub(add(imul(al, add(add(imul(al, DD(1, 0)), DD(2, 3)), imul(be, DD1_6))),
                imul(be, add(add(imul(al, DD1_6), DD(2, 4)), imul(be, DD(1, 2))))),
            scl(k2, rhsv));
    const size_t o = (size_t)(iy - 1) * ps + p;
    // D2v equation: unknown D2(v) - k2 D0(v), implicit OS v
    {
      cx<double> s2 = scl(d[10], wv[1][0]), s0 = scl(d[0], wv[1][0]);
#pragma unroll
      for (int j = 1; j < 5; j++) {
        s2 = add(s2, scl(d[10 + j], wv[1][j]));
        s0 = add(s0, scl(d[j], wv[1][j]));
      }
      const cx<double> unkn = sub(s2, scl(k2, s0));
      cx<double> impl = {0.0, 0.0};
#pragma unroll
      for (int j = 0; j < 5; j++) {
        const double os = a.ni * (d[15 + j] - 2.0 * k2 * d[10 + j] + k2 * k2 * d[j]);
        impl = add(impl, scl(os, wv[1][j]));
      }
      const cx<double> old = old_v;
      const cx<double> un = scl(a.ode0, unkn);
      const cx<double> r = sub(add(add(cx<double>{un.x / a.dt, un.y / a.dt}, impl), scl(a.ode1, expl)), scl(a.ode2, old));
      a.old[rs + o] = cx<R>{(R)expl.x, (R)expl.y};
      a.rhs[rs + o] = r;
    }
    // eta equation (or the mean flow Re u + i Re w in mode (0,0))
    {
      cx<double> ex, unkn, impl;
      if (mean) {
        ex = cx<double>{rhsu.x + meanpx, rhsw.x + meanpz};
        double a1 = d[0] * wv[0][0].x, a3 = d[0] * wv[2][0].x, b1 = d[10] * wv[0][0].x, b3 = d[10] * wv[2][0].x;
#pragma unroll
        for (int j = 1; j < 5; j++) {
          a1 += d[j] * wv[0][j].x;
          a3 += d[j] * wv[2][j].x;
          b1 += d[10 + j] * wv[0][j].x;
          b3 += d[10 + j] * wv[2][j].x;
        }
        unkn = cx<double>{a1, a3};
        impl = cx<double>{a.ni * b1, a.ni * b3};
      } else {
        ex = sub(imul(be, rhsu), imul(al, rhsw));
        cx<double> s1 = scl(d[0], wv[0][0]), s3 = scl(d[0], wv[2][0]);
#pragma unroll
        for (int j = 1; j < 5; j++) {
          s1 = add(s1, scl(d[j], wv[0][j]));
          s3 = add(s3, scl(d[j], wv[2][j]));
        }
        unkn = sub(imul(be, s1), imul(al, s3));
        impl = cx<double>{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 5; j++) {
          const double sq = a.ni * (d[10 + j] - k2 * d[j]);
          impl = add(impl, scl(sq, sub(imul(be, wv[0][j]), imul(al, wv[2][j]))));
        }
      }
      const cx<double> old = old_eta;
      const cx<double> un = scl(a.ode0, unkn);
      const cx<double> r = sub(add(add(cx<double>{un.x / a.dt, un.y / a.dt}, impl), scl(a.ode1, ex)), scl(a.ode2, old));
      a.old[o] = cx<R>{(R)ex.x, (R)ex.y};
      a.rhs[o] = r;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
      for (int q = 0; q < 6; q++) wp[q][j] = wp[q][j + 1];
#pragma unroll
      for (int c = 0; c < 3; c++) wv[c][j] = wv[c][j + 1];
    }
  }
}

template <typename R>
hipError_t launch_rhs(hipStream_t st, const RhsArgs<R> &a) {
  dim3 grid((unsigned)((a.g.npair + 255) / 256), (unsigned)((a.g.ny - 1 + a.chunk - 1) / a.chunk));
  hipLaunchKernelGGL((k_rhs<R>), grid, dim3(256), 0, st, a, a.der, a.sc);
  return hipGetLastError();
}

template hipError_t launch_rhs<double>(hipStream_t, const RhsArgs<double> &);
template hipError_t launch_rhs<float>(hipStream_t, const RhsArgs<float> &);

}  // namespace chan
