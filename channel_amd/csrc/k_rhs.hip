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
//
// With a body force, k_force_terms first forms the force terms of buildrhs
// (dnsdata.f90:643-659) from F = F0 + gain * V_src with the wall treatment
// of :616-628 (F zero on the wall rows, ghost rows that make the compact D4
// of the first and last interior rows vanish) and leaves them in the rhs
// buffer; the FORCE variant of k_rhs adds them to the explicit terms as it
// overwrites that buffer.  A five-row force window inside k_rhs itself would
// push it past the register file.
#include "chan_internal.hpp"

namespace chan {

__device__ __forceinline__ cx<double> imul(double a, cx<double> X) { return cx<double>{-(a * X.y), a * X.x}; }
__device__ __forceinline__ cx<double> neg(cx<double> X) { return cx<double>{-X.x, -X.y}; }

// ghost rows of the body force (dnsdata.f90:617-628): with F(0) = F(-1) = 0 the
// compact D4 of row 1 is d4(0)F(1) + d4(1)F(2) + d4(2)F(3), and
// F(-1) = -D4 / d4(-2); at the top F(ny+1) = -D4 / d4(2) of row ny-1.
__device__ __forceinline__ cx<double> ghost_lo(const double *d4, cx<double> f1, cx<double> f2, cx<double> f3) {
  const cx<double> s = add(add(scl(d4[2], f1), scl(d4[3], f2)), scl(d4[4], f3));
  return cx<double>{-s.x / d4[0], -s.y / d4[0]};
}
__device__ __forceinline__ cx<double> ghost_hi(const double *d4, cx<double> f3, cx<double> f2, cx<double> f1) {
  const cx<double> s = add(add(scl(d4[0], f3), scl(d4[1], f2)), scl(d4[2], f1));
  return cx<double>{-s.x / d4[4], -s.y / d4[4]};
}

// one-wave blocks: each lane walks a long row chunk, and smaller blocks spread
// the waves more evenly over the CUs (1.7% faster than 256-thread blocks in
// FP64 at 191x384x189)
constexpr int RHS_BLOCK = 64;

template <typename R, bool FORCE>
__global__ __launch_bounds__(256) void k_rhs(RhsArgs<R> a, const double *__restrict__ der,
                                             const DevScalars *__restrict__ sc) {
  // der and sc are separate restrict arguments so that their wave-uniform
  // values come through the scalar cache instead of vector registers.
  const Geom g = a.g;
  const int p = blockIdx.x * RHS_BLOCK + threadIdx.x;
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
  cx<double> wp[6][5], wv[3][5], nxo[2], nxf[2];
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
  // force terms of the eta and D2v equations, left in rhs by k_force_terms
  if constexpr (FORCE) {
    nxf[0] = a.rhs[(size_t)(y0 - 1) * ps + p];
    nxf[1] = a.rhs[rs + (size_t)(y0 - 1) * ps + p];
  }

  for (int iy = y0; iy <= y1; iy++) {
    load_row(iy + 2, 4);
    const cx<double> old_eta = nxo[0], old_v = nxo[1];
    cx<double> f_eta = {0.0, 0.0}, f_v = {0.0, 0.0};
    if constexpr (FORCE) {
      f_eta = nxf[0];
      f_v = nxf[1];
    }
    {
      const int on = min(iy, g.ny - 2);
      nxo[0] = ldold((size_t)on * ps + p);
      nxo[1] = ldold(rs + (size_t)on * ps + p);
      if constexpr (FORCE) {
        nxf[0] = a.rhs[(size_t)on * ps + p];
        nxf[1] = a.rhs[rs + (size_t)on * ps + p];
      }
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
    cx<double> expl =
        sub(add(imul(al, add(add(imul(al, DD(1, 0)), DD(2, 3)), imul(be, DD1_6))),
                imul(be, add(add(imul(al, DD1_6), DD(2, 4)), imul(be, DD(1, 2))))),
            scl(k2, rhsv));
    if constexpr (FORCE) expl = add(expl, f_v);
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
        if constexpr (FORCE) ex = add(ex, f_eta);
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
        if constexpr (FORCE) ex = add(ex, f_eta);
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

// F_c at interior row r (1 <= r <= ny-1) of pair p: the static part plus the
// feedback gain times the source velocity component.
__device__ __forceinline__ cx<double> force_at(const ForceArgs &a, int c, int s, int r, int p, int col) {
  const Geom &g = a.g;
  const size_t pstride = (size_t)g.Ny * g.ps, off = (size_t)(r + 1) * g.ps + p;
  cx<double> f = {0.0, 0.0};
  if (a.F0) f = a.F0[c * pstride + off];
  if (s >= 0) f = add(f, scl(a.gain[((size_t)c * g.Ny + (r + 1)) * (g.Nz + 1) + col], a.V[s * pstride + off]));
  return f;
}

// F of row r with the wall rows of buildrhs: zero on rows 0 and ny, ghost
// rows -1 and ny+1 from the compact D4 of rows 1 and ny-1.
__device__ __forceinline__ cx<double> force_row(const ForceArgs &a, const double *der, int c, int s, int r, int p,
                                                int col) {
  const int ny = a.g.ny;
  if (r >= 1 && r <= ny - 1) return force_at(a, c, s, r, p, col);
  if (r == -1) return ghost_lo(der + 15, force_at(a, c, s, 1, p, col), force_at(a, c, s, 2, p, col), force_at(a, c, s, 3, p, col));
  if (r == ny + 1)
    return ghost_hi(der + (size_t)(ny - 2) * 20 + 15, force_at(a, c, s, ny - 3, p, col), force_at(a, c, s, ny - 2, p, col),
                    force_at(a, c, s, ny - 1, p, col));
  return cx<double>{0.0, 0.0};
}

// The force terms of buildrhs for rows y0..y1 of pair p (one thread per pair
// and chunk of rows, five-row window of F):
//   out[0] (eta)  i kz D0(F_u) - i kx D0(F_w), or (D0 Re F_u, D0 Re F_w) in
//                 the mode (0,0) (rD0, dnsdata.f90:650-652, 657-659);
//   out[1] (D2v)  -k2 D0(F_v) - i kx D1(F_u) - i kz D1(F_w) (:643-645).
__global__ __launch_bounds__(256) void k_force_terms(ForceArgs a, const double *__restrict__ der) {
  const Geom g = a.g;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npair) return;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  const bool mean = (ix == 0 && iz == 0);
  const int col = mean ? g.Nz : iz + g.nz;
  const double al = g.alfa0 * ix, be = g.beta0 * iz;
  const double k2 = (g.alfa0 * ix) * (g.alfa0 * ix) + (g.beta0 * iz) * (g.beta0 * iz);
  const int y0 = 1 + (int)blockIdx.y * a.chunk;
  if (y0 > g.ny - 1) return;
  const int y1 = min(g.ny - 1, y0 + a.chunk - 1);
  const size_t ps = g.ps, rs = (size_t)(g.ny - 1) * ps;
  cx<double> w[3][5];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int j = 0; j < 4; j++) w[c][j] = force_row(a, der, c, a.src[c], y0 - 2 + j, p, col);
  for (int iy = y0; iy <= y1; iy++) {
#pragma unroll
    for (int c = 0; c < 3; c++) w[c][4] = force_row(a, der, c, a.src[c], iy + 2, p, col);
    const double *d = der + (size_t)(iy - 1) * 20;
    auto S = [&](int f, int c) {
      cx<double> s = scl(d[f * 5 + 0], w[c][0]);
#pragma unroll
      for (int j = 1; j < 5; j++) s = add(s, scl(d[f * 5 + j], w[c][j]));
      return s;
    };
    const cx<double> d0u = S(0, 0), d0w = S(0, 2);
    const cx<double> te = mean ? cx<double>{d0u.x, d0w.x} : sub(imul(be, d0u), imul(al, d0w));
    const cx<double> tv = sub(sub(neg(scl(k2, S(0, 1))), imul(al, S(1, 0))), imul(be, S(1, 2)));
    a.out[(size_t)(iy - 1) * ps + p] = te;
    a.out[rs + (size_t)(iy - 1) * ps + p] = tv;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int j = 0; j < 4; j++) w[c][j] = w[c][j + 1];
  }
}

// The body force that buildrhs applies to the resident field, wall rows and
// ghost rows included, in the device layout (one thread per pair and
// component).
__global__ __launch_bounds__(256) void k_force(ForceArgs a, const double *__restrict__ der) {
  const Geom g = a.g;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= g.npair) return;
  const int c = blockIdx.y;
  const int ix = p / g.Nz, iz = p % g.Nz - g.nz;
  const int col = (ix == 0 && iz == 0) ? g.Nz : iz + g.nz;
  cx<double> *o = a.out + c * (size_t)g.Ny * g.ps + p;
  for (int r = -1; r <= g.ny + 1; r++) o[(size_t)(r + 1) * g.ps] = force_row(a, der, c, a.src[c], r, p, col);
}

// Blocks of `block` pairs along x of the grid, rounded up to a multiple of 8:
// blocks (x, y) and (x, y+1) then sit a multiple of 8 apart in dispatch order,
// so round-robin dispatch puts the neighbouring row chunks of one pair block
// on one XCD and the halo rows that both read come from its L2 (k_rhs 1.3%
// faster in FP64 at 191x384x189).  The extra blocks exit at once.
static unsigned pair_blocks(const Geom &g, int block) {
  return (unsigned)(((g.npair + block - 1) / block + 7) / 8 * 8);
}

hipError_t launch_force(hipStream_t st, const ForceArgs &a) {
  dim3 grid((unsigned)((a.g.npair + 255) / 256), 3);
  hipLaunchKernelGGL(k_force, grid, dim3(256), 0, st, a, a.der);
  return hipGetLastError();
}

hipError_t launch_force_terms(hipStream_t st, const ForceArgs &a) {
  dim3 grid(pair_blocks(a.g, 256), (unsigned)((a.g.ny - 1 + a.chunk - 1) / a.chunk));
  hipLaunchKernelGGL(k_force_terms, grid, dim3(256), 0, st, a, a.der);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_rhs(hipStream_t st, const RhsArgs<R> &a) {
  dim3 grid(pair_blocks(a.g, RHS_BLOCK), (unsigned)((a.g.ny - 1 + a.chunk - 1) / a.chunk));
  if (a.force)
    hipLaunchKernelGGL((k_rhs<R, true>), grid, dim3(RHS_BLOCK), 0, st, a, a.der, a.sc);
  else
    hipLaunchKernelGGL((k_rhs<R, false>), grid, dim3(RHS_BLOCK), 0, st, a, a.der, a.sc);
  return hipGetLastError();
}

template hipError_t launch_rhs<double>(hipStream_t, const RhsArgs<double> &);
template hipError_t launch_rhs<float>(hipStream_t, const RhsArgs<float> &);

}  // namespace chan
