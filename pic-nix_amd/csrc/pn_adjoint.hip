// pn_adjoint.hip -- the adjoint form (include/picnix_adjoint.h): a 2D periodic step
// at order 1 with the hot path's operators, particles in a fixed order, and its
// hand-written discrete adjoint, in FP64 or FP32 (every kernel is templated on the
// real type R of the particles, fields and co-states; the loss and the gradient
// are reduced in FP64).
//
// Forward step (the operator order of PicApplication::push_openmp,
// pic/pic_application.cpp:219-292, on one periodic chunk):
//   B1 = B + K_B(E)            push_bfd(dt/2)
//   e, b = G(x) (E, B1)        MC gather, scaled by (q/m) dt/2
//   u' = boris(u, e, b)
//   x' = x + u' dt / gamma'
//   J = D(x, x')               Esirkepov
//   B2 = B1 + K_B(E)           push_bfd(dt/2)
//   E' = E + K_E(B2) - dt J    push_efd(dt)
//   x' wrapped into the periodic domain
// Adjoint step, for the co-state (lE', lB2, lx', lu') after the step:
//   lJ = -dt lE'
//   per particle: lx = lx' + (dD/dx)^T lJ + (dG/dx)^T le,  lu = (dboris/du)^T lu*,
//     with lu* = lu' + (df/du')^T (lx' + (dD/dx')^T lJ), f(u) = u dt / gamma(u),
//     le, lb = (dboris/d(e,b))^T lu*; the scatter G^T(le, lb) -> lEg, lB1g
//   lB2t = lB2 + K_E^T lE';  lB1t = lB2t + lB1g
//   lE = lE' + lEg + K_B^T lB2t + K_B^T lB1t;  lB = lB1t
// The per-particle Jacobians (gather weights in x, Boris in (u, e, b), current
// in (x, x')) are evaluated with forward-mode dual numbers over the same code
// as the forward, so their values are the forward's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/picnix_adjoint.h"
#include "pn_sim.hpp"

namespace picnix {

namespace {

// ---------------------------------------------------------------------------
// dual numbers: value + N directional derivatives
// ---------------------------------------------------------------------------
template <typename R>
struct Id {
  using type = R; // keeps a scalar argument out of template deduction
};

template <int N, typename R = double>
struct Dual {
  R v;
  R d[N];
  __device__ Dual() {}
  __device__ Dual(R x) : v(x)
  {
#pragma unroll
    for (int k = 0; k < N; k++)
      d[k] = R(0);
  }
  __device__ static Dual var(R x, int k)
  {
    Dual r(x);
    r.d[k] = R(1);
    return r;
  }
  __device__ Dual& operator+=(const Dual& b)
  {
    v += b.v;
#pragma unroll
    for (int k = 0; k < N; k++)
      d[k] += b.d[k];
    return *this;
  }
  __device__ Dual& operator-=(const Dual& b)
  {
    v -= b.v;
#pragma unroll
    for (int k = 0; k < N; k++)
      d[k] -= b.d[k];
    return *this;
  }
  __device__ Dual& operator*=(const Dual& b)
  {
#pragma unroll
    for (int k = 0; k < N; k++)
      d[k] = d[k] * b.v + v * b.d[k];
    v *= b.v;
    return *this;
  }
};

template <int N, typename R>
__device__ inline Dual<N, R> operator+(Dual<N, R> a, const Dual<N, R>& b)
{
  return a += b;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator-(Dual<N, R> a, const Dual<N, R>& b)
{
  return a -= b;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator-(const Dual<N, R>& a)
{
  Dual<N, R> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; k++)
    r.d[k] = -a.d[k];
  return r;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator*(Dual<N, R> a, const Dual<N, R>& b)
{
  return a *= b;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator/(const Dual<N, R>& a, const Dual<N, R>& b)
{
  Dual<N, R> r;
  const R    inv = R(1) / b.v;
  r.v            = a.v / b.v;
#pragma unroll
  for (int k = 0; k < N; k++)
    r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
  return r;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator+(const Dual<N, R>& a, typename Id<R>::type b)
{
  return a + Dual<N, R>(b);
}
template <int N, typename R>
__device__ inline Dual<N, R> operator+(typename Id<R>::type a, const Dual<N, R>& b)
{
  return Dual<N, R>(a) + b;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator-(const Dual<N, R>& a, typename Id<R>::type b)
{
  return a - Dual<N, R>(b);
}
template <int N, typename R>
__device__ inline Dual<N, R> operator-(typename Id<R>::type a, const Dual<N, R>& b)
{
  return Dual<N, R>(a) - b;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator*(const Dual<N, R>& a, typename Id<R>::type b)
{
  Dual<N, R> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; k++)
    r.d[k] = a.d[k] * b;
  return r;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator*(typename Id<R>::type a, const Dual<N, R>& b)
{
  Dual<N, R> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < N; k++)
    r.d[k] = a * b.d[k];
  return r;
}
template <int N, typename R>
__device__ inline Dual<N, R> operator/(const Dual<N, R>& a, typename Id<R>::type b)
{
  return a / Dual<N, R>(b);
}
template <int N, typename R>
__device__ inline Dual<N, R> sqrt(const Dual<N, R>& a)
{
  Dual<N, R> r;
  r.v         = ::sqrt(a.v);
  const R inv = R(0.5) / r.v;
#pragma unroll
  for (int k = 0; k < N; k++)
    r.d[k] = a.d[k] * inv;
  return r;
}

__device__ inline double val(double x) { return x; }
__device__ inline float  val(float x) { return x; }
template <int N, typename R>
__device__ inline R val(const Dual<N, R>& x)
{
  return x.v;
}

// ---------------------------------------------------------------------------
// geometry and indexing (fields [3][Ny][Nx], periodic)
// ---------------------------------------------------------------------------
struct AGeom {
  int    nx, ny;
  long   ncell;
  double dx, dy, rdx, rdy, dt, cc, rc, lx, ly;
  double q, qmdt; // macro-particle charge, (q/m) dt / 2
  long   np;
};

__device__ __forceinline__ long cell(const AGeom& g, int iy, int ix)
{
  iy = iy % g.ny;
  iy = iy < 0 ? iy + g.ny : iy;
  ix = ix % g.nx;
  ix = ix < 0 ? ix + g.nx : ix;
  return (long)iy * g.nx + ix;
}

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// Yee (push_bfd_2d / push_efd_2d, pic/engine/maxwell.hpp:231-348) and transposes
// ---------------------------------------------------------------------------
// B += K_B(E)
template <typename R>
__global__ void __launch_bounds__(256) k_bfd(AGeom g, const R* __restrict__ E,
                                             R* __restrict__ B, R cflx, R cfly)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ncell)
    return;
  const int  iy = (int)(i / g.nx), ix = (int)(i % g.nx);
  const long v = g.ncell, jx = cell(g, iy, ix - 1), jy = cell(g, iy - 1, ix);
  const R *Ex = E, *Ey = E + v, *Ez = E + 2 * v;
  B[i] += (-cfly) * (Ez[i] - Ez[jy]);
  B[v + i] += (+cflx) * (Ez[i] - Ez[jx]);
  B[2 * v + i] += (-cflx) * (Ey[i] - Ey[jx]) + (+cfly) * (Ex[i] - Ex[jy]);
}

// E += K_E(B) - dt J
template <typename R, typename A>
__global__ void __launch_bounds__(256) k_efd(AGeom g, R* __restrict__ E,
                                             const R* __restrict__ B,
                                             const A* __restrict__ J, R cflx,
                                             R cfly)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ncell)
    return;
  const int  iy = (int)(i / g.nx), ix = (int)(i % g.nx);
  const long v = g.ncell, px = cell(g, iy, ix + 1), py = cell(g, iy + 1, ix);
  const R *Bx = B, *By = B + v, *Bz = B + 2 * v;
  const R  dt = R(g.dt);
  E[i] += (+cfly) * (Bz[py] - Bz[i]) - dt * R(J[i]);
  E[v + i] += (-cflx) * (Bz[px] - Bz[i]) - dt * R(J[v + i]);
  E[2 * v + i] += (+cflx) * (By[px] - By[i]) + (-cfly) * (Bx[py] - Bx[i]) - dt * R(J[2 * v + i]);
}

// lE += K_B^T lB
template <typename R>
__global__ void __launch_bounds__(256) k_bfd_T(AGeom g, const R* __restrict__ lB,
                                               R* __restrict__ lE, R cflx, R cfly)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ncell)
    return;
  const int  iy = (int)(i / g.nx), ix = (int)(i % g.nx);
  const long v = g.ncell, px = cell(g, iy, ix + 1), py = cell(g, iy + 1, ix);
  const R *lBx = lB, *lBy = lB + v, *lBz = lB + 2 * v;
  lE[i] += cfly * (lBz[i] - lBz[py]);
  lE[v + i] += -cflx * (lBz[i] - lBz[px]);
  lE[2 * v + i] += -cfly * (lBx[i] - lBx[py]) + cflx * (lBy[i] - lBy[px]);
}

// lB += K_E^T lE
template <typename R>
__global__ void __launch_bounds__(256) k_efd_T(AGeom g, const R* __restrict__ lE,
                                               R* __restrict__ lB, R cflx, R cfly)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ncell)
    return;
  const int  iy = (int)(i / g.nx), ix = (int)(i % g.nx);
  const long v = g.ncell, mx = cell(g, iy, ix - 1), my = cell(g, iy - 1, ix);
  const R *lEx = lE, *lEy = lE + v, *lEz = lE + 2 * v;
  lB[i] += -cfly * (lEz[my] - lEz[i]);
  lB[v + i] += cflx * (lEz[mx] - lEz[i]);
  lB[2 * v + i] += cfly * (lEx[my] - lEx[i]) - cflx * (lEy[mx] - lEy[i]);
}

template <typename R, typename A>
__global__ void __launch_bounds__(256) k_axpy(long n, R a, const A* __restrict__ x,
                                              R* __restrict__ y)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    y[i] += a * R(x[i]);
}

template <typename R>
__global__ void __launch_bounds__(256) k_set(long n, R a, R* __restrict__ y)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    y[i] = a;
}

template <typename R>
__global__ void __launch_bounds__(256) k_copy(long n, const R* __restrict__ x,
                                              R* __restrict__ y)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    y[i] = x[i];
}

// ---------------------------------------------------------------------------
// per-particle pieces, written once for double and for dual numbers
// ---------------------------------------------------------------------------
// MC order-1 gather weights (velocity.hpp weights2d): integer (cell-centre) and
// half (face) grids; the first node of each stencil
// x is measured from the lower face of cell cb (cb = 0: the domain's origin; the
// FP32 form passes each particle's own cell, so that x stays small)
template <typename T>
__device__ __forceinline__ void weights1(const AGeom& g, T x, double d, double rd, T wi[2],
                                         T wh[2], int& i0, int& h0, int cb = 0)
{
  const double imin = 0.5 * d; // xmin + d/2 (odd order); xmin = 0
  i0                = digitize((double)val(x), imin, rd) + cb;
  h0                = digitize((double)val(x), 0.0, rd) + cb;
  shape_mc<1>(x, T(0.5 * d + static_cast<double>(i0 - cb) * d), T(rd), wi);
  shape_mc<1>(x, T(0.0 + static_cast<double>(h0 - cb) * d), T(rd), wh);
}

// sum_{jy} sum_{jx} F[y0 + jy][x0 + jx] wx[jx] wy[jy] (x innermost), times (q/m) dt/2
template <typename T, typename R>
__device__ __forceinline__ T interp2(const AGeom& g, const R* __restrict__ F, int y0, int x0,
                                     const T wy[2], const T wx[2])
{
  T ry(0.0);
#pragma unroll
  for (int jy = 0; jy < 2; jy++) {
    T rx(0.0);
#pragma unroll
    for (int jx = 0; jx < 2; jx++)
      rx += T(F[cell(g, y0 + jy, x0 + jx)]) * wx[jx];
    ry += rx * wy[jy];
  }
  return ry * T(g.qmdt);
}

// the six gathered components (Ex, Ey, Ez, Bx, By, Bz) at (x, y)
template <typename T, typename R>
__device__ __forceinline__ void gather(const AGeom& g, const R* __restrict__ E,
                                       const R* __restrict__ B, T x, T y, T eb[6], int cbx = 0,
                                       int cby = 0)
{
  T   wix[2], whx[2], wiy[2], why[2];
  int ix0, hx0, iy0, hy0;
  weights1(g, x, g.dx, g.rdx, wix, whx, ix0, hx0, cbx);
  weights1(g, y, g.dy, g.rdy, wiy, why, iy0, hy0, cby);
  const long v = g.ncell;
  eb[0]        = interp2(g, E, iy0, hx0, wiy, whx);
  eb[1]        = interp2(g, E + v, hy0, ix0, why, wix);
  eb[2]        = interp2(g, E + 2 * v, iy0, ix0, wiy, wix);
  eb[3]        = interp2(g, B, hy0, ix0, why, wix);
  eb[4]        = interp2(g, B + v, iy0, hx0, wiy, whx);
  eb[5]        = interp2(g, B + 2 * v, hy0, hx0, why, whx);
}

// Esirkepov order 1 in 2D (nix/esirkepov.hpp:76-143 + shift_weights :239-270,
// pic/engine/current.hpp:275-324); add(iy, ix, comp, value), comp 0..2 = Jx, Jy, Jz.
// Positions are measured from the lower faces of cell (cbx, cby), as in weights1.
template <typename T, typename Add>
__device__ __forceinline__ void esirkepov2(const AGeom& g, T x0, T y0, T z0, T x1, T y1, T z1,
                                           Add add, int cbx = 0, int cby = 0)
{
  constexpr int N = 4;
  const double  A = 1.0 / 2, Bc = 1.0 / 3;
  T             s0x[N], s0y[N], s1x[N], s1y[N];
#pragma unroll
  for (int l = 0; l < N; l++)
    s0x[l] = s0y[l] = s1x[l] = s1y[l] = T(0.0);
  const double ximin = 0.5 * g.dx, yimin = 0.5 * g.dy;
  const int    ix0   = digitize((double)val(x0), ximin, g.rdx);
  const int    iy0   = digitize((double)val(y0), yimin, g.rdy);
  shape_mc<1>(x0, T(0.5 * g.dx + static_cast<double>(ix0) * g.dx), T(g.rdx), &s0x[1]);
  shape_mc<1>(y0, T(0.5 * g.dy + static_cast<double>(iy0) * g.dy), T(g.rdy), &s0y[1]);
  const int ix1 = digitize((double)val(x1), ximin, g.rdx);
  const int iy1 = digitize((double)val(y1), yimin, g.rdy);
  shape_mc<1>(x1, T(0.5 * g.dx + static_cast<double>(ix1) * g.dx), T(g.rdx), &s1x[1]);
  shape_mc<1>(y1, T(0.5 * g.dy + static_cast<double>(iy1) * g.dy), T(g.rdy), &s1y[1]);
  if (ix1 - ix0 < 0) {
#pragma unroll
    for (int ii = 0; ii < 3; ii++)
      s1x[ii] = s1x[ii + 1];
  } else if (ix1 - ix0 > 0) {
#pragma unroll
    for (int ii = 3; ii > 0; ii--)
      s1x[ii] = s1x[ii - 1];
  }
  if (iy1 - iy0 < 0) {
#pragma unroll
    for (int ii = 0; ii < 3; ii++)
      s1y[ii] = s1y[ii + 1];
  } else if (iy1 - iy0 > 0) {
#pragma unroll
    for (int ii = 3; ii > 0; ii--)
      s1y[ii] = s1y[ii - 1];
  }
  const int ox = ix0 - 1 + cbx, oy = iy0 - 1 + cby;
  const T   vz = (z1 - z0) / T(g.dt);
#pragma unroll
  for (int l = 0; l < N; l++) {
    s1x[l] -= s0x[l];
    s1y[l] -= s0y[l];
  }
  const T qdxdt(g.q * (g.dx / g.dt));
  const T qdydt(g.q * (g.dy / g.dt));
  const T qvz = T(g.q) * vz;
#pragma unroll
  for (int jy = 0; jy < N; jy++) {
    T ww(0.0);
    T wx = -(s0y[jy] + T(A) * s1y[jy]) * qdxdt;
#pragma unroll
    for (int jx = 0; jx < N - 1; jx++) {
      ww += s1x[jx] * wx;
      add(oy + jy, ox + jx + 1, 0, ww);
    }
  }
#pragma unroll
  for (int jx = 0; jx < N; jx++) {
    T ww(0.0);
    T wy = -(s0x[jx] + T(A) * s1x[jx]) * qdydt;
#pragma unroll
    for (int jy = 0; jy < N - 1; jy++) {
      ww += s1y[jy] * wy;
      add(oy + jy + 1, ox + jx, 1, ww);
    }
  }
#pragma unroll
  for (int jy = 0; jy < N; jy++)
#pragma unroll
    for (int jx = 0; jx < N; jx++) {
      T wz = ((T(1.0) * s0x[jx] + T(A) * s1x[jx]) * s0y[jy] +
              (T(A) * s0x[jx] + T(Bc) * s1x[jx]) * s1y[jy]) *
             qvz;
      add(oy + jy, ox + jx, 2, wz);
    }
}

// ---------------------------------------------------------------------------
// forward particle step: gather + Boris + position + current + periodic wrap
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_particles(AGeom g, R* __restrict__ P,
                                                   const R* __restrict__ E,
                                                   const R* __restrict__ B1,
                                                   R* __restrict__ J)
{
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np)
    return;
  const long n  = g.np;
  R          x = P[p], y = P[n + p], z = P[2 * n + p];
  R          ux = P[3 * n + p], uy = P[4 * n + p], uz = P[5 * n + p];
  R          eb[6];
  gather<R>(g, E, B1, x, y, eb);
  push_boris(ux, uy, uz, eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], R(g.cc));
  const R    gm = lorentz_factor(ux, uy, uz, R(g.rc));
  const R    dt = R(g.dt) / gm;
  const R    x1 = x + ux * dt, y1 = y + uy * dt, z1 = z + uz * dt;
  const long v  = g.ncell;
  esirkepov2<R>(g, x, y, z, x1, y1, z1, [&](int iy, int ix, int c, R w) {
    unsafeAtomicAdd(&J[c * v + cell(g, iy, ix)], w);
  });
  // XtensorParticle::set_boundary_periodic (nix/xtensor_particle.hpp:359-376)
  const R lx = R(g.lx), ly = R(g.ly);
  P[p]         = x1 + (x1 < 0) * lx - (x1 >= lx) * lx;
  P[n + p]     = y1 + (y1 < 0) * ly - (y1 >= ly) * ly;
  P[2 * n + p] = z1;
  P[3 * n + p] = ux;
  P[4 * n + p] = uy;
  P[5 * n + p] = uz;
}

// ---------------------------------------------------------------------------
// adjoint particle step.  P: the pre-step state; L: the particle co-state (in:
// after the step, out: before it); lE1: the co-state of E after the step (lJ =
// -dt lE1); lEg / lB1g: scatter of the gather's adjoint.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_particles_T(AGeom g, const R* __restrict__ P,
                                                     const R* __restrict__ E,
                                                     const R* __restrict__ B1,
                                                     const R* __restrict__ lE1,
                                                     R* __restrict__ L,
                                                     R* __restrict__ lEg,
                                                     R* __restrict__ lB1g)
{
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np)
    return;
  const long n = g.np, v = g.ncell;
  const R    x = P[p], y = P[n + p], z = P[2 * n + p];
  const R    u0[3] = {P[3 * n + p], P[4 * n + p], P[5 * n + p]};
  // gather and its position derivative
  using D2 = Dual<2, R>;
  D2 ebd[6];
  gather<D2>(g, E, B1, D2::var(x, 0), D2::var(y, 1), ebd);
  // Boris and its Jacobian in (u, e, b)
  using D9 = Dual<9, R>;
  D9 uu[3] = {D9::var(u0[0], 0), D9::var(u0[1], 1), D9::var(u0[2], 2)};
  D9 eb9[6];
#pragma unroll
  for (int c = 0; c < 6; c++)
    eb9[c] = D9::var(ebd[c].v, 3 + c);
  push_boris<D9>(uu[0], uu[1], uu[2], eb9[0], eb9[1], eb9[2], eb9[3], eb9[4], eb9[5], D9(R(g.cc)));
  const R u1[3] = {uu[0].v, uu[1].v, uu[2].v};
  const R gm    = lorentz_factor(u1[0], u1[1], u1[2], R(g.rc));
  const R dtg   = R(g.dt) / gm;
  const R x1 = x + u1[0] * dtg, y1 = y + u1[1] * dtg, z1 = z + u1[2] * dtg;
  // current: sum over its nodes of lJ dJ/d(x, y, z, x', y', z')
  using D6 = Dual<6, R>;
  R gJ[6] = {0, 0, 0, 0, 0, 0};
  esirkepov2<D6>(g, D6::var(x, 0), D6::var(y, 1), D6::var(z, 2), D6::var(x1, 3), D6::var(y1, 4),
                 D6::var(z1, 5), [&](int iy, int ix, int c, const D6& w) {
                   const R lj = -R(g.dt) * lE1[c * v + cell(g, iy, ix)];
#pragma unroll
                   for (int k = 0; k < 6; k++)
                     gJ[k] += lj * w.d[k];
                 });
  // position: x' = x + f(u'), f(u) = u dt / gamma(u)
  R lx1[3], lpos[3];
  for (int k = 0; k < 3; k++) {
    lx1[k]  = L[k * n + p] + gJ[3 + k];
    lpos[k] = lx1[k] + gJ[k];
  }
  R       lu1[3];
  const R rc2 = R(g.rc) * R(g.rc);
  for (int i = 0; i < 3; i++) {
    R s = L[(3 + i) * n + p];
    for (int j = 0; j < 3; j++) {
      const R dfji = dtg * ((i == j ? R(1) : R(0)) - u1[j] * u1[i] * rc2 / (gm * gm));
      s += dfji * lx1[j];
    }
    lu1[i] = s;
  }
  // Boris transpose
  R lu0[3], leb[6];
  for (int k = 0; k < 3; k++)
    lu0[k] = uu[0].d[k] * lu1[0] + uu[1].d[k] * lu1[1] + uu[2].d[k] * lu1[2];
  for (int c = 0; c < 6; c++)
    leb[c] = uu[0].d[3 + c] * lu1[0] + uu[1].d[3 + c] * lu1[1] + uu[2].d[3 + c] * lu1[2];
  // gather transpose: scatter to the fields, and the weights' position derivative
  for (int c = 0; c < 6; c++) {
    lpos[0] += leb[c] * ebd[c].d[0];
    lpos[1] += leb[c] * ebd[c].d[1];
  }
  {
    R   wix[2], whx[2], wiy[2], why[2];
    int ix0, hx0, iy0, hy0;
    weights1(g, x, g.dx, g.rdx, wix, whx, ix0, hx0);
    weights1(g, y, g.dy, g.rdy, wiy, why, iy0, hy0);
    auto scatter = [&](R* F, int y0, int x0, const R* wy, const R* wx, R l) {
      const R a = l * R(g.qmdt);
      for (int jy = 0; jy < 2; jy++)
        for (int jx = 0; jx < 2; jx++)
          unsafeAtomicAdd(&F[cell(g, y0 + jy, x0 + jx)], a * wy[jy] * wx[jx]);
    };
    scatter(lEg, iy0, hx0, wiy, whx, leb[0]);
    scatter(lEg + v, hy0, ix0, why, wix, leb[1]);
    scatter(lEg + 2 * v, iy0, ix0, wiy, wix, leb[2]);
    scatter(lB1g, hy0, ix0, why, wix, leb[3]);
    scatter(lB1g + v, iy0, hx0, wiy, whx, leb[4]);
    scatter(lB1g + 2 * v, hy0, hx0, why, whx, leb[5]);
  }
  for (int k = 0; k < 3; k++) {
    L[k * n + p]       = lpos[k];
    L[(3 + k) * n + p] = lu0[k];
  }
}

// ---------------------------------------------------------------------------
// The FP32 form keeps every particle cell-relative, as the hot path's FP32 form
// does: P = [rx, ry, z, ux, uy, uz] (FP32; rx, ry in [0, 1), the offset inside
// the particle's cell in cell units) then [cx, cy] (int32), 8 x np 4-byte slots.
// The kernels work in the particle's own frame (x measured from its cell's lower
// face), so positions never exceed two cells and keep FP32's relative precision
// at any domain size; the co-states are the same as the FP64 form's (per unit
// length).  z enters only through z' - z, so the frame puts z at 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wrap_cell(float& r, int& c, int n)
{
  const float f = floorf(r);
  r -= f;
  if (r >= 1.0f)
    r = 0.99999994f;
  c += (int)f;
  c %= n;
  c = c < 0 ? c + n : c;
}

__global__ void __launch_bounds__(256) k_particles32(AGeom g, float* __restrict__ P,
                                                     const float* __restrict__ E,
                                                     const float* __restrict__ B1,
                                                     double* __restrict__ J)
{
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np)
    return;
  const long  n  = g.np;
  int*        C  = reinterpret_cast<int*>(P + 6 * n);
  const int   cx = C[p], cy = C[n + p];
  const float dx = (float)g.dx, dy = (float)g.dy;
  const float x = P[p] * dx, y = P[n + p] * dy;
  float       ux = P[3 * n + p], uy = P[4 * n + p], uz = P[5 * n + p];
  float       eb[6];
  gather<float>(g, E, B1, x, y, eb, cx, cy);
  push_boris(ux, uy, uz, eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], (float)g.cc);
  const float gm = lorentz_factor(ux, uy, uz, (float)g.rc);
  const float dt = (float)g.dt / gm;
  const float x1 = x + ux * dt, y1 = y + uy * dt, dz = uz * dt;
  const long  v  = g.ncell;
  esirkepov2<float>(
      g, x, y, 0.0f, x1, y1, dz,
      [&](int iy, int ix, int c, float w) {
        unsafeAtomicAdd(&J[c * v + cell(g, iy, ix)], (double)w);
      },
      cx, cy);
  float rx = x1 * (float)g.rdx, ry = y1 * (float)g.rdy;
  int   nx = cx, ny = cy;
  wrap_cell(rx, nx, g.nx);
  wrap_cell(ry, ny, g.ny);
  P[p]         = rx;
  P[n + p]     = ry;
  P[2 * n + p] += dz;
  P[3 * n + p] = ux;
  P[4 * n + p] = uy;
  P[5 * n + p] = uz;
  C[p]         = nx;
  C[n + p]     = ny;
}

// the adjoint particle step of the FP32 form (k_particles_T in the particle's frame)
__global__ void __launch_bounds__(256) k_particles32_T(AGeom g, const float* __restrict__ P,
                                                       const float* __restrict__ E,
                                                       const float* __restrict__ B1,
                                                       const float* __restrict__ lE1,
                                                       float* __restrict__ L,
                                                       double* __restrict__ lEg,
                                                       double* __restrict__ lB1g)
{
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np)
    return;
  using R          = float;
  const long  n    = g.np, v = g.ncell;
  const int*  C    = reinterpret_cast<const int*>(P + 6 * n);
  const int   cx   = C[p], cy = C[n + p];
  const R     x    = P[p] * (R)g.dx, y = P[n + p] * (R)g.dy;
  const R     u0[3] = {P[3 * n + p], P[4 * n + p], P[5 * n + p]};
  using D2         = Dual<2, R>;
  D2 ebd[6];
  gather<D2>(g, E, B1, D2::var(x, 0), D2::var(y, 1), ebd, cx, cy);
  using D9 = Dual<9, R>;
  D9 uu[3] = {D9::var(u0[0], 0), D9::var(u0[1], 1), D9::var(u0[2], 2)};
  D9 eb9[6];
#pragma unroll
  for (int c = 0; c < 6; c++)
    eb9[c] = D9::var(ebd[c].v, 3 + c);
  push_boris<D9>(uu[0], uu[1], uu[2], eb9[0], eb9[1], eb9[2], eb9[3], eb9[4], eb9[5], D9(R(g.cc)));
  const R u1[3] = {uu[0].v, uu[1].v, uu[2].v};
  const R gm    = lorentz_factor(u1[0], u1[1], u1[2], R(g.rc));
  const R dtg   = R(g.dt) / gm;
  const R x1 = x + u1[0] * dtg, y1 = y + u1[1] * dtg, z1 = u1[2] * dtg;
  using D6 = Dual<6, R>;
  R gJ[6] = {0, 0, 0, 0, 0, 0};
  esirkepov2<D6>(
      g, D6::var(x, 0), D6::var(y, 1), D6::var(R(0), 2), D6::var(x1, 3), D6::var(y1, 4),
      D6::var(z1, 5),
      [&](int iy, int ix, int c, const D6& w) {
        const R lj = -R(g.dt) * lE1[c * v + cell(g, iy, ix)];
#pragma unroll
        for (int k = 0; k < 6; k++)
          gJ[k] += lj * w.d[k];
      },
      cx, cy);
  R lx1[3], lpos[3];
  for (int k = 0; k < 3; k++) {
    lx1[k]  = L[k * n + p] + gJ[3 + k];
    lpos[k] = lx1[k] + gJ[k];
  }
  R       lu1[3];
  const R rc2 = R(g.rc) * R(g.rc);
  for (int i = 0; i < 3; i++) {
    R s = L[(3 + i) * n + p];
    for (int j = 0; j < 3; j++) {
      const R dfji = dtg * ((i == j ? R(1) : R(0)) - u1[j] * u1[i] * rc2 / (gm * gm));
      s += dfji * lx1[j];
    }
    lu1[i] = s;
  }
  R lu0[3], leb[6];
  for (int k = 0; k < 3; k++)
    lu0[k] = uu[0].d[k] * lu1[0] + uu[1].d[k] * lu1[1] + uu[2].d[k] * lu1[2];
  for (int c = 0; c < 6; c++)
    leb[c] = uu[0].d[3 + c] * lu1[0] + uu[1].d[3 + c] * lu1[1] + uu[2].d[3 + c] * lu1[2];
  for (int c = 0; c < 6; c++) {
    lpos[0] += leb[c] * ebd[c].d[0];
    lpos[1] += leb[c] * ebd[c].d[1];
  }
  {
    R   wix[2], whx[2], wiy[2], why[2];
    int ix0, hx0, iy0, hy0;
    weights1(g, x, g.dx, g.rdx, wix, whx, ix0, hx0, cx);
    weights1(g, y, g.dy, g.rdy, wiy, why, iy0, hy0, cy);
    auto scatter = [&](double* F, int y0, int x0, const R* wy, const R* wx, R l) {
      const R a = l * R(g.qmdt);
      for (int jy = 0; jy < 2; jy++)
        for (int jx = 0; jx < 2; jx++)
          unsafeAtomicAdd(&F[cell(g, y0 + jy, x0 + jx)], (double)(a * wy[jy] * wx[jx]));
    };
    scatter(lEg, iy0, hx0, wiy, whx, leb[0]);
    scatter(lEg + v, hy0, ix0, why, wix, leb[1]);
    scatter(lEg + 2 * v, iy0, ix0, wiy, wix, leb[2]);
    scatter(lB1g, hy0, ix0, why, wix, leb[3]);
    scatter(lB1g + v, iy0, hx0, wiy, whx, leb[4]);
    scatter(lB1g + 2 * v, hy0, hx0, why, whx, leb[5]);
  }
  for (int k = 0; k < 3; k++) {
    L[k * n + p]       = lpos[k];
    L[(3 + k) * n + p] = lu0[k];
  }
}

// ---------------------------------------------------------------------------
// initial state from theta, and dL/dtheta from the initial co-state
// base: [x0, y0, N_x, N_y, N_z]; modes: [M][3] (kx, ky, phi) in radians per length
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_init(AGeom g, const double* __restrict__ base,
                                              const double* __restrict__ modes, int nm,
                                              const double* __restrict__ theta, double vth,
                                              R* __restrict__ P)
{
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np)
    return;
  const long n  = g.np;
  double     x0 = base[p], y0 = base[n + p];
  double     x = x0, y = y0;
  for (int k = 0; k < nm; k++) {
    const double kx = modes[3 * k], ky = modes[3 * k + 1], ph = modes[3 * k + 2];
    const double s  = sin(kx * x0 + ky * y0 + ph) / (kx * kx + ky * ky);
    x -= theta[1 + k] * kx * s;
    y -= theta[1 + k] * ky * s;
  }
  x            = x - floor(x / g.lx) * g.lx;
  y            = y - floor(y / g.ly) * g.ly;
  x            = x >= g.lx ? x - g.lx : x;
  y            = y >= g.ly ? y - g.ly : y;
  if constexpr (sizeof(R) == sizeof(float)) {
    // the FP32 form: cell-relative offsets of the FP64 initial state
    int*         C  = reinterpret_cast<int*>(P + 6 * n);
    const double ux = x * g.rdx, uy = y * g.rdy;
    const double fx = floor(ux), fy = floor(uy);
    float        rx = (float)(ux - fx), ry = (float)(uy - fy);
    int          cx = (int)fx, cy = (int)fy;
    wrap_cell(rx, cx, g.nx);
    wrap_cell(ry, cy, g.ny);
    P[p]     = rx;
    P[n + p] = ry;
    C[p]     = cx;
    C[n + p] = cy;
  } else {
    P[p]     = R(x);
    P[n + p] = R(y);
  }
  P[2 * n + p] = R(0);
  P[3 * n + p] = R(theta[0] * vth * base[2 * n + p]);
  P[4 * n + p] = R(vth * base[3 * n + p]);
  P[5 * n + p] = R(vth * base[4 * n + p]);
}

// per-block partial sums of dL/dtheta_j = sum_p lx_p . dx_p/dtheta_j
template <typename R>
__global__ void __launch_bounds__(256) k_grad(AGeom g, const double* __restrict__ base,
                                              const double* __restrict__ modes, int nm,
                                              double vth, const R* __restrict__ L,
                                              double* __restrict__ partial)
{
  __shared__ double red[256];
  const long p  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n  = g.np;
  const int  nt = 1 + nm;
  for (int j = 0; j < nt; j++) {
    double acc = 0.0;
    if (p < n) {
      if (j == 0) {
        acc = (double)L[3 * n + p] * vth * base[2 * n + p];
      } else {
        const double x0 = base[p], y0 = base[n + p];
        const double kx = modes[3 * (j - 1)], ky = modes[3 * (j - 1) + 1], ph = modes[3 * (j - 1) + 2];
        const double s  = sin(kx * x0 + ky * y0 + ph) / (kx * kx + ky * ky);
        acc             = -((double)L[p] * kx + (double)L[n + p] * ky) * s;
      }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st)
        red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      partial[(long)j * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

// per-block partial sums of 1/2 B^2
template <typename R>
__global__ void __launch_bounds__(256) k_benergy(AGeom g, const R* __restrict__ B,
                                                 double* __restrict__ partial)
{
  __shared__ double red[256];
  const long i   = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double     acc = 0.0;
  if (i < g.ncell) {
    const long v = g.ncell;
    const double bx = B[i], by = B[v + i], bz = B[2 * v + i];
    acc = 0.5 * (bx * bx + by * by + bz * bz);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partial[blockIdx.x] = red[0];
}

} // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct AdjState {
  pn_adj_params       p{};
  AGeom               g{};
  bool                fp32   = false; // real type of the particles, fields and co-states
  hipStream_t         stream = nullptr;
  std::vector<double> modes; // [M][3] kx, ky, phi
  std::vector<double> mnum;  // [M][3] m_x, m_y, phi
  double *d_base = nullptr, *d_modes = nullptr, *d_theta = nullptr;
  // state and co-state, elements of the form's real type
  void *d_P = nullptr, *d_E = nullptr, *d_B = nullptr, *d_B1 = nullptr;
  void *d_L = nullptr, *d_lE = nullptr, *d_lB = nullptr, *d_lB2 = nullptr;
  // particle-to-grid sums (current, gathered co-states): FP64 in both forms,
  // so that an FP32 grid value is rounded once, after all particles added in
  double *d_J = nullptr, *d_lEg = nullptr, *d_lB1g = nullptr;
  // checkpoints (every K steps) and the recomputed segment
  std::vector<void*> ckpt;
  void*              d_seg    = nullptr;
  int                seg_cap  = 0;
  int                nsteps   = 0;
  int                K        = 0;
  bool               recorded = false;
  std::string        err;
  std::vector<double> theta;

  size_t esz() const { return fp32 ? sizeof(float) : sizeof(double); }
  // particle state: 6 reals per particle, plus the FP32 form's 2 cell indices
  size_t pbytes() const { return (fp32 ? 8 : 6) * (size_t)g.np * esz(); }
  // particle co-state: 6 reals per particle
  size_t lbytes() const { return 6 * (size_t)g.np * esz(); }
  size_t fbytes() const { return 3 * (size_t)g.ncell * esz(); }
  size_t abytes() const { return 3 * (size_t)g.ncell * sizeof(double); }
  size_t sbytes() const { return pbytes() + 2 * fbytes(); }
};

} // namespace picnix

struct pn_adj {
  picnix::AdjState s;
};

namespace picnix {
namespace {

void chk(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}
#define ACHK(call) ::picnix::chk((call), #call)

void* dnew(size_t bytes)
{
  void* p = nullptr;
  ACHK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
  return p;
}

template <typename R>
R* as(void* p)
{
  return static_cast<R*>(p);
}

template <typename R>
void bfd(AdjState& a, void* E, void* B)
{
  const AGeom& g = a.g;
  hipLaunchKernelGGL(k_bfd<R>, dim3(nblk(g.ncell, 256)), dim3(256), 0, a.stream, g, as<R>(E),
                     as<R>(B), R(g.cc * (0.5 * g.dt) / g.dx), R(g.cc * (0.5 * g.dt) / g.dy));
}

// one forward step; B1 keeps B after the first half push
template <typename R>
void step(AdjState& a)
{
  const AGeom& g = a.g;
  const long   f = 3 * g.ncell;
  hipLaunchKernelGGL(k_copy<R>, dim3(nblk(f, 256)), dim3(256), 0, a.stream, f, as<R>(a.d_B),
                     as<R>(a.d_B1));
  bfd<R>(a, a.d_E, a.d_B1);
  hipLaunchKernelGGL(k_set<double>, dim3(nblk(f, 256)), dim3(256), 0, a.stream, f, 0.0, a.d_J);
  if constexpr (std::is_same<R, float>::value)
    hipLaunchKernelGGL(k_particles32, dim3(nblk(g.np, 256)), dim3(256), 0, a.stream, g,
                       as<R>(a.d_P), as<R>(a.d_E), as<R>(a.d_B1), a.d_J);
  else
    hipLaunchKernelGGL(k_particles<R>, dim3(nblk(g.np, 256)), dim3(256), 0, a.stream, g,
                       as<R>(a.d_P), as<R>(a.d_E), as<R>(a.d_B1), a.d_J);
  hipLaunchKernelGGL(k_copy<R>, dim3(nblk(f, 256)), dim3(256), 0, a.stream, f, as<R>(a.d_B1),
                     as<R>(a.d_B));
  bfd<R>(a, a.d_E, a.d_B);
  hipLaunchKernelGGL((k_efd<R, double>), dim3(nblk(g.ncell, 256)), dim3(256), 0, a.stream, g,
                     as<R>(a.d_E), as<R>(a.d_B), a.d_J, R(g.cc * g.dt / g.dx),
                     R(g.cc * g.dt / g.dy));
}

void step_any(AdjState& a)
{
  if (a.fp32)
    step<float>(a);
  else
    step<double>(a);
}

void save(AdjState& a, void* dst)
{
  char* d = static_cast<char*>(dst);
  ACHK(hipMemcpyAsync(d, a.d_P, a.pbytes(), hipMemcpyDeviceToDevice, a.stream));
  ACHK(hipMemcpyAsync(d + a.pbytes(), a.d_E, a.fbytes(), hipMemcpyDeviceToDevice, a.stream));
  ACHK(hipMemcpyAsync(d + a.pbytes() + a.fbytes(), a.d_B, a.fbytes(), hipMemcpyDeviceToDevice,
                      a.stream));
}

void restore(AdjState& a, const void* src)
{
  const char* d = static_cast<const char*>(src);
  ACHK(hipMemcpyAsync(a.d_P, d, a.pbytes(), hipMemcpyDeviceToDevice, a.stream));
  ACHK(hipMemcpyAsync(a.d_E, d + a.pbytes(), a.fbytes(), hipMemcpyDeviceToDevice, a.stream));
  ACHK(hipMemcpyAsync(a.d_B, d + a.pbytes() + a.fbytes(), a.fbytes(), hipMemcpyDeviceToDevice,
                      a.stream));
}

double benergy(AdjState& a)
{
  const AGeom& g  = a.g;
  const int    nb = nblk(g.ncell, 256);
  double*      d  = static_cast<double*>(dnew(nb * sizeof(double)));
  if (a.fp32)
    hipLaunchKernelGGL(k_benergy<float>, dim3(nb), dim3(256), 0, a.stream, g, as<float>(a.d_B), d);
  else
    hipLaunchKernelGGL(k_benergy<double>, dim3(nb), dim3(256), 0, a.stream, g, as<double>(a.d_B),
                       d);
  std::vector<double> h(nb);
  ACHK(hipMemcpyAsync(h.data(), d, nb * sizeof(double), hipMemcpyDeviceToHost, a.stream));
  ACHK(hipStreamSynchronize(a.stream));
  ACHK(hipFree(d));
  double s = 0;
  for (double v : h)
    s += v;
  return s;
}

// one adjoint step for the pre-step state stored at `pre`
template <typename R>
void step_T(AdjState& a, const void* pre)
{
  const AGeom& g = a.g;
  const long   f = 3 * g.ncell;
  const dim3   gf(nblk(f, 256)), gc(nblk(g.ncell, 256)), b(256);
  restore(a, pre); // P, E, B before the step
  hipLaunchKernelGGL(k_copy<R>, gf, b, 0, a.stream, f, as<R>(a.d_B), as<R>(a.d_B1));
  bfd<R>(a, a.d_E, a.d_B1);
  hipLaunchKernelGGL(k_set<double>, gf, b, 0, a.stream, f, 0.0, a.d_lEg);
  hipLaunchKernelGGL(k_set<double>, gf, b, 0, a.stream, f, 0.0, a.d_lB1g);
  if constexpr (std::is_same<R, float>::value)
    hipLaunchKernelGGL(k_particles32_T, dim3(nblk(g.np, 256)), b, 0, a.stream, g, as<R>(a.d_P),
                       as<R>(a.d_E), as<R>(a.d_B1), as<R>(a.d_lE), as<R>(a.d_L), a.d_lEg,
                       a.d_lB1g);
  else
    hipLaunchKernelGGL(k_particles_T<R>, dim3(nblk(g.np, 256)), b, 0, a.stream, g, as<R>(a.d_P),
                       as<R>(a.d_E), as<R>(a.d_B1), as<R>(a.d_lE), as<R>(a.d_L), a.d_lEg,
                       a.d_lB1g);
  const R cflx = R(g.cc * g.dt / g.dx), cfly = R(g.cc * g.dt / g.dy);
  const R hx = R(g.cc * (0.5 * g.dt) / g.dx), hy = R(g.cc * (0.5 * g.dt) / g.dy);
  // lB2t = lB2 + K_E^T lE'   (d_lB holds lB2 on entry)
  hipLaunchKernelGGL(k_efd_T<R>, gc, b, 0, a.stream, g, as<R>(a.d_lE), as<R>(a.d_lB), cflx, cfly);
  // lB1t = lB2t + lB1g
  hipLaunchKernelGGL(k_copy<R>, gf, b, 0, a.stream, f, as<R>(a.d_lB), as<R>(a.d_lB2));
  hipLaunchKernelGGL((k_axpy<R, double>), gf, b, 0, a.stream, f, R(1), a.d_lB1g, as<R>(a.d_lB));
  // lE = lE' + lEg + K_B^T lB2t + K_B^T lB1t
  hipLaunchKernelGGL((k_axpy<R, double>), gf, b, 0, a.stream, f, R(1), a.d_lEg, as<R>(a.d_lE));
  hipLaunchKernelGGL(k_bfd_T<R>, gc, b, 0, a.stream, g, as<R>(a.d_lB2), as<R>(a.d_lE), hx, hy);
  hipLaunchKernelGGL(k_bfd_T<R>, gc, b, 0, a.stream, g, as<R>(a.d_lB), as<R>(a.d_lE), hx, hy);
  // lB = lB1t (already in d_lB)
}

// host FP64 values to the form's real type on the device, and back
void upload(AdjState& a, void* dst, const double* src, size_t n)
{
  if (!a.fp32) {
    ACHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
    return;
  }
  std::vector<float> t(src, src + n);
  ACHK(hipMemcpy(dst, t.data(), n * sizeof(float), hipMemcpyHostToDevice));
}

void download(AdjState& a, double* dst, const void* src, size_t n)
{
  if (!a.fp32) {
    ACHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
    return;
  }
  std::vector<float> t(n);
  ACHK(hipMemcpy(t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
  std::copy(t.begin(), t.end(), dst);
}

// the FP32 form's particles from and to absolute FP64 values [6][np]
void upload_cells(AdjState& a, const double* soa)
{
  const long         n = a.g.np;
  std::vector<float> f(8 * (size_t)n);
  int*               C = reinterpret_cast<int*>(f.data() + 6 * n);
  for (long i = 0; i < n; i++) {
    for (int k = 0; k < 2; k++) {
      const double u  = soa[(size_t)k * n + i] * (k == 0 ? a.g.rdx : a.g.rdy);
      const double fl = std::floor(u);
      const int    nc = k == 0 ? a.g.nx : a.g.ny;
      float        r  = (float)(u - fl);
      int          c  = (int)fl;
      if (r >= 1.0f)
        r = 0.99999994f;
      c = ((c % nc) + nc) % nc;
      f[(size_t)k * n + i]  = r;
      C[(size_t)k * n + i] = c;
    }
    for (int k = 2; k < 6; k++)
      f[(size_t)k * n + i] = (float)soa[(size_t)k * n + i];
  }
  ACHK(hipMemcpy(a.d_P, f.data(), a.pbytes(), hipMemcpyHostToDevice));
}

void download_cells(AdjState& a, double* soa)
{
  const long         n = a.g.np;
  std::vector<float> f(8 * (size_t)n);
  ACHK(hipMemcpy(f.data(), a.d_P, a.pbytes(), hipMemcpyDeviceToHost));
  const int* C = reinterpret_cast<const int*>(f.data() + 6 * n);
  for (long i = 0; i < n; i++) {
    soa[i]     = ((double)C[i] + (double)f[i]) * a.g.dx;
    soa[n + i] = ((double)C[n + i] + (double)f[n + i]) * a.g.dy;
    for (int k = 2; k < 6; k++)
      soa[(size_t)k * n + i] = (double)f[(size_t)k * n + i];
  }
}

template <typename F>
int guard(pn_adj* h, F f)
{
  if (!h)
    return -1;
  try {
    return f(h->s);
  } catch (const HipError& e) {
    h->s.err = e.msg;
    return -3;
  } catch (const std::exception& e) {
    h->s.err = e.what();
    return -2;
  }
}

} // namespace
} // namespace picnix

using picnix::AdjState;

extern "C" {

int pn_adj_create(const pn_adj_params* p, pn_adj** out)
{
  if (!p || !out)
    return -1;
  *out   = nullptr;
  auto h = new pn_adj();
  int  rc = picnix::guard(h, [&](AdjState& a) {
    if (p->Nx <= 1 || p->Ny <= 1 || p->nppc <= 0 || p->delh <= 0 || p->delt <= 0 || p->cc <= 0 ||
        p->nmodes < 0 || p->nmodes > PN_ADJ_MAX_MODES || p->kmax <= 0) {
      a.err = "invalid adjoint-form parameters";
      return -1;
    }
    if (p->precision != 0 && p->precision != 1) {
      a.err = "precision must be 0 (FP64) or 1 (FP32)";
      return -1;
    }
    a.p           = *p;
    a.fp32        = p->precision == 1;
    picnix::AGeom& g = a.g;
    g.nx          = p->Nx;
    g.ny          = p->Ny;
    g.ncell       = (long)p->Nx * p->Ny;
    g.dx = g.dy   = p->delh;
    g.rdx = g.rdy = 1 / p->delh;
    g.dt          = p->delt;
    g.cc          = p->cc;
    g.rc          = 1 / p->cc;
    g.lx          = p->Nx * p->delh;
    g.ly          = p->Ny * p->delh;
    const double m = p->ro / p->nppc;
    g.q           = p->qm * m;
    g.qmdt        = 0.5 * g.q / m * p->delt; // as PushConsts::qmdt
    g.np          = (long)g.ncell * p->nppc;
    ACHK(hipStreamCreate(&a.stream));
    // seeded base draws (std::mt19937_64, libstdc++ uniform and normal), then modes
    std::mt19937_64                        mt(p->seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::normal_distribution<double>       nor(0.0, 1.0);
    const long                             n = g.np;
    std::vector<double>                    base(5 * (size_t)n);
    for (long i = 0; i < n; i++) {
      base[i]     = uni(mt) * g.lx;
      base[n + i] = uni(mt) * g.ly;
    }
    for (long i = 0; i < n; i++)
      for (int c = 0; c < 3; c++)
        base[(2 + c) * (size_t)n + i] = nor(mt);
    std::uniform_int_distribution<int> mi(-p->kmax, p->kmax);
    for (int k = 0; k < p->nmodes; k++) {
      int mx = 0, my = 0;
      while (mx == 0 && my == 0) {
        mx = mi(mt);
        my = mi(mt);
      }
      const double ph = 2 * M_PI * uni(mt);
      a.mnum.insert(a.mnum.end(), {(double)mx, (double)my, ph});
      a.modes.insert(a.modes.end(), {2 * M_PI * mx / g.lx, 2 * M_PI * my / g.ly, ph});
    }
    a.d_base  = static_cast<double*>(picnix::dnew(base.size() * sizeof(double)));
    a.d_modes = static_cast<double*>(picnix::dnew(std::max<size_t>(a.modes.size(), 1) * sizeof(double)));
    a.d_theta = static_cast<double*>(picnix::dnew((1 + PN_ADJ_MAX_MODES) * sizeof(double)));
    ACHK(hipMemcpy(a.d_base, base.data(), base.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!a.modes.empty())
      ACHK(hipMemcpy(a.d_modes, a.modes.data(), a.modes.size() * sizeof(double),
                     hipMemcpyHostToDevice));
    a.d_P   = picnix::dnew(a.pbytes());
    a.d_L   = picnix::dnew(a.lbytes());
    for (void** f : {&a.d_E, &a.d_B, &a.d_B1, &a.d_lE, &a.d_lB, &a.d_lB2})
      *f = picnix::dnew(a.fbytes());
    for (double** f : {&a.d_J, &a.d_lEg, &a.d_lB1g})
      *f = static_cast<double*>(picnix::dnew(a.abytes()));
    ACHK(hipMemset(a.d_E, 0, a.fbytes()));
    ACHK(hipMemset(a.d_B, 0, a.fbytes()));
    ACHK(hipMemset(a.d_P, 0, a.pbytes()));
    return 0;
  });
  *out = h;
  return rc;
}

int pn_adj_destroy(pn_adj* h)
{
  if (!h)
    return -1;
  AdjState& a = h->s;
  for (void* p : {(void*)a.d_base, (void*)a.d_modes, (void*)a.d_theta, a.d_P, a.d_E, a.d_B,
                  a.d_B1, (void*)a.d_J, a.d_L, a.d_lE, a.d_lB, (void*)a.d_lEg, (void*)a.d_lB1g,
                  a.d_lB2, a.d_seg})
    if (p)
      (void)hipFree(p);
  for (void* p : a.ckpt)
    (void)hipFree(p);
  if (a.stream)
    (void)hipStreamDestroy(a.stream);
  delete h;
  return 0;
}

const char* pn_adj_last_error(const pn_adj* h) { return h ? h->s.err.c_str() : "null handle"; }

int pn_adj_modes(pn_adj* h, double* out)
{
  return picnix::guard(h, [&](AdjState& a) {
    if (!out)
      return -1;
    std::copy(a.mnum.begin(), a.mnum.end(), out);
    return 0;
  });
}

int pn_adj_forward(pn_adj* h, const double* theta, int nsteps, int K, double* loss)
{
  return picnix::guard(h, [&](AdjState& a) {
    if (!theta || nsteps < 0 || K <= 0) {
      a.err = "invalid theta, step count or checkpoint interval";
      return -1;
    }
    const picnix::AGeom& g  = a.g;
    const int            nm = a.p.nmodes;
    a.theta.assign(theta, theta + 1 + nm);
    ACHK(hipMemcpyAsync(a.d_theta, theta, (1 + nm) * sizeof(double), hipMemcpyHostToDevice,
                        a.stream));
    if (a.fp32)
      hipLaunchKernelGGL(picnix::k_init<float>, dim3(picnix::nblk(g.np, 256)), dim3(256), 0,
                         a.stream, g, a.d_base, a.d_modes, nm, a.d_theta, a.p.vth,
                         static_cast<float*>(a.d_P));
    else
      hipLaunchKernelGGL(picnix::k_init<double>, dim3(picnix::nblk(g.np, 256)), dim3(256), 0,
                         a.stream, g, a.d_base, a.d_modes, nm, a.d_theta, a.p.vth,
                         static_cast<double*>(a.d_P));
    ACHK(hipMemsetAsync(a.d_E, 0, a.fbytes(), a.stream));
    ACHK(hipMemsetAsync(a.d_B, 0, a.fbytes(), a.stream));
    const int nck = (nsteps + K - 1) / K + 1;
    while ((int)a.ckpt.size() < nck)
      a.ckpt.push_back(picnix::dnew(a.sbytes()));
    for (int s = 0; s < nsteps; s++) {
      if (s % K == 0)
        picnix::save(a, a.ckpt[s / K]);
      picnix::step_any(a);
    }
    ACHK(hipGetLastError());
    a.nsteps   = nsteps;
    a.K        = K;
    a.recorded = true;
    const double L = picnix::benergy(a);
    if (loss)
      *loss = L;
    return 0;
  });
}

int pn_adj_backward(pn_adj* h, double* grad)
{
  return picnix::guard(h, [&](AdjState& a) {
    if (!a.recorded || !grad) {
      a.err = "pn_adj_backward needs a pn_adj_forward first";
      return -6;
    }
    const picnix::AGeom& g  = a.g;
    const int            nm = a.p.nmodes;
    const long           f  = 3 * g.ncell;
    // final state: L = 1/2 sum B^2 -> lB = B, everything else 0
    if (a.fp32)
      hipLaunchKernelGGL(picnix::k_copy<float>, dim3(picnix::nblk(f, 256)), dim3(256), 0,
                         a.stream, f, static_cast<float*>(a.d_B), static_cast<float*>(a.d_lB));
    else
      hipLaunchKernelGGL(picnix::k_copy<double>, dim3(picnix::nblk(f, 256)), dim3(256), 0,
                         a.stream, f, static_cast<double*>(a.d_B), static_cast<double*>(a.d_lB));
    ACHK(hipMemsetAsync(a.d_lE, 0, a.fbytes(), a.stream));
    ACHK(hipMemsetAsync(a.d_L, 0, a.lbytes(), a.stream));
    if (a.seg_cap < a.K) {
      if (a.d_seg)
        ACHK(hipFree(a.d_seg));
      a.d_seg   = picnix::dnew(a.sbytes() * (size_t)a.K);
      a.seg_cap = a.K;
    }
    const int nseg = (a.nsteps + a.K - 1) / a.K;
    for (int sg = nseg - 1; sg >= 0; sg--) {
      const int s0 = sg * a.K, s1 = std::min(a.nsteps, s0 + a.K);
      // recompute the segment from its checkpoint, keeping every pre-step state
      picnix::restore(a, a.ckpt[sg]);
      for (int s = s0; s < s1; s++) {
        picnix::save(a, static_cast<char*>(a.d_seg) + a.sbytes() * (size_t)(s - s0));
        if (s + 1 < s1)
          picnix::step_any(a);
      }
      for (int s = s1 - 1; s >= s0; s--) {
        const void* pre = static_cast<char*>(a.d_seg) + a.sbytes() * (size_t)(s - s0);
        if (a.fp32)
          picnix::step_T<float>(a, pre);
        else
          picnix::step_T<double>(a, pre);
      }
    }
    // the initial fields do not depend on theta; the particles' co-state does
    const int nb = picnix::nblk(g.np, 256);
    double*   d  = static_cast<double*>(picnix::dnew((size_t)nb * (1 + nm) * sizeof(double)));
    if (a.fp32)
      hipLaunchKernelGGL(picnix::k_grad<float>, dim3(nb), dim3(256), 0, a.stream, g, a.d_base,
                         a.d_modes, nm, a.p.vth, static_cast<const float*>(a.d_L), d);
    else
      hipLaunchKernelGGL(picnix::k_grad<double>, dim3(nb), dim3(256), 0, a.stream, g, a.d_base,
                         a.d_modes, nm, a.p.vth, static_cast<const double*>(a.d_L), d);
    std::vector<double> hpart((size_t)nb * (1 + nm));
    ACHK(hipMemcpyAsync(hpart.data(), d, hpart.size() * sizeof(double), hipMemcpyDeviceToHost,
                        a.stream));
    ACHK(hipStreamSynchronize(a.stream));
    ACHK(hipFree(d));
    for (int j = 0; j <= nm; j++) {
      double s = 0;
      for (int b = 0; b < nb; b++)
        s += hpart[(size_t)j * nb + b];
      grad[j] = s;
    }
    // the particle state was overwritten by the recomputation: the forward's
    // final state is no longer held
    a.recorded = false;
    return 0;
  });
}

int pn_adj_set_state(pn_adj* h, const double* xu6, int64_t np, const double* E, const double* B)
{
  return picnix::guard(h, [&](AdjState& a) {
    if (np != a.g.np || !xu6) {
      a.err = "particle count differs from Nx * Ny * nppc";
      return -1;
    }
    std::vector<double> soa(6 * (size_t)np);
    for (int64_t i = 0; i < np; i++)
      for (int c = 0; c < 6; c++)
        soa[(size_t)c * np + i] = xu6[6 * i + c];
    if (a.fp32)
      picnix::upload_cells(a, soa.data());
    else
      picnix::upload(a, a.d_P, soa.data(), 6 * (size_t)np);
    if (E)
      picnix::upload(a, a.d_E, E, 3 * (size_t)a.g.ncell);
    if (B)
      picnix::upload(a, a.d_B, B, 3 * (size_t)a.g.ncell);
    a.recorded = false;
    return 0;
  });
}

int pn_adj_get_state(pn_adj* h, double* xu6, double* E, double* B)
{
  return picnix::guard(h, [&](AdjState& a) {
    ACHK(hipStreamSynchronize(a.stream));
    if (xu6) {
      const long          np = a.g.np;
      std::vector<double> soa(6 * (size_t)np);
      if (a.fp32)
        picnix::download_cells(a, soa.data());
      else
        picnix::download(a, soa.data(), a.d_P, 6 * (size_t)np);
      for (long i = 0; i < np; i++)
        for (int c = 0; c < 6; c++)
          xu6[6 * i + c] = soa[(size_t)c * np + i];
    }
    if (E)
      picnix::download(a, E, a.d_E, 3 * (size_t)a.g.ncell);
    if (B)
      picnix::download(a, B, a.d_B, 3 * (size_t)a.g.ncell);
    return 0;
  });
}

int pn_adj_step(pn_adj* h, int nsteps)
{
  return picnix::guard(h, [&](AdjState& a) {
    for (int s = 0; s < nsteps; s++)
      picnix::step_any(a);
    ACHK(hipGetLastError());
    ACHK(hipStreamSynchronize(a.stream));
    a.recorded = false;
    return 0;
  });
}

int pn_adj_num_particles(pn_adj* h, int64_t* np)
{
  return picnix::guard(h, [&](AdjState& a) {
    if (!np)
      return -1;
    *np = a.g.np;
    return 0;
  });
}

} // extern "C"
