// picnix_oracle.cpp -- FP64 CPU oracle for the pic-nix per-timestep hot path.
//
// TEST INFRASTRUCTURE ONLY (see picnix_oracle.h).  Every routine restates the
// reference algorithm at the cited file:line, in the reference's own data layout
// and operation order, compiled with -ffp-contract=off so that each expression
// is evaluated exactly as written (one IEEE rounding per operation).
//
// Coverage of the checks that pin each function (tests/test_oracle_*.py):
//   shape_mc1..4        closed-form B-splines (nix/unittest/test_primitives.cpp:271-530)
//   digitize            definition + test_primitives.cpp:73-117 cases
//   lorentz_factor      closed form
//   push_boris          |u| conservation for E=0, gyro phase 2*atan(|t|), E x B drift,
//                       independent textbook Boris on tiny cases
//   esirkepov deposit   unit charge + discrete continuity (test_esirkepov.cpp:480-566)
//                       and the non-crossing reduction Jx = q vx <S>
//   gather              naive triple loop + exact reproduction of linear fields
//   Yee E/B             vacuum plane-wave dispersion, div B preserved
//   count/sort          permutation, per-bin membership, brute-force stable_sort
//   halo/migration      chunked oracle vs a single-chunk run of the same domain
//   mt19937_64/uniform/normal   the C++ standard's 10000th-value check + libstdc++
//   gilbert chunk map   check_index / check_locality properties (test_sfc.cpp:11-118)
#include "picnix_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int SIMD_WIDTH = 8; // nix/nix.hpp:102-105 (NIX_SIMD_WIDTH default)
constexpr int NC         = 7; // nix/particle.hpp:19 (x,y,z,ux,uy,uz,id)

// ---------------------------------------------------------------------------
// primitives (nix/primitives.hpp)
// ---------------------------------------------------------------------------

// nix/primitives.hpp:46-58
inline int digitize(double x, double xmin, double rdx)
{
  return static_cast<int>(std::floor((x - xmin) * rdx));
}

// nix/primitives.hpp:157-161
inline double lorentz_factor(double ux, double uy, double uz, double rc)
{
  return std::sqrt(1 + (ux * ux + uy * uy + uz * uz) * rc * rc);
}

// nix/primitives.hpp:163-189
inline void push_boris(double& ux, double& uy, double& uz, double ex, double ey, double ez,
                       double bx, double by, double bz, double cc)
{
  ux += ex;
  uy += ey;
  uz += ez;
  double gm = 1 / std::sqrt(cc * cc + ux * ux + uy * uy + uz * uz);
  bx *= gm;
  by *= gm;
  bz *= gm;
  double bb = 2.0 / (1.0 + bx * bx + by * by + bz * bz);
  double vx = ux + (uy * bz - uz * by);
  double vy = uy + (uz * bx - ux * bz);
  double vz = uz + (ux * by - uy * bx);
  ux += (vy * bz - vz * by) * bb + ex;
  uy += (vz * bx - vx * bz) * bb + ey;
  uz += (vx * by - vy * bx) * bb + ez;
}

// nix/primitives.hpp:191-220
inline void push_vay(double& ux, double& uy, double& uz, double ex, double ey, double ez,
                     double bx, double by, double bz, double cc)
{
  double gm = 1 / std::sqrt(cc * cc + ux * ux + uy * uy + uz * uz);
  double vx = ux + 2 * ex + gm * (uy * bz - uz * by);
  double vy = uy + 2 * ey + gm * (uz * bx - ux * bz);
  double vz = uz + 2 * ez + gm * (ux * by - uy * bx);
  gm        = (cc * cc + vx * vx + vy * vy + vz * vz);
  double bb = bx * bx + by * by + bz * bz;
  double bu = bx * vx + by * vy + bz * vz;
  double xx = gm - bb;
  double yy = bb + bu * bu;
  gm        = 1 / std::sqrt(0.5 * (xx + std::sqrt(xx * xx + 4 * yy)));
  bx *= gm;
  by *= gm;
  bz *= gm;
  bu = bx * vx + by * vy + bz * vz;
  bb = 1.0 / (1.0 + bx * bx + by * by + bz * bz);
  ux = (vx + bu * bx + (vy * bz - vz * by)) * bb;
  uy = (vy + bu * by + (vz * bx - vx * bz)) * bb;
  uz = (vz + bu * bz + (vx * by - vy * bx)) * bb;
}

// nix/primitives.hpp:222-253
inline void push_higuera_cary(double& ux, double& uy, double& uz, double ex, double ey, double ez,
                              double bx, double by, double bz, double cc)
{
  ux += ex;
  uy += ey;
  uz += ez;
  double gm = cc * cc + ux * ux + uy * uy + uz * uz;
  double bb = bx * bx + by * by + bz * bz;
  double bu = bx * ux + by * uy + bz * uz;
  double xx = gm - bb;
  double yy = bb + bu * bu;
  gm        = 1 / std::sqrt(0.5 * (xx + std::sqrt(xx * xx + 4 * yy)));
  bx *= gm;
  by *= gm;
  bz *= gm;
  bb        = 2.0 / (1.0 + bx * bx + by * by + bz * bz);
  double vx = ux + (uy * bz - uz * by);
  double vy = uy + (uz * bx - ux * bz);
  double vz = uz + (ux * by - uy * bx);
  ux += (vy * bz - vz * by) * bb + ex;
  uy += (vz * bx - vx * bz) * bb + ey;
  uz += (vx * by - vy * bx) * bb + ez;
}

// nix/primitives.hpp:256-328 (shape_mc1..4), dispatched as shape_mc<Order> (:514-529)
template <int Order>
inline void shape_mc(double x, double X, double rdx, double* s)
{
  double delta = (x - X) * rdx;
  if constexpr (Order == 1) {
    s[0] = 1 - delta;
    s[1] = delta;
  } else if constexpr (Order == 2) {
    double w0 = delta;
    double w1 = 0.5 - w0;
    double w2 = 0.5 + w0;
    s[0]      = 0.50 * w1 * w1;
    s[1]      = 0.75 - w0 * w0;
    s[2]      = 0.50 * w2 * w2;
  } else if constexpr (Order == 3) {
    const double a       = 1 / 6.0;
    double       w1      = delta;
    double       w2      = 1 - delta;
    double       w1_pow2 = w1 * w1;
    double       w2_pow2 = w2 * w2;
    double       w1_pow3 = w1_pow2 * w1;
    double       w2_pow3 = w2_pow2 * w2;
    s[0]                 = a * w2_pow3;
    s[1]                 = a * (4 - 6 * w1_pow2 + 3 * w1_pow3);
    s[2]                 = a * (4 - 6 * w2_pow2 + 3 * w2_pow3);
    s[3]                 = a * w1_pow3;
  } else {
    const double a       = 1 / 384.0;
    const double b       = 1 / 96.0;
    const double c       = 115 / 192.0;
    const double d       = 1 / 8.0;
    double       w1      = 1 + delta;
    double       w2      = 1 - delta;
    double       w3      = 1 + delta * 2;
    double       w4      = 1 - delta * 2;
    double       w0_pow2 = delta * delta;
    double       w1_pow2 = w1 * w1;
    double       w2_pow2 = w2 * w2;
    double       w1_pow3 = w1_pow2 * w1;
    double       w2_pow3 = w2_pow2 * w2;
    double       w1_pow4 = w1_pow3 * w1;
    double       w2_pow4 = w2_pow3 * w2;
    double       w3_pow4 = w3 * w3 * w3 * w3;
    double       w4_pow4 = w4 * w4 * w4 * w4;
    s[0]                 = a * w4_pow4;
    s[1] = b * (55 + 20 * w1 - 120 * w1_pow2 + 80 * w1_pow3 - 16 * w1_pow4);
    s[2] = c + d * w0_pow2 * (2 * w0_pow2 - 5);
    s[3] = b * (55 + 20 * w2 - 120 * w2_pow2 + 80 * w2_pow3 - 16 * w2_pow4);
    s[4] = a * w3_pow4;
  }
}

template <typename T>
inline T wt_abs(T v)
{
  return std::fabs(v);
}

// WT shape functions (nix/primitives.hpp:331-566; Lu et al., J. Comput. Phys. 413,
// 109388 (2020)): the shape of order O is the B-spline of degree O-1 smeared by
// a top hat of half width dt = c dt / dx.  Left/middle/right branches selected
// by the 0/1 factors t1..t4 exactly as the reference combines them.
template <int Order, typename T>
inline void shape_wt(T x, T X, T rdx, T dt, T rdt, T* s)
{
  const T delta = (x - X) * rdx;
  if constexpr (Order == 1) {
    // top-hat (half width dt) smeared nearest-grid-point weight
    T ss = T(0.25) * rdt * (1 + 2 * dt - 2 * delta);
    ss   = T(0) < ss ? ss : T(0); // max(0, .) then min(1, .), as std::max / std::min
    ss   = ss < T(1) ? ss : T(1);
    s[0] = ss;
    s[1] = 1 - ss;
  } else if constexpr (Order == 2) {
    const T t1 = delta < -dt ? T(1) : T(0);
    const T t2 = 1 - t1;
    const T t3 = delta < +dt ? T(1) : T(0);
    const T t4 = 1 - t3;
    const T w0 = wt_abs(delta);
    const T w1 = dt - delta;
    const T w2 = dt + delta;
    // left (t1), middle (t2 t3) and right (t4) branches of the piecewise weight
    const T l[3] = {w0, 1 - w0, T(0)};
    const T m[3] = {T(0.25) * rdt * w1 * w1, T(0.50) * rdt * (dt * (2 - dt) - w0 * w0),
                    T(0.25) * rdt * w2 * w2};
    const T r[3] = {l[2], l[1], l[0]};
    for (int k = 0; k < 3; k++)
      s[k] = l[k] * t1 + m[k] * t2 * t3 + r[k] * t4;
  } else if constexpr (Order == 3) {
    const T a    = T(1) / T(96.0);
    const T b    = T(1) / T(24.0);
    const T c    = T(1) / T(12.0);
    const T adt  = a * rdt;
    const T t1   = delta < T(0.5) - dt ? T(1) : T(0);
    const T t2   = 1 - t1;
    const T t3   = delta < T(0.5) + dt ? T(1) : T(0);
    const T t4   = 1 - t3;
    const T w0   = delta;
    const T w1   = 1 - delta;
    const T w3   = 1 - 2 * delta;
    const T w4   = 1 + 2 * delta;
    const T w5   = 2 * dt + w3;
    const T w6   = 2 * dt - w3;
    const T w7   = 3 - 2 * delta;
    const T w0p2 = w0 * w0;
    const T w1p2 = w1 * w1;
    const T w3p2 = w3 * w3;
    const T w3p3 = w3p2 * w3;
    const T w4p2 = w4 * w4;
    const T w5p3 = w5 * w5 * w5;
    const T w6p3 = w6 * w6 * w6;
    const T w7p2 = w7 * w7;
    const T dtp2 = dt * dt;
    const T dtp3 = dtp2 * dt;
    const T dt4  = 4 * dtp2;
    const T sodd  = adt * (-8 * dtp3 - 6 * dt * w3p2);
    const T seven = adt * (-36 * dtp2 * w3 - 3 * w3p3);
    const T l[4] = {b * (dt4 + 3 * w3p2), c * (9 - dt4 - 12 * w0p2), b * (dt4 + 3 * w4p2), T(0)};
    const T m[4] = {adt * w5p3, sodd + seven + w1, sodd - seven + w0, adt * w6p3};
    const T r[4] = {T(0), b * (dt4 + 3 * w7p2), c * (9 - dt4 - 12 * w1p2), b * (dt4 + 3 * w3p2)};
    for (int k = 0; k < 4; k++)
      s[k] = l[k] * t1 + m[k] * t2 * t3 + r[k] * t4;
  } else {
    const T a    = T(1) / T(48.0);
    const T b    = T(1) / T(24.0);
    const T c    = T(1) / T(12.0);
    const T d    = T(1) / T(6.0);
    const T adt  = a * rdt;
    const T bdt  = b * rdt;
    const T cdt  = c * rdt;
    const T t1   = delta < -dt ? T(1) : T(0);
    const T t2   = 1 - t1;
    const T t3   = delta < +dt ? T(1) : T(0);
    const T t4   = 1 - t3;
    const T w0   = wt_abs(delta);
    const T w1   = 1 - w0;
    const T w2   = 1 - delta;
    const T w3   = 1 + delta;
    const T w4   = dt - delta;
    const T w5   = dt + delta;
    const T w0p2 = w0 * w0;
    const T w0p3 = w0p2 * w0;
    const T w0p4 = w0p3 * w0;
    const T w1p2 = w1 * w1;
    const T w1p3 = w1p2 * w1;
    const T w2p3 = w2 * w2 * w2;
    const T w3p3 = w3 * w3 * w3;
    const T w4p4 = w4 * w4 * w4 * w4;
    const T w5p4 = w5 * w5 * w5 * w5;
    const T dtp2 = dt * dt;
    const T dtp3 = dtp2 * dt;
    const T dtp4 = dtp3 * dt;
    const T ss1  = -dtp4 - 6 * w0p2 * dtp2 - w0p4;
    const T ss2  = 3 * dtp4 - 8 * dtp3 + 18 * w0p2 * dtp2 + (16 - 24 * w0p2) * dt + 3 * w0p4;
    const T l[5] = {d * w0 * (w0p2 + dtp2), d * (4 - 6 * w1p2 + 3 * w1p3 + (1 - 3 * w0) * dtp2),
                    d * (4 - 6 * w0p2 + 3 * w0p3 - (2 - 3 * w0) * dtp2), d * w1 * (w1p2 + dtp2),
                    T(0)};
    const T m[5] = {adt * w4p4, cdt * (ss1 + 2 * dtp3 * w3 + 2 * dt * (-6 * delta + w3p3)),
                    bdt * ss2, cdt * (ss1 + 2 * dtp3 * w2 + 2 * dt * (+6 * delta + w2p3)),
                    adt * w5p4};
    const T r[5] = {l[4], l[3], l[2], l[1], l[0]};
    for (int k = 0; k < 5; k++)
      s[k] = l[k] * t1 + m[k] * t2 * t3 + r[k] * t4;
  }
}

// ---------------------------------------------------------------------------
// Esirkepov primitives (nix/esirkepov.hpp)
// ---------------------------------------------------------------------------

// nix/esirkepov.hpp:239-270 (scalar branch)
template <int Dim, int Order>
inline void shift_weights(const int shift[Dim], double ss[Dim][Order + 3])
{
  for (int dir = 0; dir < Dim; dir++) {
    if (shift[dir] < 0) {
      for (int ii = 0; ii < Order + 2; ii++)
        ss[dir][ii] = ss[dir][ii + 1];
    } else if (shift[dir] > 0) {
      for (int ii = Order + 2; ii > 0; ii--)
        ss[dir][ii] = ss[dir][ii - 1];
    }
  }
}

// nix/esirkepov.hpp:145-237 (ro3d, ds3d, jx3d, jy3d, jz3d) + deposit3d :323-337
template <int Order>
inline void deposit3d(double dxdt, double dydt, double dzdt, double qs, double ss[2][3][Order + 3],
                      double* cur /* [N][N][N][4] */)
{
  constexpr int N = Order + 3;
  auto          C = [&](int jz, int jy, int jx, int k) -> double& {
    return cur[((jz * N + jy) * N + jx) * 4 + k];
  };
  for (int jz = 0; jz < N; jz++)
    for (int jy = 0; jy < N; jy++)
      for (int jx = 0; jx < N; jx++)
        C(jz, jy, jx, 0) += qs * ss[1][0][jx] * ss[1][1][jy] * ss[1][2][jz];
  for (int dir = 0; dir < 3; dir++)
    for (int l = 0; l < N; l++)
      ss[1][dir][l] -= ss[0][dir][l];
  const double A = 1.0 / 2;
  const double B = 1.0 / 3;
  {
    double qdxdt = qs * dxdt;
    for (int jz = 0; jz < N; jz++)
      for (int jy = 0; jy < N; jy++) {
        double ww = 0;
        double wx = -((1 * ss[0][1][jy] + A * ss[1][1][jy]) * ss[0][2][jz] +
                      (A * ss[0][1][jy] + B * ss[1][1][jy]) * ss[1][2][jz]) *
                    qdxdt;
        for (int jx = 0; jx < N - 1; jx++) {
          ww += ss[1][0][jx] * wx;
          C(jz, jy, jx + 1, 1) += ww;
        }
      }
  }
  {
    double qdydt = qs * dydt;
    for (int jz = 0; jz < N; jz++)
      for (int jx = 0; jx < N; jx++) {
        double ww = 0;
        double wy = -((1 * ss[0][2][jz] + A * ss[1][2][jz]) * ss[0][0][jx] +
                      (A * ss[0][2][jz] + B * ss[1][2][jz]) * ss[1][0][jx]) *
                    qdydt;
        for (int jy = 0; jy < N - 1; jy++) {
          ww += ss[1][1][jy] * wy;
          C(jz, jy + 1, jx, 2) += ww;
        }
      }
  }
  {
    double qdzdt = qs * dzdt;
    for (int jy = 0; jy < N; jy++)
      for (int jx = 0; jx < N; jx++) {
        double ww = 0;
        double wz = -((1 * ss[0][0][jx] + A * ss[1][0][jx]) * ss[0][1][jy] +
                      (A * ss[0][0][jx] + B * ss[1][0][jx]) * ss[1][1][jy]) *
                    qdzdt;
        for (int jz = 0; jz < N - 1; jz++) {
          ww += ss[1][2][jz] * wz;
          C(jz + 1, jy, jx, 3) += ww;
        }
      }
  }
}

// nix/esirkepov.hpp:76-143 (ro2d .. jz2d) + deposit2d :287-302
// nix/esirkepov.hpp:18-67,277-291 (ro1d, ds1d, jx1d, jy1d, jz1d)
template <int Order>
inline void deposit1d(double dxdt, double vy, double vz, double qs, double ss[2][1][Order + 3],
                      double* cur /* [N][4] */)
{
  constexpr int N = Order + 3;
  for (int jx = 0; jx < N; jx++)
    cur[jx * 4 + 0] += qs * ss[1][0][jx];
  for (int l = 0; l < N; l++)
    ss[1][0][l] -= ss[0][0][l];
  const double A = 1.0 / 2;
  {
    double ww = 0;
    double wx = -(qs * dxdt);
    for (int jx = 0; jx < N - 1; jx++) {
      ww += ss[1][0][jx] * wx;
      cur[(jx + 1) * 4 + 1] += ww;
    }
  }
  {
    const double qvy = qs * vy;
    for (int jx = 0; jx < N; jx++)
      cur[jx * 4 + 2] += (ss[0][0][jx] + A * ss[1][0][jx]) * qvy;
  }
  {
    const double qvz = qs * vz;
    for (int jx = 0; jx < N; jx++)
      cur[jx * 4 + 3] += (ss[0][0][jx] + A * ss[1][0][jx]) * qvz;
  }
}

template <int Order>
inline void deposit2d(double dxdt, double dydt, double vz, double qs, double ss[2][2][Order + 3],
                      double* cur /* [N][N][4] */)
{
  constexpr int N = Order + 3;
  auto C = [&](int jy, int jx, int k) -> double& { return cur[(jy * N + jx) * 4 + k]; };
  for (int jy = 0; jy < N; jy++)
    for (int jx = 0; jx < N; jx++)
      C(jy, jx, 0) += qs * ss[1][0][jx] * ss[1][1][jy];
  for (int dir = 0; dir < 2; dir++)
    for (int l = 0; l < N; l++)
      ss[1][dir][l] -= ss[0][dir][l];
  const double A = 1.0 / 2;
  const double B = 1.0 / 3;
  {
    double qdxdt = qs * dxdt;
    for (int jy = 0; jy < N; jy++) {
      double ww = 0;
      double wx = -(ss[0][1][jy] + A * ss[1][1][jy]) * qdxdt;
      for (int jx = 0; jx < N - 1; jx++) {
        ww += ss[1][0][jx] * wx;
        C(jy, jx + 1, 1) += ww;
      }
    }
  }
  {
    double qdydt = qs * dydt;
    for (int jx = 0; jx < N; jx++) {
      double ww = 0;
      double wy = -(ss[0][0][jx] + A * ss[1][0][jx]) * qdydt;
      for (int jy = 0; jy < N - 1; jy++) {
        ww += ss[1][1][jy] * wy;
        C(jy + 1, jx, 2) += ww;
      }
    }
  }
  {
    double qvz = qs * vz;
    for (int jy = 0; jy < N; jy++)
      for (int jx = 0; jx < N; jx++) {
        double wz = ((1 * ss[0][0][jx] + A * ss[1][0][jx]) * ss[0][1][jy] +
                     (A * ss[0][0][jx] + B * ss[1][0][jx]) * ss[1][1][jy]) *
                    qvz;
        C(jy, jx, 3) += wz;
      }
  }
}

// ---------------------------------------------------------------------------
// random numbers: std::mt19937_64 and libstdc++'s uniform/normal distributions
// (GCC 11.4 libstdc++, <bits/random.h>, <bits/random.tcc>); the reference calls
// them at example/thermal/main.cpp:72-75,100-105 through the aliases at
// nix/random.hpp:9-10.
// ---------------------------------------------------------------------------
struct MT64 {
  // ISO C++ [rand.predef]: mersenne_twister_engine<uint_fast64_t, 64, 312, 156, 31,
  //   0xb5026f5aa96619e9, 29, 0x5555555555555555, 17, 0x71d67fffeda60000, 37,
  //   0xfff7eee000000000, 43, 6364136223846793005>
  uint64_t mt[312];
  int      idx;
  explicit MT64(uint64_t seed = 5489u) { this->seed(seed); }
  void     seed(uint64_t s)
  {
    mt[0] = s;
    for (int i = 1; i < 312; i++)
      mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    idx = 312;
  }
  uint64_t operator()()
  {
    if (idx >= 312) {
      const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
      for (int i = 0; i < 312; i++) {
        uint64_t x = (mt[i] & UM) | (mt[(i + 1) % 312] & LM);
        uint64_t y = x >> 1;
        if (x & 1)
          y ^= 0xB5026F5AA96619E9ULL;
        mt[i] = mt[(i + 156) % 312] ^ y;
      }
      idx = 0;
    }
    uint64_t z = mt[idx++];
    z ^= (z >> 29) & 0x5555555555555555ULL;
    z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
    z ^= (z << 37) & 0xFFF7EEE000000000ULL;
    z ^= (z >> 43);
    return z;
  }
};

// libstdc++ generate_canonical<double, 53>(mt19937_64): one 64-bit draw
// (k = max(1, (53 + 64 - 1) / 64) = 1), converted to double and divided by
// 2^64; a result that rounds to 1.0 is replaced by nextafter(1, 0).
inline double canonical(MT64& g)
{
  double sum = static_cast<double>(g());
  double tmp = 18446744073709551616.0; // 2^64
  double ret = sum / tmp;
  if (ret >= 1.0)
    ret = std::nextafter(1.0, 0.0);
  return ret;
}

// libstdc++ uniform_real_distribution<double>(a, b): a + (b - a) * canonical
inline double uniform01(MT64& g) { return canonical(g) * (1.0 - 0.0) + 0.0; }

// libstdc++ normal_distribution<double>: Marsaglia polar method, one cached value
struct Normal {
  bool   saved_available = false;
  double saved           = 0;
  double operator()(MT64& g, double mean = 0.0, double stddev = 1.0)
  {
    double ret;
    if (saved_available) {
      saved_available = false;
      ret             = saved;
    } else {
      double x, y, r2;
      do {
        x  = 2.0 * canonical(g) - 1.0;
        y  = 2.0 * canonical(g) - 1.0;
        r2 = x * x + y * y;
      } while (r2 > 1.0 || r2 == 0.0);
      const double mult = std::sqrt(-2 * std::log(r2) / r2);
      saved             = x * mult;
      saved_available   = true;
      ret               = y * mult;
    }
    return ret * stddev + mean;
  }
};

// libstdc++ gamma_distribution<double> (Marsaglia and Tsang, ACM TOMS 26 (2000)
// 363-372, as written in bits/random.tcc): the normal member keeps its cached
// second value between calls.
struct Gamma {
  double alpha, beta, malpha, a2;
  Normal nd;
  Gamma(double a, double b) : alpha(a), beta(b)
  {
    malpha          = alpha < 1.0 ? alpha + 1.0 : alpha;
    const double a1 = malpha - 1.0 / 3.0;
    a2              = 1.0 / std::sqrt(9.0 * a1);
  }
  double operator()(MT64& g)
  {
    double       u, v, n;
    const double a1 = malpha - 1.0 / 3.0;
    do {
      do {
        n = nd(g);
        v = 1.0 + a2 * n;
      } while (v <= 0.0);
      v = v * v * v;
      u = canonical(g);
    } while (u > 1.0 - 0.0331 * n * n * n * n &&
             (std::log(u) > (0.5 * n * n + a1 * (1.0 - v + std::log(v)))));
    if (alpha == malpha)
      return a1 * v * beta;
    do
      u = canonical(g);
    while (u == 0.0);
    return std::pow(u, 1.0 / alpha) * a1 * v * beta;
  }
};

// libstdc++ poisson_distribution<int> for mean < 12: multiply uniforms until the
// product drops to exp(-mean) (bits/random.tcc; Devroye 1986, Ch. X).  Means of
// 12 and above take libstdc++'s rejection branch, which is not restated here.
inline int poisson_small(MT64& g, double lm_thr)
{
  int    x    = 0;
  double prod = 1.0;
  do {
    prod *= canonical(g);
    x += 1;
  } while (prod > lm_thr);
  return x - 1;
}

// nix::MaxwellJuttner (nix/random.hpp:27-163; Zenitani, Phys. Plasmas 22, 042116
// (2015)): Lorentz factor by the modified Canfield method, isotropic direction,
// then the flipping-method boost along x.
struct MaxwellJuttner {
  static constexpr double a = 0.56;
  static constexpr double b = 0.35;
  double                  temperature, drift;
  double                  table[5];
  std::vector<Gamma>      gamma;
  MaxwellJuttner(double T, double u0) : temperature(T), drift(u0)
  {
    table[0] = 0;
    table[1] = std::sqrt(M_PI);
    table[2] = a * std::sqrt(2 * T);
    table[3] = b * 1.5 * std::sqrt(M_PI) * T;
    table[4] = std::sqrt(2 * T) * (2 * T);
    for (int i = 1; i < 5; i++) // std::partial_sum
      table[i] = table[i - 1] + table[i];
    const double last = table[4];
    for (int i = 0; i < 5; i++) // normalised by the total (the lambda reads back())
      table[i] = table[i] / last;
    gamma = {Gamma(1.5, T), Gamma(2.0, T), Gamma(2.5, T), Gamma(3.0, T)};
  }
  void set_drift(double u0) { drift = u0; }
  void operator()(MT64& g, double& ux, double& uy, double& uz)
  {
    double xx     = -1;
    bool   status = false;
    do {
      double r1    = uniform01(g);
      double r2    = uniform01(g);
      int    index = (int)(std::upper_bound(table, table + 5, r1) - table) - 1;
      xx           = gamma[index](g);
      double R1    = (xx + 1) * std::sqrt(xx + 2);
      double R2 = std::sqrt(2) + a * std::sqrt(xx) + b * std::sqrt(2) * xx + std::pow(xx, 1.5);
      status    = (r2 > R1 / R2);
    } while (status);
    double uu = std::sqrt(xx * (xx + 2));
    double r3 = uniform01(g);
    double r4 = uniform01(g);
    ux        = uu * (2 * r3 - 1);
    uy        = uu * 2 * std::sqrt(r3 * (1 - r3)) * std::cos(2 * M_PI * r4);
    uz        = uu * 2 * std::sqrt(r3 * (1 - r3)) * std::sin(2 * M_PI * r4);
    // lorentz_boost (nix/random.hpp:144-156)
    double rr = uniform01(g);
    double Gm = std::sqrt(1 + drift * drift);
    double Vx = drift / Gm;
    double gm = std::sqrt(1 + ux * ux + uy * uy + uz * uz);
    double vx = ux / gm;
    ux        = (-Vx * vx > rr) ? -ux : +ux;
    ux        = Gm * (ux + Vx * gm);
  }
};

// ---------------------------------------------------------------------------
// chunk map: generalized Hilbert ("gilbert") curve, nix/sfc.cpp:31-141,200-509
// (published algorithm of J. Cerveny, https://github.com/jakubcerveny/gilbert)
// ---------------------------------------------------------------------------
inline int sgn(int x) { return x == 0 ? 0 : (x > 0 ? 1 : -1); }

struct Gilbert {
  int              Nz, Ny, Nx;
  std::vector<int> index; // [z][y][x] -> id
  int              id = 0;
  void             put(int x, int y, int z) { index[(z * Ny + y) * Nx + x] = id++; }

  void g2d(int x, int y, int ax, int ay, int bx, int by, bool xz)
  {
    // 2D curve over a (x, y) plane; when xz is set the second axis is z
    int w = std::abs(ax + ay), h = std::abs(bx + by);
    int dax = sgn(ax), day = sgn(ay), dbx = sgn(bx), dby = sgn(by);
    auto P  = [&](int px, int py) {
      if (xz)
        put(px, 0, py);
      else
        put(px, py, 0);
    };
    if (h == 1) {
      for (int i = 0; i < w; i++) {
        P(x, y);
        x += dax;
        y += day;
      }
      return;
    }
    if (w == 1) {
      for (int i = 0; i < h; i++) {
        P(x, y);
        x += dbx;
        y += dby;
      }
      return;
    }
    int ax2 = ax / 2, ay2 = ay / 2, bx2 = bx / 2, by2 = by / 2;
    int w2 = std::abs(ax2 + ay2), h2 = std::abs(bx2 + by2);
    if (2 * w > 3 * h) {
      if ((w2 % 2) && (w > 2)) {
        ax2 += dax;
        ay2 += day;
      }
      g2d(x, y, ax2, ay2, bx, by, xz);
      g2d(x + ax2, y + ay2, ax - ax2, ay - ay2, bx, by, xz);
    } else {
      if ((h2 % 2) && (h > 2)) {
        bx2 += dbx;
        by2 += dby;
      }
      g2d(x, y, bx2, by2, ax2, ay2, xz);
      g2d(x + bx2, y + by2, ax, ay, bx - bx2, by - by2, xz);
      g2d(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby), -bx2, -by2, -(ax - ax2),
          -(ay - ay2), xz);
    }
  }

  void g3d(int x, int y, int z, int ax, int ay, int az, int bx, int by, int bz, int cx, int cy,
           int cz)
  {
    int w = std::abs(ax + ay + az), h = std::abs(bx + by + bz), d = std::abs(cx + cy + cz);
    int dax = sgn(ax), day = sgn(ay), daz = sgn(az);
    int dbx = sgn(bx), dby = sgn(by), dbz = sgn(bz);
    int dcx = sgn(cx), dcy = sgn(cy), dcz = sgn(cz);
    if (h == 1 && d == 1) {
      for (int i = 0; i < w; i++) {
        put(x, y, z);
        x += dax;
        y += day;
        z += daz;
      }
      return;
    }
    if (d == 1 && w == 1) {
      for (int i = 0; i < h; i++) {
        put(x, y, z);
        x += dbx;
        y += dby;
        z += dbz;
      }
      return;
    }
    if (w == 1 && h == 1) {
      for (int i = 0; i < d; i++) {
        put(x, y, z);
        x += dcx;
        y += dcy;
        z += dcz;
      }
      return;
    }
    int ax2 = ax / 2, ay2 = ay / 2, az2 = az / 2;
    int bx2 = bx / 2, by2 = by / 2, bz2 = bz / 2;
    int cx2 = cx / 2, cy2 = cy / 2, cz2 = cz / 2;
    int w2 = std::abs(ax2 + ay2 + az2), h2 = std::abs(bx2 + by2 + bz2),
        d2 = std::abs(cx2 + cy2 + cz2);
    if ((w2 % 2) && (w > 2)) {
      ax2 += dax;
      ay2 += day;
      az2 += daz;
    }
    if ((h2 % 2) && (h > 2)) {
      bx2 += dbx;
      by2 += dby;
      bz2 += dbz;
    }
    if ((d2 % 2) && (d > 2)) {
      cx2 += dcx;
      cy2 += dcy;
      cz2 += dcz;
    }
    int ax3 = ax - ax2, ay3 = ay - ay2, az3 = az - az2;
    int bx3 = bx - bx2, by3 = by - by2, bz3 = bz - bz2;
    int cx3 = cx - cx2, cy3 = cy - cy2, cz3 = cz - cz2;
    if ((2 * w > 3 * h) && (2 * w > 3 * d)) {
      g3d(x, y, z, ax2, ay2, az2, bx, by, bz, cx, cy, cz);
      g3d(x + ax2, y + ay2, z + az2, ax3, ay3, az3, bx, by, bz, cx, cy, cz);
    } else if (3 * h > 4 * d) {
      g3d(x, y, z, bx2, by2, bz2, cx, cy, cz, ax2, ay2, az2);
      g3d(x + bx2, y + by2, z + bz2, ax, ay, az, bx3, by3, bz3, cx, cy, cz);
      int x3 = x + bx2 + ax - (dax + dbx), y3 = y + by2 + ay - (day + dby),
          z3 = z + bz2 + az - (daz + dbz);
      g3d(x3, y3, z3, -bx2, -by2, -bz2, cx, cy, cz, -ax3, -ay3, -az3);
    } else if (3 * d > 4 * h) {
      g3d(x, y, z, cx2, cy2, cz2, ax2, ay2, az2, bx, by, bz);
      g3d(x + cx2, y + cy2, z + cz2, ax, ay, az, bx, by, bz, cx3, cy3, cz3);
      int x3 = x + cx2 + ax - (dax + dcx), y3 = y + cy2 + ay - (day + dcy),
          z3 = z + cz2 + az - (daz + dcz);
      g3d(x3, y3, z3, -cx2, -cy2, -cz2, -ax3, -ay3, -az3, bx, by, bz);
    } else {
      g3d(x, y, z, bx2, by2, bz2, cx2, cy2, cz2, ax2, ay2, az2);
      int x2 = x + bx2, y2 = y + by2, z2 = z + bz2;
      g3d(x2, y2, z2, cx, cy, cz, ax2, ay2, az2, bx3, by3, bz3);
      int x3 = x2 + cx - (dbx + dcx), y3 = y2 + cy - (dby + dcy), z3 = z2 + cz - (dbz + dcz);
      g3d(x3, y3, z3, ax, ay, az, -bx2, -by2, -bz2, -cx3, -cy3, -cz3);
      int x4 = x3 + ax - (dax - dbx), y4 = y3 + ay - (day - dby), z4 = z3 + az - (daz - dbz);
      g3d(x4, y4, z4, -cx, -cy, -cz, -ax3, -ay3, -az3, bx3, by3, bz3);
      int x5 = x4 - cx - (dbx - dcx), y5 = y4 - cy - (dby - dcy), z5 = z4 - cz - (dbz - dcz);
      g3d(x5, y5, z5, -bx2, -by2, -bz2, cx2, cy2, cz2, -ax3, -ay3, -az3);
    }
  }
};

// nix/sfc.cpp:100-141 (get_map3d) with the 1D / 2D special cases
std::vector<int> chunk_map(int Cz, int Cy, int Cx)
{
  Gilbert g{Cz, Cy, Cx, std::vector<int>(Cz * Cy * Cx, 0)};
  if (Cz != 1 && Cy != 1 && Cx != 1) {
    if (Cx >= Cy && Cx >= Cz)
      g.g3d(0, 0, 0, Cx, 0, 0, 0, Cy, 0, 0, 0, Cz);
    else if (Cy >= Cx && Cy >= Cz)
      g.g3d(0, 0, 0, 0, Cy, 0, Cx, 0, 0, 0, 0, Cz);
    else
      g.g3d(0, 0, 0, 0, 0, Cz, Cx, 0, 0, 0, Cy, 0);
  } else if (Cz == 1 && Cy != 1 && Cx != 1) {
    // nix/sfc.cpp:43-55 via get_map2d(Ny=Cy, Nx=Cx)
    if (Cx >= Cy)
      g.g2d(0, 0, Cx, 0, 0, Cy, false);
    else
      g.g2d(0, 0, 0, Cy, Cx, 0, false);
  } else if (Cz != 1 && Cy == 1 && Cx != 1) {
    if (Cx >= Cz)
      g.g2d(0, 0, Cx, 0, 0, Cz, true);
    else
      g.g2d(0, 0, 0, Cz, Cx, 0, true);
  } else if (Cz != 1 && Cy != 1 && Cx == 1) {
    // (y, z) plane: walk as a 2D map over (Nz, Ny) with "x" = y
    Gilbert h{1, Cz, Cy, std::vector<int>(Cz * Cy, 0)};
    if (Cy >= Cz)
      h.g2d(0, 0, Cy, 0, 0, Cz, false);
    else
      h.g2d(0, 0, 0, Cz, Cy, 0, false);
    for (int iz = 0; iz < Cz; iz++)
      for (int iy = 0; iy < Cy; iy++)
        g.index[iz * Cy + iy] = h.index[iz * Cy + iy];
  } else {
    // 1D maps (or a single chunk): id = linear coordinate (nix/sfc.cpp:31-37)
    for (int i = 0; i < Cz * Cy * Cx; i++)
      g.index[i] = i;
  }
  return g.index;
}


// ---------------------------------------------------------------------------
// particle container (nix/particle.hpp, nix/xtensor_particle.hpp)
// ---------------------------------------------------------------------------
struct Particles {
  int                  Np = 0;
  int                  Ng = 0;
  double               q = 0, m = 0;
  std::vector<double>  xu, xv; // [cap][7]
  std::vector<int32_t> gindex; // [cap]
  std::vector<int32_t> pindex; // [Ng + 1]
  std::vector<int32_t> pcount; // [Ng + 1][8]

  void reserve(int n)
  {
    if ((size_t)n * NC > xu.size()) {
      xu.resize((size_t)n * NC, 0.0);
      xv.resize((size_t)n * NC, 0.0);
      gindex.resize(n, 0);
    }
  }
  double* XU(int ip) { return &xu[(size_t)ip * NC]; }
  double* XV(int ip) { return &xv[(size_t)ip * NC]; }
};

// ---------------------------------------------------------------------------
// chunk (nix/chunk.hpp/.cpp, pic/pic_chunk.hpp/.cpp)
// ---------------------------------------------------------------------------
struct Chunk {
  int    id;
  int    coord[3];  // (cz, cy, cx)
  int    offset[3]; // cell offset (z, y, x)
  int    dims[3];   // interior cells (z, y, x)
  bool   has_dim[3];
  int    Nb;
  int    Lb[3], Ub[3]; // (z, y, x)
  int    n[3];         // allocated extent incl. ghosts (z, y, x)
  double del[3];       // cell size (z, y, x)
  double lim[3][2];    // physical range (z, y, x), nix/chunk.cpp:210-237
  double glim[3][2];   // global physical range
  double rng[3][2];    // get_{z,y,x}range: +-max for absent dimensions (nix/chunk.hpp:295-335)
  int    nb_id[27]; // -1: no neighbour (a physical boundary, MPI_PROC_NULL)

  // example/shock: physical walls in x and the state the example keeps in
  // option["random_seed"] / option["boundary"] (example/shock/main.cpp:19-24,212-226)
  bool    wall_lo = false, wall_hi = false;
  int     random_seed = 0;
  int64_t lastid      = 0;
  int     influx[2] = {0, 0}, efflux[2] = {0, 0};
  MT64    mt;

  std::vector<double>    uf, uj, ff, um;
  std::vector<Particles> up;

  size_t cell(int iz, int iy, int ix) const { return ((size_t)iz * n[1] + iy) * n[2] + ix; }
  double& UF(int iz, int iy, int ix, int k) { return uf[cell(iz, iy, ix) * 6 + k]; }
  double& UJ(int iz, int iy, int ix, int k) { return uj[cell(iz, iy, ix) * 4 + k]; }
  double& FF(int iz, int iy, int ix, int l, int k) { return ff[(cell(iz, iy, ix) * 3 + l) * 3 + k]; }
  double& UM(int iz, int iy, int ix, int is, int k, int Ns)
  {
    return um[(cell(iz, iy, ix) * Ns + is) * 14 + k];
  }

  // send/recv bounds, nix/chunk.cpp:118-208 (set_index_bounds)
  void bounds(int d, int idx, bool send, int& lo, int& hi) const
  {
    if (send) {
      lo = (idx == 2) ? Ub[d] - Nb + 1 : Lb[d];
      hi = (idx == 0) ? Lb[d] + Nb - 1 : Ub[d];
    } else {
      lo = (idx == 0) ? Lb[d] - Nb : (idx == 1 ? Lb[d] : Ub[d] + 1);
      hi = (idx == 0) ? Lb[d] - 1 : (idx == 1 ? Ub[d] : Ub[d] + Nb);
    }
  }
  int ilo(int d) const { return has_dim[d] ? 0 : 1; }
  int ihi(int d) const { return has_dim[d] ? 2 : 1; }
};

// example/shock/main.cpp:56-92 derived constants
struct ShockParams {
  int    nppc = 0;
  double delt = 0, u0 = 0, vte = 0, vti = 0, taper = 0;
  double E0[3] = {0, 0, 0}, B0[3] = {0, 0, 0};
  double me = 0, qe = 0, mi = 0, qi = 0;
};

struct Sim {
  ShockParams        sh;
  pno_params         p;
  int                dim;
  int                order;
  int                Ns;
  std::vector<Chunk> chunks; // indexed by chunk id
  std::vector<int>   id_of_coord;
};

int g_threads = 0;

template <typename F>
void for_chunks(Sim& s, F f)
{
  const int nc = (int)s.chunks.size();
#ifdef _OPENMP
  int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
  for (int i = 0; i < nc; i++)
    f(s.chunks[i]);
}

// --------------------------------------------------------------------------
// count / sort / periodic wrap (nix/xtensor_particle.hpp:231-376)
// --------------------------------------------------------------------------
inline int flatindex(const Chunk& c, int iz, int iy, int ix)
{
  const int sx = 1;
  const int sy = sx * (c.Ub[2] - c.Lb[2] + 2);
  const int sz = sy * (c.Ub[1] - c.Lb[1] + 2);
  return iz * sz + iy * sy + ix * sx;
}

// nix/xtensor_particle.hpp:324-357
void count(const Chunk& c, Particles& P, int lbp, int ubp, bool reset, int order)
{
  const int    is_odd  = (order % 2 == 1) ? 1 : 0;
  const int    oob     = P.Ng;
  const double xoffset = c.lim[2][0] - 0.5 * c.del[2] * is_odd;
  const double yoffset = c.lim[1][0] - 0.5 * c.del[1] * is_odd;
  const double zoffset = c.lim[0][0] - 0.5 * c.del[0] * is_odd;
  const double rdx = 1 / c.del[2], rdy = 1 / c.del[1], rdz = 1 / c.del[0];
  const double xmin = c.rng[2][0], xmax = c.rng[2][1];
  const double ymin = c.rng[1][0], ymax = c.rng[1][1];
  const double zmin = c.rng[0][0], zmax = c.rng[0][1];

  if (reset)
    std::fill(P.pcount.begin(), P.pcount.end(), 0);

  for (int ip = lbp; ip <= ubp; ip++) {
    const double* x  = P.XU(ip);
    int           ix = c.has_dim[2] ? digitize(x[0], xoffset, rdx) : 0;
    int           iy = c.has_dim[1] ? digitize(x[1], yoffset, rdy) : 0;
    int           iz = c.has_dim[0] ? digitize(x[2], zoffset, rdz) : 0;
    int           ii = flatindex(c, iz, iy, ix);
    ii = c.has_dim[2] && (x[0] < xmin || x[0] >= xmax) ? oob : ii;
    ii = c.has_dim[1] && (x[1] < ymin || x[1] >= ymax) ? oob : ii;
    ii = c.has_dim[0] && (x[2] < zmin || x[2] >= zmax) ? oob : ii;
    // increment (xtensor_particle.hpp:245-252)
    P.gindex[ip] = ii;
    P.pcount[(size_t)ii * SIMD_WIDTH + ip % SIMD_WIDTH]++;
  }
}

// nix/xtensor_particle.hpp:260-321
void sort(Particles& P)
{
  const int Ng = P.Ng;
  auto      PC = [&](int ii, int jj) -> int32_t& { return P.pcount[(size_t)ii * SIMD_WIDTH + jj]; };
  for (int ii = 0; ii < Ng + 1; ii++)
    for (int jj = 0; jj < SIMD_WIDTH - 1; jj++)
      PC(ii, jj + 1) += PC(ii, jj);
  for (int ii = 0; ii < Ng; ii++)
    for (int jj = 0; jj < SIMD_WIDTH; jj++)
      PC(ii + 1, jj) += PC(ii, SIMD_WIDTH - 1);
  P.pindex[0] = 0;
  for (int ii = 0; ii < Ng; ii++)
    P.pindex[ii + 1] = PC(ii, SIMD_WIDTH - 1);
  for (int ii = 0; ii < Ng + 1; ii++)
    for (int jj = SIMD_WIDTH - 1; jj > 0; jj--)
      PC(ii, jj) = PC(ii, jj - 1);
  for (int ii = 0; ii < Ng + 1; ii++)
    PC(ii, 0) = P.pindex[ii];
  for (int ip = 0; ip < P.Np; ip++) {
    int ii = P.gindex[ip];
    int jj = ip % SIMD_WIDTH;
    int jp = PC(ii, jj);
    std::memcpy(P.XV(jp), P.XU(ip), NC * sizeof(double));
    PC(ii, jj)++;
  }
  std::swap(P.xu, P.xv);
  P.Np = P.pindex[Ng];
}

// nix/xtensor_particle.hpp:359-376
void set_boundary_periodic(const Chunk& c, Particles& P, int lbp, int ubp)
{
  const double X1 = c.glim[2][0], X2 = c.glim[2][1];
  const double Y1 = c.glim[1][0], Y2 = c.glim[1][1];
  const double Z1 = c.glim[0][0], Z2 = c.glim[0][1];
  const double X  = c.has_dim[2] * (X2 - X1);
  const double Y  = c.has_dim[1] * (Y2 - Y1);
  const double Z  = c.has_dim[0] * (Z2 - Z1);
  for (int ip = lbp; ip <= ubp; ip++) {
    double* x = P.XU(ip);
    x[0] += (x[0] < X1) * X - (x[0] >= X2) * X;
    x[1] += (x[1] < Y1) * Y - (x[1] >= Y2) * Y;
    x[2] += (x[2] < Z1) * Z - (x[2] >= Z2) * Z;
  }
}

// --------------------------------------------------------------------------
// Maxwell solver (pic/engine/maxwell.hpp)
// --------------------------------------------------------------------------
// pic/engine/maxwell.hpp:44-61
void init_friedman(Chunk& c)
{
  const int Nb = c.Nb;
  for (int iz = c.Lb[0] - Nb; iz <= c.Ub[0] + Nb; iz++)
    for (int iy = c.Lb[1] - Nb; iy <= c.Ub[1] + Nb; iy++)
      for (int ix = c.Lb[2] - Nb; ix <= c.Ub[2] + Nb; ix++)
        for (int dir = 0; dir < 3; dir++) {
          c.FF(iz, iy, ix, 0, dir) = c.UF(iz, iy, ix, dir);
          c.FF(iz, iy, ix, 1, dir) = c.UF(iz, iy, ix, dir);
          c.FF(iz, iy, ix, 2, dir) = c.UF(iz, iy, ix, dir);
        }
}

// pic/engine/maxwell.hpp:148-201 (1D), :292-348 (2D) and :442-496 (3D)
void push_bfd(Chunk& c, int dim, double cc, double theta, double delt)
{
  const int    Nb   = c.Nb;
  const double A    = 1 + 0.5 * theta;
  const double B    = -theta * (1 - 0.5 * theta);
  const double C    = 0.5 * theta * (1 - theta) * (1 - theta);
  const double cflx = cc * delt / c.del[2];
  const double cfly = cc * delt / c.del[1];
  const double cflz = cc * delt / c.del[0];
  const int    lbx = c.Lb[2], ubx = c.Ub[2], lby = c.Lb[1], uby = c.Ub[1], lbz = c.Lb[0],
            ubz = c.Ub[0];
  const int zlo = dim == 3 ? lbz - Nb : lbz, zhi = dim == 3 ? ubz + Nb : lbz;
  const int ylo = dim >= 2 ? lby - Nb : lby, yhi = dim >= 2 ? uby + Nb : lby;

  for (int iz = zlo; iz <= zhi; iz++)
    for (int iy = ylo; iy <= yhi; iy++)
      for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
        for (int k = 0; k < 3; k++)
          c.FF(iz, iy, ix, 0, k) =
              A * c.UF(iz, iy, ix, k) + B * c.FF(iz, iy, ix, 1, k) + C * c.FF(iz, iy, ix, 2, k);

  if (dim == 3) {
    for (int iz = lbz - Nb + 1; iz <= ubz + Nb; iz++)
      for (int iy = lby - Nb + 1; iy <= uby + Nb; iy++)
        for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
          c.UF(iz, iy, ix, 3) += (-cfly) * (c.FF(iz, iy, ix, 0, 2) - c.FF(iz, iy - 1, ix, 0, 2)) +
                                 (+cflz) * (c.FF(iz, iy, ix, 0, 1) - c.FF(iz - 1, iy, ix, 0, 1));
    for (int iz = lbz - Nb + 1; iz <= ubz + Nb; iz++)
      for (int iy = lby - Nb; iy <= uby + Nb; iy++)
        for (int ix = lbx - Nb + 1; ix <= ubx + Nb; ix++)
          c.UF(iz, iy, ix, 4) += (-cflz) * (c.FF(iz, iy, ix, 0, 0) - c.FF(iz - 1, iy, ix, 0, 0)) +
                                 (+cflx) * (c.FF(iz, iy, ix, 0, 2) - c.FF(iz, iy, ix - 1, 0, 2));
    for (int iz = lbz - Nb; iz <= ubz + Nb; iz++)
      for (int iy = lby - Nb + 1; iy <= uby + Nb; iy++)
        for (int ix = lbx - Nb + 1; ix <= ubx + Nb; ix++)
          c.UF(iz, iy, ix, 5) += (-cflx) * (c.FF(iz, iy, ix, 0, 1) - c.FF(iz, iy, ix - 1, 0, 1)) +
                                 (+cfly) * (c.FF(iz, iy, ix, 0, 0) - c.FF(iz, iy - 1, ix, 0, 0));
  } else if (dim == 1) {
    const int iz = lbz, iy = lby;
    for (int ix = lbx - Nb + 1; ix <= ubx + Nb; ix++)
      c.UF(iz, iy, ix, 4) += (+cflx) * (c.FF(iz, iy, ix, 0, 2) - c.FF(iz, iy, ix - 1, 0, 2));
    for (int ix = lbx - Nb + 1; ix <= ubx + Nb; ix++)
      c.UF(iz, iy, ix, 5) += (-cflx) * (c.FF(iz, iy, ix, 0, 1) - c.FF(iz, iy, ix - 1, 0, 1));
  } else {
    const int iz = lbz;
    for (int iy = lby - Nb + 1; iy <= uby + Nb; iy++)
      for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
        c.UF(iz, iy, ix, 3) += (-cfly) * (c.FF(iz, iy, ix, 0, 2) - c.FF(iz, iy - 1, ix, 0, 2));
    for (int iy = lby - Nb; iy <= uby + Nb; iy++)
      for (int ix = lbx - Nb + 1; ix <= ubx + Nb; ix++)
        c.UF(iz, iy, ix, 4) += (+cflx) * (c.FF(iz, iy, ix, 0, 2) - c.FF(iz, iy, ix - 1, 0, 2));
    for (int iy = lby - Nb + 1; iy <= uby + Nb; iy++)
      for (int ix = lbx - Nb + 1; ix <= ubx + Nb; ix++)
        c.UF(iz, iy, ix, 5) += (-cflx) * (c.FF(iz, iy, ix, 0, 1) - c.FF(iz, iy, ix - 1, 0, 1)) +
                               (+cfly) * (c.FF(iz, iy, ix, 0, 0) - c.FF(iz, iy - 1, ix, 0, 0));
  }
}

// pic/engine/maxwell.hpp:85-146 (1D), :231-290 (2D) and :383-440 (3D)
void push_efd(Chunk& c, int dim, double cc, double theta, double delt)
{
  const int    Nb   = c.Nb;
  const double cflx = cc * delt / c.del[2];
  const double cfly = cc * delt / c.del[1];
  const double cflz = cc * delt / c.del[0];
  const int    lbx = c.Lb[2], ubx = c.Ub[2], lby = c.Lb[1], uby = c.Ub[1], lbz = c.Lb[0],
            ubz = c.Ub[0];
  const int zlo = dim == 3 ? lbz - Nb : lbz, zhi = dim == 3 ? ubz + Nb : lbz;
  const int ylo = dim >= 2 ? lby - Nb : lby, yhi = dim >= 2 ? uby + Nb : lby;

  for (int iz = zlo; iz <= zhi; iz++)
    for (int iy = ylo; iy <= yhi; iy++)
      for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
        for (int k = 0; k < 3; k++) {
          c.FF(iz, iy, ix, 2, k) = c.FF(iz, iy, ix, 1, k) + theta * c.FF(iz, iy, ix, 2, k);
          c.FF(iz, iy, ix, 1, k) = c.UF(iz, iy, ix, k);
        }

  if (dim == 3) {
    for (int iz = lbz - Nb; iz <= ubz + Nb - 1; iz++)
      for (int iy = lby - Nb; iy <= uby + Nb - 1; iy++)
        for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
          c.UF(iz, iy, ix, 0) += (+cfly) * (c.UF(iz, iy + 1, ix, 5) - c.UF(iz, iy, ix, 5)) +
                                 (-cflz) * (c.UF(iz + 1, iy, ix, 4) - c.UF(iz, iy, ix, 4)) -
                                 delt * c.UJ(iz, iy, ix, 1);
    for (int iz = lbz - Nb; iz <= ubz + Nb - 1; iz++)
      for (int iy = lby - Nb; iy <= uby + Nb; iy++)
        for (int ix = lbx - Nb; ix <= ubx + Nb - 1; ix++)
          c.UF(iz, iy, ix, 1) += (+cflz) * (c.UF(iz + 1, iy, ix, 3) - c.UF(iz, iy, ix, 3)) +
                                 (-cflx) * (c.UF(iz, iy, ix + 1, 5) - c.UF(iz, iy, ix, 5)) -
                                 delt * c.UJ(iz, iy, ix, 2);
    for (int iz = lbz - Nb; iz <= ubz + Nb; iz++)
      for (int iy = lby - Nb; iy <= uby + Nb - 1; iy++)
        for (int ix = lbx - Nb; ix <= ubx + Nb - 1; ix++)
          c.UF(iz, iy, ix, 2) += (+cflx) * (c.UF(iz, iy, ix + 1, 4) - c.UF(iz, iy, ix, 4)) +
                                 (-cfly) * (c.UF(iz, iy + 1, ix, 3) - c.UF(iz, iy, ix, 3)) -
                                 delt * c.UJ(iz, iy, ix, 3);
  } else if (dim == 1) {
    const int iz = lbz, iy = lby;
    for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
      c.UF(iz, iy, ix, 0) += -delt * c.UJ(iz, iy, ix, 1);
    for (int ix = lbx - Nb; ix <= ubx + Nb - 1; ix++)
      c.UF(iz, iy, ix, 1) +=
          (-cflx) * (c.UF(iz, iy, ix + 1, 5) - c.UF(iz, iy, ix, 5)) - delt * c.UJ(iz, iy, ix, 2);
    for (int ix = lbx - Nb; ix <= ubx + Nb - 1; ix++)
      c.UF(iz, iy, ix, 2) +=
          (+cflx) * (c.UF(iz, iy, ix + 1, 4) - c.UF(iz, iy, ix, 4)) - delt * c.UJ(iz, iy, ix, 3);
  } else {
    const int iz = lbz;
    for (int iy = lby - Nb; iy <= uby + Nb - 1; iy++)
      for (int ix = lbx - Nb; ix <= ubx + Nb; ix++)
        c.UF(iz, iy, ix, 0) += (+cfly) * (c.UF(iz, iy + 1, ix, 5) - c.UF(iz, iy, ix, 5)) -
                               delt * c.UJ(iz, iy, ix, 1);
    for (int iy = lby - Nb; iy <= uby + Nb; iy++)
      for (int ix = lbx - Nb; ix <= ubx + Nb - 1; ix++)
        c.UF(iz, iy, ix, 1) += (-cflx) * (c.UF(iz, iy, ix + 1, 5) - c.UF(iz, iy, ix, 5)) -
                               delt * c.UJ(iz, iy, ix, 2);
    for (int iy = lby - Nb; iy <= uby + Nb - 1; iy++)
      for (int ix = lbx - Nb; ix <= ubx + Nb - 1; ix++)
        c.UF(iz, iy, ix, 2) += (+cflx) * (c.UF(iz, iy, ix + 1, 4) - c.UF(iz, iy, ix, 4)) +
                               (-cfly) * (c.UF(iz, iy + 1, ix, 3) - c.UF(iz, iy, ix, 3)) -
                               delt * c.UJ(iz, iy, ix, 3);
  }
}

// pic/engine/maxwell.hpp:63-83 (1D), :203-229 (2D), :350-381 (3D); the x buffers are
// zero for periodic runs
void get_diverror(Chunk& c, int dim, double& efd, double& bfd)
{
  const double rdx = 1 / c.del[2], rdy = 1 / c.del[1], rdz = 1 / c.del[0];
  efd = 0;
  bfd = 0;
  // cells next to a physical boundary are skipped (pic/pic_engine.hpp:42-71)
  const int lbx = c.Lb[2] + (c.wall_lo ? c.Nb : 0);
  const int ubx = c.Ub[2] - (c.wall_hi ? c.Nb : 0);
  if (dim == 3) {
    for (int iz = c.Lb[0]; iz <= c.Ub[0]; iz++)
      for (int iy = c.Lb[1]; iy <= c.Ub[1]; iy++)
        for (int ix = lbx; ix <= ubx; ix++) {
          efd += (c.UF(iz, iy, ix + 1, 0) - c.UF(iz, iy, ix, 0)) * rdx +
                 (c.UF(iz, iy + 1, ix, 1) - c.UF(iz, iy, ix, 1)) * rdy +
                 (c.UF(iz + 1, iy, ix, 2) - c.UF(iz, iy, ix, 2)) * rdz - c.UJ(iz, iy, ix, 0);
          bfd += (c.UF(iz, iy, ix, 3) - c.UF(iz, iy, ix - 1, 3)) * rdx +
                 (c.UF(iz, iy, ix, 4) - c.UF(iz, iy - 1, ix, 4)) * rdy +
                 (c.UF(iz, iy, ix, 5) - c.UF(iz - 1, iy, ix, 5)) * rdz;
        }
  } else if (dim == 1) {
    const int iz = c.Lb[0], iy = c.Lb[1];
    for (int ix = lbx; ix <= ubx; ix++) {
      efd += (c.UF(iz, iy, ix + 1, 0) - c.UF(iz, iy, ix, 0)) * rdx - c.UJ(iz, iy, ix, 0);
      bfd += (c.UF(iz, iy, ix, 3) - c.UF(iz, iy, ix - 1, 3)) * rdx;
    }
  } else {
    const int iz = c.Lb[0];
    for (int iy = c.Lb[1]; iy <= c.Ub[1]; iy++)
      for (int ix = lbx; ix <= ubx; ix++) {
        efd += (c.UF(iz, iy, ix + 1, 0) - c.UF(iz, iy, ix, 0)) * rdx +
               (c.UF(iz, iy + 1, ix, 1) - c.UF(iz, iy, ix, 1)) * rdy - c.UJ(iz, iy, ix, 0);
        bfd += (c.UF(iz, iy, ix, 3) - c.UF(iz, iy, ix - 1, 3)) * rdx +
               (c.UF(iz, iy, ix, 4) - c.UF(iz, iy - 1, ix, 4)) * rdy;
      }
  }
}

// pic/pic_chunk.cpp:407-439
void get_energy(Chunk& c, int Ns, double cc, double& efd, double& bfd, double* particle)
{
  efd = 0.0;
  bfd = 0.0;
  for (int is = 0; is < Ns; is++)
    particle[is] = 0.0;
  for (int iz = c.Lb[0]; iz <= c.Ub[0]; iz++)
    for (int iy = c.Lb[1]; iy <= c.Ub[1]; iy++)
      for (int ix = c.Lb[2]; ix <= c.Ub[2]; ix++) {
        efd += 0.5 * (c.UF(iz, iy, ix, 0) * c.UF(iz, iy, ix, 0) +
                      c.UF(iz, iy, ix, 1) * c.UF(iz, iy, ix, 1) +
                      c.UF(iz, iy, ix, 2) * c.UF(iz, iy, ix, 2));
        bfd += 0.5 * (c.UF(iz, iy, ix, 3) * c.UF(iz, iy, ix, 3) +
                      c.UF(iz, iy, ix, 4) * c.UF(iz, iy, ix, 4) +
                      c.UF(iz, iy, ix, 5) * c.UF(iz, iy, ix, 5));
      }
  for (int iz = c.Lb[0]; iz <= c.Ub[0]; iz++)
    for (int iy = c.Lb[1]; iy <= c.Ub[1]; iy++)
      for (int ix = c.Lb[2]; ix <= c.Ub[2]; ix++)
        for (int is = 0; is < Ns; is++)
          particle[is] += c.UM(iz, iy, ix, is, 4, Ns) * cc - c.UM(iz, iy, ix, is, 0, Ns) * cc * cc;
}

// --------------------------------------------------------------------------
// particle operators (pic/engine/velocity.hpp, position.hpp, current.hpp, moment.hpp)
// --------------------------------------------------------------------------
struct Geo {
  // per-chunk constants shared by the particle engines (BaseVelocity ctor, velocity.hpp:61-79)
  int    lbx, lby, lbz;
  double dx, dy, dz, xmin, ymin, zmin;
  explicit Geo(const Chunk& c)
  {
    lbx  = c.Lb[2];
    lby  = c.Lb[1];
    lbz  = c.Lb[0];
    dx   = c.del[2];
    dy   = c.del[1];
    dz   = c.del[0];
    xmin = c.lim[2][0];
    ymin = c.lim[1][0];
    zmin = c.lim[0][0];
  }
};

// weights (velocity.hpp:223-359, MC interpolation) and the scalar gather
// (velocity.hpp:420-450 push_scalar3d; :390-418 push_scalar2d; nix/interp.hpp:14-130)
template <int Dim, int Order>
void push_velocity_chunk(Chunk& c, int Ns, double cc, int pusher, int interp, double delt)
{
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int S      = Order + 1; // points used by the scalar gather (Stencil + 2)
  Geo           g(c);
  for (int is = 0; is < Ns; is++) {
    Particles& P    = c.up[is];
    double     qmdt = 0.5 * P.q / P.m * delt;
    for (int ip = 0; ip < P.Np; ip++) {
      double* xu     = P.XU(ip);
      double  wix[S + 1] = {0}, wiy[S + 1] = {0}, wiz[S + 1] = {0};
      double  whx[S + 1] = {0}, why[S + 1] = {0}, whz[S + 1] = {0};

      double ximin  = g.xmin + 0.5 * g.dx * is_odd;
      double xhmin  = g.xmin + 0.5 * g.dx * is_odd - 0.5 * g.dx;
      double yimin  = g.ymin + 0.5 * g.dy * is_odd;
      double yhmin  = g.ymin + 0.5 * g.dy * is_odd - 0.5 * g.dy;
      double zimin  = g.zmin + 0.5 * g.dz * is_odd;
      double zhmin  = g.zmin + 0.5 * g.dz * is_odd - 0.5 * g.dz;
      double xigrid = g.xmin + 0.5 * g.dx, xhgrid = g.xmin;
      double yigrid = g.ymin + 0.5 * g.dy, yhgrid = g.ymin;
      double zigrid = g.zmin + 0.5 * g.dz, zhgrid = g.zmin;
      double rdx = 1 / g.dx, rdy = 1 / g.dy, rdz = 1 / g.dz;

      int ix0 = digitize(xu[0], ximin, rdx);
      int iy0 = digitize(xu[1], yimin, rdy);
      int hx0 = digitize(xu[0], xhmin, rdx);
      int hy0 = digitize(xu[1], yhmin, rdy);
      // integer-grid weights: MC, or WT with c dt / dx (velocity.hpp:240-250, 283-293)
      const double cdtx = cc * delt / g.dx, cdty = cc * delt / g.dy;
      if (interp == 1) {
        shape_wt<Order>(xu[0], xigrid + static_cast<double>(ix0) * g.dx, rdx, cdtx, 1 / cdtx, wix);
        shape_wt<Order>(xu[1], yigrid + static_cast<double>(iy0) * g.dy, rdy, cdty, 1 / cdty, wiy);
      } else {
        shape_mc<Order>(xu[0], xigrid + static_cast<double>(ix0) * g.dx, rdx, wix);
        shape_mc<Order>(xu[1], yigrid + static_cast<double>(iy0) * g.dy, rdy, wiy);
      }
      shape_mc<Order>(xu[0], xhgrid + static_cast<double>(hx0) * g.dx, rdx, whx);
      shape_mc<Order>(xu[1], yhgrid + static_cast<double>(hy0) * g.dy, rdy, why);
      ix0 += g.lbx - (Order / 2);
      iy0 += g.lby - (Order / 2);
      hx0 += g.lbx - (Order / 2);
      hy0 += g.lby - (Order / 2);

      double eb[6];
      if constexpr (Dim == 3) {
        int iz0 = digitize(xu[2], zimin, rdz);
        int hz0 = digitize(xu[2], zhmin, rdz);
        const double cdtz = cc * delt / g.dz;
        if (interp == 1)
          shape_wt<Order>(xu[2], zigrid + static_cast<double>(iz0) * g.dz, rdz, cdtz, 1 / cdtz, wiz);
        else
          shape_mc<Order>(xu[2], zigrid + static_cast<double>(iz0) * g.dz, rdz, wiz);
        shape_mc<Order>(xu[2], zhgrid + static_cast<double>(hz0) * g.dz, rdz, whz);
        iz0 += g.lbz - (Order / 2);
        hz0 += g.lbz - (Order / 2);
        // nix/interp.hpp:89-110 (interp3d_impl_sorted)
        auto interp = [&](int z0, int y0, int x0, int k, const double* wz, const double* wy,
                          const double* wx) {
          double rz = 0;
          for (int jz = 0; jz < S; jz++) {
            double ry = 0;
            for (int jy = 0; jy < S; jy++) {
              double rx = 0;
              for (int jx = 0; jx < S; jx++)
                rx += c.UF(z0 + jz, y0 + jy, x0 + jx, k) * wx[jx];
              ry += rx * wy[jy];
            }
            rz += ry * wz[jz];
          }
          return rz * qmdt;
        };
        eb[0] = interp(iz0, iy0, hx0, 0, wiz, wiy, whx);
        eb[1] = interp(iz0, hy0, ix0, 1, wiz, why, wix);
        eb[2] = interp(hz0, iy0, ix0, 2, whz, wiy, wix);
        eb[3] = interp(hz0, hy0, ix0, 3, whz, why, wix);
        eb[4] = interp(hz0, iy0, hx0, 4, whz, wiy, whx);
        eb[5] = interp(iz0, hy0, hx0, 5, wiz, why, whx);
      } else if constexpr (Dim == 1) {
        const int iz0 = g.lbz, iy1 = g.lby;
        // nix/interp.hpp:14-26 (interp1d_impl_sorted), velocity.hpp:362-388 (push_scalar1d)
        auto interp = [&](int x0, int k, const double* wx) {
          double rx = 0;
          for (int jx = 0; jx < S; jx++)
            rx += c.UF(iz0, iy1, x0 + jx, k) * wx[jx];
          return rx * qmdt;
        };
        eb[0] = interp(hx0, 0, whx);
        eb[1] = interp(ix0, 1, wix);
        eb[2] = interp(ix0, 2, wix);
        eb[3] = interp(ix0, 3, wix);
        eb[4] = interp(hx0, 4, whx);
        eb[5] = interp(hx0, 5, whx);
      } else {
        const int iz0 = g.lbz;
        // nix/interp.hpp:50-63 (interp2d_impl_sorted)
        auto interp = [&](int y0, int x0, int k, const double* wy, const double* wx) {
          double ry = 0;
          for (int jy = 0; jy < S; jy++) {
            double rx = 0;
            for (int jx = 0; jx < S; jx++)
              rx += c.UF(iz0, y0 + jy, x0 + jx, k) * wx[jx];
            ry += rx * wy[jy];
          }
          return ry * qmdt;
        };
        eb[0] = interp(iy0, hx0, 0, wiy, whx);
        eb[1] = interp(hy0, ix0, 1, why, wix);
        eb[2] = interp(iy0, ix0, 2, wiy, wix);
        eb[3] = interp(hy0, ix0, 3, why, wix);
        eb[4] = interp(iy0, hx0, 4, wiy, whx);
        eb[5] = interp(hy0, hx0, 5, why, whx);
      }
      // BaseVelocity::push_impl, velocity.hpp:81-99
      if (pusher == 0)
        push_boris(xu[3], xu[4], xu[5], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
      else if (pusher == 1)
        push_vay(xu[3], xu[4], xu[5], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
      else
        push_higuera_cary(xu[3], xu[4], xu[5], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
    }
  }
}

// pic/engine/position.hpp:55-68,117-130 then Position::set_boundary (pic/pic_engine.hpp:292-303)
void shock_boundary_particle(Sim& s, Chunk& c);

void push_position_chunk(Sim& s, Chunk& c, int Ns, double cc, int order, double delt)
{
  double rc = 1 / cc;
  for (int is = 0; is < Ns; is++) {
    Particles& P = c.up[is];
    for (int ip = 0; ip < P.Np; ip++) {
      double* xu = P.XU(ip);
      double* xv = P.XV(ip);
      for (int i = 0; i < NC; i++)
        xv[i] = xu[i];
      double gm = lorentz_factor(xu[3], xu[4], xu[5], rc);
      double dt = delt / gm;
      xu[0] += xu[3] * dt;
      xu[1] += xu[4] * dt;
      xu[2] += xu[5] * dt;
    }
  }
  // set_boundary_particle: a no-op for periodic runs (pic/pic_chunk.cpp:483-485),
  // the wall and upstream boundary for the shock
  if (s.p.problem == 1)
    shock_boundary_particle(s, c);
  for (int is = 0; is < Ns; is++)
    count(c, c.up[is], 0, c.up[is].Np - 1, true, order);
}

// pic/engine/current.hpp:326-395 (local3d) / :275-324 (local2d); scalar path :59-98
template <int Dim, int Order>
void local_current(const Geo& g, const double* xv, const double* xu, double qs, double delt,
                   double* cur, int origin[3])
{
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int N      = Order + 3;
  double        rdx = 1 / g.dx, rdy = 1 / g.dy, rdz = 1 / g.dz;
  double        ximin = g.xmin + 0.5 * g.dx * is_odd;
  double        yimin = g.ymin + 0.5 * g.dy * is_odd;
  double        zimin = g.zmin + 0.5 * g.dz * is_odd;
  double        xgrid = g.xmin + 0.5 * g.dx;
  double        ygrid = g.ymin + 0.5 * g.dy;
  double        zgrid = g.zmin + 0.5 * g.dz;
  double        dxdt  = g.dx / delt;
  double        dydt  = g.dy / delt;
  double        dzdt  = g.dz / delt;

  if constexpr (Dim == 3) {
    double ss[2][3][N] = {};
    int    ix0         = digitize(xv[0], ximin, rdx);
    int    iy0         = digitize(xv[1], yimin, rdy);
    int    iz0         = digitize(xv[2], zimin, rdz);
    shape_mc<Order>(xv[0], xgrid + static_cast<double>(ix0) * g.dx, rdx, &ss[0][0][1]);
    shape_mc<Order>(xv[1], ygrid + static_cast<double>(iy0) * g.dy, rdy, &ss[0][1][1]);
    shape_mc<Order>(xv[2], zgrid + static_cast<double>(iz0) * g.dz, rdz, &ss[0][2][1]);
    int ix1 = digitize(xu[0], ximin, rdx);
    int iy1 = digitize(xu[1], yimin, rdy);
    int iz1 = digitize(xu[2], zimin, rdz);
    shape_mc<Order>(xu[0], xgrid + static_cast<double>(ix1) * g.dx, rdx, &ss[1][0][1]);
    shape_mc<Order>(xu[1], ygrid + static_cast<double>(iy1) * g.dy, rdy, &ss[1][1][1]);
    shape_mc<Order>(xu[2], zgrid + static_cast<double>(iz1) * g.dz, rdz, &ss[1][2][1]);
    int shift[3] = {ix1 - ix0, iy1 - iy0, iz1 - iz0};
    shift_weights<3, Order>(shift, ss[1]);
    deposit3d<Order>(dxdt, dydt, dzdt, qs, ss, cur);
    origin[0] = iz0 + g.lbz - (Order / 2) - 1;
    origin[1] = iy0 + g.lby - (Order / 2) - 1;
    origin[2] = ix0 + g.lbx - (Order / 2) - 1;
  } else if constexpr (Dim == 1) {
    // pic/engine/current.hpp:207-258 (local1d)
    double ss[2][1][N] = {};
    double vy          = (xu[1] - xv[1]) / delt;
    double vz          = (xu[2] - xv[2]) / delt;
    int    ix0         = digitize(xv[0], ximin, rdx);
    shape_mc<Order>(xv[0], xgrid + static_cast<double>(ix0) * g.dx, rdx, &ss[0][0][1]);
    int ix1 = digitize(xu[0], ximin, rdx);
    shape_mc<Order>(xu[0], xgrid + static_cast<double>(ix1) * g.dx, rdx, &ss[1][0][1]);
    int shift[1] = {ix1 - ix0};
    shift_weights<1, Order>(shift, ss[1]);
    deposit1d<Order>(dxdt, vy, vz, qs, ss, cur);
    origin[0] = g.lbz;
    origin[1] = g.lby;
    origin[2] = ix0 + g.lbx - (Order / 2) - 1;
    (void)yimin;
    (void)zimin;
    (void)ygrid;
    (void)zgrid;
    (void)dydt;
    (void)dzdt;
    (void)rdy;
    (void)rdz;
  } else {
    double ss[2][2][N] = {};
    double vz          = (xu[2] - xv[2]) / delt;
    int    ix0         = digitize(xv[0], ximin, rdx);
    int    iy0         = digitize(xv[1], yimin, rdy);
    shape_mc<Order>(xv[0], xgrid + static_cast<double>(ix0) * g.dx, rdx, &ss[0][0][1]);
    shape_mc<Order>(xv[1], ygrid + static_cast<double>(iy0) * g.dy, rdy, &ss[0][1][1]);
    int ix1 = digitize(xu[0], ximin, rdx);
    int iy1 = digitize(xu[1], yimin, rdy);
    shape_mc<Order>(xu[0], xgrid + static_cast<double>(ix1) * g.dx, rdx, &ss[1][0][1]);
    shape_mc<Order>(xu[1], ygrid + static_cast<double>(iy1) * g.dy, rdy, &ss[1][1][1]);
    int shift[2] = {ix1 - ix0, iy1 - iy0};
    shift_weights<2, Order>(shift, ss[1]);
    deposit2d<Order>(dxdt, dydt, vz, qs, ss, cur);
    origin[0] = g.lbz;
    origin[1] = iy0 + g.lby - (Order / 2) - 1;
    origin[2] = ix0 + g.lbx - (Order / 2) - 1;
  }
}

// pic/engine/current.hpp:59-98 (call_scalar_impl) + append_current{2,3}d (primitives.hpp:678-834)
template <int Dim, int Order>
void deposit_current_chunk(Chunk& c, int Ns, double delt)
{
  constexpr int N  = Order + 3;
  constexpr int NT = (Dim == 3 ? N * N * N : (Dim == 2 ? N * N : N)) * 4;
  Geo           g(c);
  std::fill(c.uj.begin(), c.uj.end(), 0.0);
  for (int is = 0; is < Ns; is++) {
    Particles& P = c.up[is];
    for (int ip = 0; ip < P.Np; ip++) {
      double cur[NT] = {0};
      int    o[3];
      local_current<Dim, Order>(g, P.XV(ip), P.XU(ip), P.q, delt, cur, o);
      if constexpr (Dim == 3) {
        for (int jz = 0; jz < N; jz++)
          for (int jy = 0; jy < N; jy++)
            for (int jx = 0; jx < N; jx++)
              for (int k = 0; k < 4; k++)
                c.UJ(o[0] + jz, o[1] + jy, o[2] + jx, k) += cur[((jz * N + jy) * N + jx) * 4 + k];
      } else if constexpr (Dim == 2) {
        for (int jy = 0; jy < N; jy++)
          for (int jx = 0; jx < N; jx++)
            for (int k = 0; k < 4; k++)
              c.UJ(o[0], o[1] + jy, o[2] + jx, k) += cur[(jy * N + jx) * 4 + k];
      } else {
        for (int jx = 0; jx < N; jx++)
          for (int k = 0; k < 4; k++)
            c.UJ(o[0], o[1], o[2] + jx, k) += cur[jx * 4 + k];
      }
    }
  }
}

// pic/engine/moment.hpp:74-98 (scalar path), local2d/local3d :271-394
template <int Dim, int Order>
void deposit_moment_chunk(Chunk& c, int Ns, double cc)
{
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int S      = Order + 1;
  Geo           g(c);
  std::fill(c.um.begin(), c.um.end(), 0.0);
  for (int is = 0; is < Ns; is++) {
    Particles& P  = c.up[is];
    double     ms = P.m;
    for (int ip = 0; ip < P.Np; ip++) {
      double* xu = P.XU(ip);
      double  wx[S] = {0}, wy[S] = {0}, wz[S] = {0};
      double  rc    = 1 / cc;
      double  rdx = 1 / g.dx, rdy = 1 / g.dy, rdz = 1 / g.dz;
      int     ix = digitize(xu[0], g.xmin + 0.5 * g.dx * is_odd, rdx);
      shape_mc<Order>(xu[0], (g.xmin + 0.5 * g.dx) + static_cast<double>(ix) * g.dx, rdx, wx);
      int iy = g.lby, Sy = 1;
      if constexpr (Dim >= 2) {
        iy = digitize(xu[1], g.ymin + 0.5 * g.dy * is_odd, rdy);
        shape_mc<Order>(xu[1], (g.ymin + 0.5 * g.dy) + static_cast<double>(iy) * g.dy, rdy, wy);
        iy += g.lby - (Order / 2);
        Sy = S;
      } else {
        wy[0] = 1.0;
      }
      int iz = g.lbz, Sz = 1;
      if constexpr (Dim == 3) {
        iz = digitize(xu[2], g.zmin + 0.5 * g.dz * is_odd, rdz);
        shape_mc<Order>(xu[2], (g.zmin + 0.5 * g.dz) + static_cast<double>(iz) * g.dz, rdz, wz);
        iz += g.lbz - (Order / 2);
        Sz = S;
      } else {
        wz[0] = 1.0;
      }
      ix += g.lbx - (Order / 2);
      for (int jz = 0; jz < Sz; jz++)
        for (int jy = 0; jy < Sy; jy++)
          for (int jx = 0; jx < S; jx++) {
            // moment.hpp local1d/local2d/local3d: the product of the present axes' weights
            double ww = (Dim == 3) ? ms * wx[jx] * wy[jy] * wz[jz]
                                   : (Dim == 2 ? ms * wx[jx] * wy[jy] : ms * wx[jx]);
            double gm = lorentz_factor(xu[3], xu[4], xu[5], rc);
            double m[14];
            m[0]  = ww;
            m[1]  = ww * xu[3] / gm;
            m[2]  = ww * xu[4] / gm;
            m[3]  = ww * xu[5] / gm;
            m[8]  = ww * xu[3];
            m[9]  = ww * xu[4];
            m[10] = ww * xu[5];
            m[4]  = ww * gm * cc;
            m[5]  = ww * xu[3] * xu[3] / gm;
            m[6]  = ww * xu[4] * xu[4] / gm;
            m[7]  = ww * xu[5] * xu[5] / gm;
            m[11] = ww * xu[3] * xu[4] / gm;
            m[12] = ww * xu[4] * xu[5] / gm;
            m[13] = ww * xu[5] * xu[3] / gm;
            for (int k = 0; k < 14; k++)
              c.UM(iz + jz, iy + jy, ix + jx, is, k, Ns) += m[k];
          }
    }
  }
}

// --------------------------------------------------------------------------
// halo exchanges (nix/chunk.hpp:392-543, nix/xtensor_halo3d.hpp) as in-memory copies
// --------------------------------------------------------------------------
inline int dir_index(int iz, int iy, int ix) { return 9 * iz + 3 * iy + ix; }

// copy a [z][y][x][ncomp] box between two ghost-inclusive arrays
template <typename Op>
void box_op(const Chunk& src, const std::vector<double>& sa, const int slo[3], Chunk& dst,
            std::vector<double>& da, const int dlo[3], const int ext[3], int ncomp, Op op)
{
  for (int z = 0; z < ext[0]; z++)
    for (int y = 0; y < ext[1]; y++)
      for (int x = 0; x < ext[2]; x++) {
        size_t si = src.cell(slo[0] + z, slo[1] + y, slo[2] + x) * ncomp;
        size_t di = dst.cell(dlo[0] + z, dlo[1] + y, dlo[2] + x) * ncomp;
        for (int k = 0; k < ncomp; k++)
          op(da[di + k], sa[si + k]);
      }
}

// Emf (XtensorHaloField3D: send interior edge -> neighbour ghost, copy),
// Cur / Mom (XtensorHaloCurrent3D / Moment3D: send ghost -> neighbour interior edge, add).
// Receiver R unpacks slot (iz,iy,ix) from neighbour S = nb_R(slot), whose message was
// packed for direction -slot (MPI matching of nix/chunk.hpp:465-524).
void set_boundary_emf(const Sim& s, Chunk& c);
void set_boundary_mom(const Sim& s, Chunk& c);

void exchange_field(Sim& s, int mode)
{
  const bool add = (mode != 0);
  // the split-independent order resets the walls before the neighbours copy them
  if (mode == 0 && s.p.problem == 1 && s.p.emf_order == 1)
    for_chunks(s, [&](Chunk& c) { set_boundary_emf(s, c); });
  for_chunks(s, [&](Chunk& R) {
    for (int iz = R.ilo(0); iz <= R.ihi(0); iz++)
      for (int iy = R.ilo(1); iy <= R.ihi(1); iy++)
        for (int ix = R.ilo(2); ix <= R.ihi(2); ix++) {
          if (iz == 1 && iy == 1 && ix == 1)
            continue;
          if (R.nb_id[dir_index(iz, iy, ix)] < 0)
            continue; // MPI_PROC_NULL: nothing is sent or received
          const Chunk& S   = s.chunks[R.nb_id[dir_index(iz, iy, ix)]];
          const int    o[3] = {2 - iz, 2 - iy, 2 - ix}; // sender's slot
          const int    r[3] = {iz, iy, ix};
          int          slo[3], dlo[3], ext[3];
          for (int d = 0; d < 3; d++) {
            int lo, hi, lo2, hi2;
            if (!add) {
              S.bounds(d, o[d], true, lo, hi);   // sender interior edge
              R.bounds(d, r[d], false, lo2, hi2); // receiver ghost
            } else {
              S.bounds(d, o[d], false, lo, hi);  // sender ghost
              R.bounds(d, r[d], true, lo2, hi2); // receiver interior edge
            }
            slo[d] = lo;
            dlo[d] = lo2;
            ext[d] = hi - lo + 1;
          }
          if (mode == 0)
            box_op(S, S.uf, slo, R, R.uf, dlo, ext, 6, [](double& a, double b) { a = b; });
          else if (mode == 1)
            box_op(S, S.uj, slo, R, R.uj, dlo, ext, 4, [](double& a, double b) { a += b; });
          else
            box_op(S, S.um, slo, R, R.um, dlo, ext, 14 * s.Ns,
                   [](double& a, double b) { a += b; });
        }
  });
  // physical boundary after the unpack (pic/pic_chunk.cpp:339-340 -> set_boundary_field)
  if (s.p.problem == 1) {
    if (mode == 0)
      for_chunks(s, [&](Chunk& c) { set_boundary_emf(s, c); });
    else if (mode == 2)
      for_chunks(s, [&](Chunk& c) { set_boundary_mom(s, c); });
  }
}

// particle migration: XtensorHaloParticle3D::{pre_pack, pack, pre_unpack, unpack,
// post_unpack} (nix/xtensor_halo3d.hpp:192-499)
void shock_inject_particle(Sim& s, Chunk& c);

void exchange_particle(Sim& s)
{
  const int nc = (int)s.chunks.size();
  // set_boundary_pack(BoundaryParticle) injects first (pic/pic_chunk.cpp:305-309)
  if (s.p.problem == 1)
    for_chunks(s, [&](Chunk& c) { shock_inject_particle(s, c); });
  // packed leavers: [chunk][species][27 directions] -> flat records
  std::vector<std::vector<std::vector<double>>> sendbuf(nc);
  for_chunks(s, [&](Chunk& S) {
    auto& buf = sendbuf[S.id];
    buf.assign(s.Ns * 27, std::vector<double>());
    for (int is = 0; is < s.Ns; is++) {
      Particles& P = S.up[is];
      for (int ip = 0; ip < P.Np; ip++) {
        const double* x  = P.XU(ip);
        int           iz = (x[2] >= S.rng[0][1]) - (x[2] < S.rng[0][0]) + 1;
        int           iy = (x[1] >= S.rng[1][1]) - (x[1] < S.rng[1][0]) + 1;
        int           ix = (x[0] >= S.rng[2][1]) - (x[0] < S.rng[2][0]) + 1;
        if (ix == 1 && iy == 1 && iz == 1)
          continue;
        auto& b = buf[is * 27 + dir_index(iz, iy, ix)];
        b.insert(b.end(), x, x + NC);
      }
    }
  });
  for_chunks(s, [&](Chunk& R) {
    for (int is = 0; is < s.Ns; is++) {
      Particles& P       = R.up[is];
      int        np_prev = P.Np;
      int        nrecv   = 0;
      for (int iz = R.ilo(0); iz <= R.ihi(0); iz++)
        for (int iy = R.ilo(1); iy <= R.ihi(1); iy++)
          for (int ix = R.ilo(2); ix <= R.ihi(2); ix++) {
            if (iz == 1 && iy == 1 && ix == 1)
              continue;
            int   sid = R.nb_id[dir_index(iz, iy, ix)];
            if (sid < 0)
              continue;
            auto& b   = sendbuf[sid][is * 27 + dir_index(2 - iz, 2 - iy, 2 - ix)];
            nrecv += (int)(b.size() / NC);
          }
      P.reserve(np_prev + nrecv);
      int pos = np_prev;
      for (int iz = R.ilo(0); iz <= R.ihi(0); iz++)
        for (int iy = R.ilo(1); iy <= R.ihi(1); iy++)
          for (int ix = R.ilo(2); ix <= R.ihi(2); ix++) {
            if (iz == 1 && iy == 1 && ix == 1)
              continue;
            int   sid = R.nb_id[dir_index(iz, iy, ix)];
            if (sid < 0)
              continue;
            auto& b   = sendbuf[sid][is * 27 + dir_index(2 - iz, 2 - iy, 2 - ix)];
            if (!b.empty())
              std::memcpy(P.XU(pos), b.data(), b.size() * sizeof(double));
            pos += (int)(b.size() / NC);
          }
      set_boundary_periodic(R, P, np_prev, pos - 1);
      count(R, P, np_prev, pos - 1, false, s.order);
      P.Np = pos;
    }
    for (int is = 0; is < s.Ns; is++)
      sort(R.up[is]);
  });
}

// --------------------------------------------------------------------------
// dispatch helpers
// --------------------------------------------------------------------------
template <template <int, int> class F, typename... Args>
void dispatch(int dim, int order, Args&&... args)
{
  if (dim == 3) {
    switch (order) {
    case 1: F<3, 1>::run(args...); break;
    case 2: F<3, 2>::run(args...); break;
    case 3: F<3, 3>::run(args...); break;
    default: F<3, 4>::run(args...); break;
    }
  } else if (dim == 2) {
    switch (order) {
    case 1: F<2, 1>::run(args...); break;
    case 2: F<2, 2>::run(args...); break;
    case 3: F<2, 3>::run(args...); break;
    default: F<2, 4>::run(args...); break;
    }
  } else {
    switch (order) {
    case 1: F<1, 1>::run(args...); break;
    case 2: F<1, 2>::run(args...); break;
    case 3: F<1, 3>::run(args...); break;
    default: F<1, 4>::run(args...); break;
    }
  }
}

template <int D, int O>
struct VelocityOp {
  static void run(Chunk& c, int Ns, double cc, int pusher, int interp, double delt)
  {
    push_velocity_chunk<D, O>(c, Ns, cc, pusher, interp, delt);
  }
};
template <int D, int O>
struct CurrentOp {
  static void run(Chunk& c, int Ns, double delt) { deposit_current_chunk<D, O>(c, Ns, delt); }
};
template <int D, int O>
struct MomentOp {
  static void run(Chunk& c, int Ns, double cc) { deposit_moment_chunk<D, O>(c, Ns, cc); }
};

// --------------------------------------------------------------------------
// example/shock: reflecting wall at x = 0 and upstream injection at x = Lx on the
// chunks whose x neighbour is MPI_PROC_NULL (set_periodicity(1, 1, 0),
// example/shock/main.cpp:550-558)
// --------------------------------------------------------------------------
// example/shock/main.cpp:50-54
double initial_flow_profile(const Chunk& c, double x, double L)
{
  double xmin = c.glim[2][0] + c.Nb * c.del[2];
  return -(x - xmin < L ? std::sin(0.5 * M_PI * (x - xmin) / L) : 1.0);
}

// density extension n(x) (1 everywhere for density = 1, the reference's loader)
double shock_density(const pno_params& p, double x, double delh)
{
  if (p.density == 1.0)
    return 1.0;
  double xc = x / delh;
  if (xc < p.ramp_x)
    return 1.0;
  if (xc < p.ramp_x + p.ramp_w)
    return 1.0 + (p.density - 1.0) * (xc - p.ramp_x) / p.ramp_w;
  return p.density;
}

int shock_npc(const pno_params& p, double x, double delh)
{
  return (int)std::floor(p.nppc * shock_density(p, x, delh) + 0.5);
}

// MainChunk::reset_random_number (example/shock/main.cpp:19-24)
void reset_random_number(Chunk& c)
{
  c.random_seed += 1;
  c.mt.seed((uint64_t)(int64_t)c.random_seed);
}

// MainChunk::generate_injection_particle (example/shock/main.cpp:26-48), c = 1
void generate_injection_particle(Chunk& c, MaxwellJuttner& mj, double delt, double& x,
                                 double& ux, double& uy, double& uz)
{
  const double cc = 1.0;
  const double rc = 1.0 / cc;
  mj(c.mt, ux, uy, uz);
  double xx = c.lim[2][1] + cc * delt * uniform01(c.mt);
  double dx = ux * delt / lorentz_factor(ux, uy, uz, rc);
  while (xx + dx >= c.lim[2][1]) {
    mj(c.mt, ux, uy, uz);
    xx = c.lim[2][1] + cc * delt * uniform01(c.mt);
    dx = ux * delt / lorentz_factor(ux, uy, uz, rc);
  }
  x = xx + dx;
}

// MainChunk::set_boundary_emf (example/shock/main.cpp:232-340): conducting wall
// over the first Nb cells and the lower ghosts; fixed upstream fields over the
// last Nb cells and the upper ghosts.
void set_boundary_emf(const Sim& s, Chunk& c)
{
  const int Nb  = c.Nb;
  const int Lbx = c.Lb[2], Ubx = c.Ub[2], Lby = c.Lb[1], Uby = c.Ub[1], Lbz = c.Lb[0],
            Ubz = c.Ub[0];
  if (c.wall_lo) {
    for (int ix = 0; ix < 2 * Nb; ix++) {
      int ix1 = Lbx - ix + Nb - 1;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb; iy++) {
          c.UF(iz, iy, ix1, 1) = 0;
          c.UF(iz, iy, ix1, 2) = 0;
        }
    }
    for (int ix = 0; ix < 2 * Nb; ix++) {
      int ix1 = Lbx - ix + Nb - 1;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb - 1; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb - 1; iy++)
          c.UF(iz, iy, ix1, 0) = 0;
    }
    for (int ix = 0; ix < 2 * Nb; ix++) {
      int ix1 = Lbx - ix + Nb - 1;
      int ix2 = Lbx + Nb;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb; iy++) {
          c.UF(iz, iy, ix1, 4) = c.UF(iz, iy, ix2, 4);
          c.UF(iz, iy, ix1, 5) = c.UF(iz, iy, ix2, 5);
        }
    }
    for (int ix = 0; ix < 2 * Nb; ix++) {
      int ix1 = Lbx - ix + Nb - 1;
      int ix2 = Lbx + Nb;
      for (int iz = Lbz - Nb + 1; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb + 1; iy <= Uby + Nb; iy++)
          c.UF(iz, iy, ix1, 3) = c.UF(iz, iy, ix2, 3);
    }
  }
  if (c.wall_hi) {
    const double* E0 = s.sh.E0;
    const double* B0 = s.sh.B0;
    for (int ix = 0; ix < 2 * Nb; ix++) {
      int ix1 = Ubx + ix - Nb + 1;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb; iy++) {
          c.UF(iz, iy, ix1, 1) = E0[1];
          c.UF(iz, iy, ix1, 2) = E0[2];
        }
    }
    for (int ix = 0; ix < 2 * Nb - 1; ix++) {
      int ix1 = Ubx + ix - Nb + 2;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb - 1; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb - 1; iy++)
          c.UF(iz, iy, ix1, 0) = E0[0];
    }
    for (int ix = 0; ix < 2 * Nb - 1; ix++) {
      int ix1 = Ubx + ix - Nb + 2;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb; iy++) {
          c.UF(iz, iy, ix1, 4) = B0[1];
          c.UF(iz, iy, ix1, 5) = B0[2];
        }
    }
    for (int ix = 0; ix < 2 * Nb; ix++) {
      int ix1 = Ubx + ix - Nb + 1;
      for (int iz = Lbz - Nb + 1; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb + 1; iy <= Uby + Nb; iy++)
          c.UF(iz, iy, ix1, 3) = B0[0];
    }
  }
}

// MainChunk::set_boundary_mom (example/shock/main.cpp:342-383).  It indexes the
// 5-D moment array um(z, y, x, species, k) with four indices; xtensor matches
// too few indices to the trailing axes (nix/thirdparty/xtensor/xstrides.hpp:
// 471-476), so um(iz, iy, ix, ik) is the element at flat offset
// iz*s_y + iy*s_x + ix*s_species + ik of the z = 0 plane.  Restated literally.
void set_boundary_mom(const Sim& s, Chunk& c)
{
  const int Nb  = c.Nb;
  const int Lbx = c.Lb[2], Ubx = c.Ub[2], Lby = c.Lb[1], Uby = c.Ub[1], Lbz = c.Lb[0],
            Ubz = c.Ub[0];
  const size_t sk = 1, ss = 14, sx = 14 * (size_t)s.Ns, sy = sx * c.n[2];
  auto at = [&](int a, int b, int d, int k) -> double& {
    return c.um[(size_t)a * sy + (size_t)b * sx + (size_t)d * ss + (size_t)k * sk];
  };
  if (c.wall_lo)
    for (int ix = 0; ix < Nb; ix++) {
      int ix1 = Lbx - ix + Nb - 1;
      int ix2 = Lbx + Nb;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb; iy++)
          for (int ik = 0; ik < 14; ik++)
            at(iz, iy, ix1, ik) = at(iz, iy, ix2, ik);
    }
  if (c.wall_hi)
    for (int ix = 0; ix < Nb; ix++) {
      int ix1 = Ubx + ix - Nb + 1;
      int ix2 = Ubx - Nb;
      for (int iz = Lbz - Nb; iz <= Ubz + Nb; iz++)
        for (int iy = Lby - Nb; iy <= Uby + Nb; iy++)
          for (int ik = 0; ik < 14; ik++)
            at(iz, iy, ix1, ik) = at(iz, iy, ix2, ik);
    }
}

// MainChunk::set_boundary_particle (example/shock/main.cpp:398-443), called by
// Position before the count (pic/pic_engine.hpp:292-303)
void shock_boundary_particle(Sim& s, Chunk& c)
{
  if (c.wall_lo)
    for (int is = 0; is < s.Ns; is++) {
      Particles& P = c.up[is];
      for (int ip = 0; ip < P.Np; ip++) {
        double* xu = P.XU(ip);
        if (xu[0] < c.glim[2][0]) {
          xu[0] = -xu[0] + 2 * c.glim[2][0];
          xu[3] = -xu[3];
          xu[4] = -xu[4];
          xu[5] = -xu[5];
        }
      }
    }
  if (c.wall_hi) {
    reset_random_number(c);
    MaxwellJuttner mj[2] = {MaxwellJuttner(s.sh.vte * s.sh.vte, -s.sh.u0),
                            MaxwellJuttner(s.sh.vti * s.sh.vti, -s.sh.u0)};
    for (int is = 0; is < s.Ns; is++) {
      int        efflux = 0;
      Particles& P      = c.up[is];
      for (int ip = 0; ip < P.Np; ip++) {
        double* xu = P.XU(ip);
        if (xu[0] >= c.glim[2][1]) {
          double x, ux, uy, uz;
          generate_injection_particle(c, mj[is], s.sh.delt, x, ux, uy, uz);
          xu[0] = x;
          xu[3] = ux;
          xu[4] = uy;
          xu[5] = uz;
          efflux += 1;
        }
      }
      c.efflux[is] = efflux;
    }
  }
}

// MainChunk::inject_particle (example/shock/main.cpp:445-535), called from
// set_boundary_pack(BoundaryParticle) (pic/pic_chunk.cpp:305-309)
void shock_inject_particle(Sim& s, Chunk& c)
{
  if (!c.wall_hi)
    return;
  const double rc   = 1.0 / 1.0;
  const double u0   = s.sh.u0;
  const double npu  = s.p.density == 1.0 ? (double)s.sh.nppc
                                          : (double)shock_npc(s.p, c.glim[2][1], c.del[2]);
  const double flux = npu * std::abs(u0) / std::sqrt(1 + u0 * u0 * rc * rc) * s.sh.delt / c.del[2];
  const int    Lbx = c.Lb[2], Lby = c.Lb[1], Uby = c.Ub[1], Lbz = c.Lb[0], Ubz = c.Ub[0];
  (void)Lbx;
  std::vector<int>     np_inj((size_t)c.n[0] * c.n[1], 0);
  std::vector<int64_t> id64(s.Ns, c.lastid);
  int                  np_inj_total = 0;
  reset_random_number(c);
  if (!(flux < 12))
    return; // libstdc++'s rejection branch (mean >= 12) is not restated
  const double   lm_thr = std::exp(-flux);
  MaxwellJuttner mj[2]  = {MaxwellJuttner(s.sh.vte * s.sh.vte, -u0),
                          MaxwellJuttner(s.sh.vti * s.sh.vti, -u0)};
  for (int iz = Lbz; iz <= Ubz; iz++)
    for (int iy = Lby; iy <= Uby; iy++) {
      np_inj[(size_t)iz * c.n[1] + iy] = poisson_small(c.mt, lm_thr);
      np_inj_total += np_inj[(size_t)iz * c.n[1] + iy];
    }
  for (int is = 0; is < s.Ns; is++)
    c.up[is].reserve(c.up[is].Np + np_inj_total + 1);
  for (int iz = Lbz; iz <= Ubz; iz++)
    for (int iy = Lby; iy <= Uby; iy++)
      for (int is = 0; is < s.Ns; is++) {
        Particles& P       = c.up[is];
        int        np_prev = P.Np;
        int        np_next = P.Np + np_inj[(size_t)iz * c.n[1] + iy];
        for (int ip = np_prev; ip < np_next; ip++) {
          double x, ux, uy, uz;
          generate_injection_particle(c, mj[is], s.sh.delt, x, ux, uy, uz);
          double  y  = c.lim[1][0] + (iy - Lby + uniform01(c.mt)) * c.del[1];
          double  z  = c.lim[0][0] + (iz - Lbz + uniform01(c.mt)) * c.del[0];
          double* xu = P.XU(ip);
          xu[0]      = x;
          xu[1]      = y;
          xu[2]      = z;
          xu[3]      = ux;
          xu[4]      = uy;
          xu[5]      = uz;
          std::memcpy(&xu[6], &id64[is], sizeof(int64_t));
          id64[is]++;
        }
        if (np_next > np_prev)
          count(c, P, np_prev, np_next - 1, false, s.order);
        P.Np = np_next;
      }
  c.lastid += np_inj_total;
  for (int is = 0; is < s.Ns; is++)
    c.influx[is] = np_inj_total;
}

// --------------------------------------------------------------------------
// set-up: PicApplication/Application chunk set-up + example loaders
// --------------------------------------------------------------------------
// example/shock/main.cpp:56-92: cc = 1 and Ns = 2 are fixed by the example
void setup_shock_params(Sim& s)
{
  pno_params& p = s.p;
  p.cc          = 1.0;
  p.Ns          = 2;
  ShockParams& h = s.sh;
  const double cc = 1.0;
  const double u0 = p.u0;
  double gamma = std::sqrt(1.0 + (u0 * u0) / (cc * cc));
  double me    = 1.0 / p.nppc;
  double qe    = -p.wp / p.nppc * std::sqrt(gamma);
  double mi    = me * p.mime;
  double qi    = -qe;
  double b0    = cc * std::sqrt(p.sigma) / std::abs(qe / me);
  double vae   = cc * std::sqrt(p.sigma);
  double vai   = cc * std::sqrt(p.sigma / p.mime);
  h.nppc       = p.nppc;
  h.delt       = p.delt;
  h.u0         = u0;
  h.taper      = p.taper;
  h.vte        = vae * std::sqrt(0.5 * p.betae);
  h.vti        = vai * std::sqrt(0.5 * p.betai);
  h.B0[0]      = b0 * std::cos(p.theta / 180 * M_PI);
  h.B0[1]      = b0 * std::sin(p.theta / 180 * M_PI) * std::cos(p.phi / 180 * M_PI);
  h.B0[2]      = b0 * std::sin(p.theta / 180 * M_PI) * std::sin(p.phi / 180 * M_PI);
  h.E0[0]      = 0;
  h.E0[1]      = +(-u0) * h.B0[2] / (gamma * cc);
  h.E0[2]      = -(-u0) * h.B0[1] / (gamma * cc);
  h.me         = me;
  h.qe         = qe;
  h.mi         = mi;
  h.qi         = qi;
}

// example/shock/main.cpp:94-137: fields over the ghost-inclusive block (the
// position uses the array index, xi = (ix + 0.5) dx + xmin)
void init_shock_fields(Sim& s, Chunk& c)
{
  const ShockParams& h  = s.sh;
  const double       cc = 1.0;
  const double       xmin = c.lim[2][0];
  for (int iz = 0; iz < c.n[0]; iz++)
    for (int iy = 0; iy < c.n[1]; iy++)
      for (int ix = 0; ix < c.n[2]; ix++) {
        double xi = (ix + 0.5) * c.del[2] + xmin;
        double Ux = h.u0 * initial_flow_profile(c, xi, h.taper);
        double gm = std::sqrt(1 + Ux * Ux / (cc * cc));
        c.UF(iz, iy, ix, 0) = h.E0[0];
        c.UF(iz, iy, ix, 1) = +Ux * h.B0[2] / (gm * cc);
        c.UF(iz, iy, ix, 2) = -Ux * h.B0[1] / (gm * cc);
        c.UF(iz, iy, ix, 3) = h.B0[0];
        c.UF(iz, iy, ix, 4) = h.B0[1];
        c.UF(iz, iy, ix, 5) = h.B0[2];
      }
}

// example/shock/main.cpp:139-210: nppc particles per cell with a drifting
// Maxwell-Juttner velocity, electrons and ions at the same positions, one
// mt19937_64 per chunk
void load_shock_particles(Sim& s, Chunk& c, const int gdims[3])
{
  const pno_params&  p = s.p;
  const ShockParams& h = s.sh;
  c.mt.seed((uint64_t)(int64_t)c.random_seed);
  MaxwellJuttner mj[2] = {MaxwellJuttner(h.vte * h.vte, 0.0), MaxwellJuttner(h.vti * h.vti, 0.0)};
  const int npc_max = p.density == 1.0 ? p.nppc
                                       : (int)std::floor(p.nppc * std::max(1.0, p.density) + 0.5);
  int       mp    = npc_max * c.dims[0] * c.dims[1] * c.dims[2];
  int       cz    = gdims[0] / c.dims[0];
  int       cy    = gdims[1] / c.dims[1];
  int64_t   maxid = std::numeric_limits<int64_t>::max() / (cy * cz);
  int64_t   ciz   = c.offset[0] / c.dims[0];
  int64_t   ciy   = c.offset[1] / c.dims[1];
  int64_t   cix   = c.offset[2] / c.dims[2];
  c.lastid        = (ciy + ciz * cy) * maxid + cix * mp;
  for (int is = 0; is < 2; is++) {
    Particles& P = c.up[is];
    P.m          = is == 0 ? h.me : h.mi;
    P.q          = is == 0 ? h.qe : h.qi;
    P.reserve((int)(mp * (1 + p.buffer_ratio)) + 1);
  }
  int ip = 0;
  for (int iz = c.Lb[0]; iz <= c.Ub[0]; iz++)
    for (int iy = c.Lb[1]; iy <= c.Ub[1]; iy++)
      for (int ix = c.Lb[2]; ix <= c.Ub[2]; ix++) {
        int npc = p.nppc;
        if (p.density != 1.0)
          npc = shock_npc(p, c.lim[2][0] + (ix - c.Lb[2] + 0.5) * c.del[2], c.del[2]);
        for (int jp = 0; jp < npc; ip++, jp++) {
          double x = c.lim[2][0] + (ix - c.Lb[2] + uniform01(c.mt)) * c.del[2];
          double y = c.lim[1][0] + (iy - c.Lb[1] + uniform01(c.mt)) * c.del[1];
          double z = c.lim[0][0] + (iz - c.Lb[0] + uniform01(c.mt)) * c.del[0];
          double U = h.u0 * initial_flow_profile(c, x, h.taper);
          for (int is = 0; is < 2; is++) {
            mj[is].set_drift(U);
            double ux, uy, uz;
            mj[is](c.mt, ux, uy, uz);
            double* xu = c.up[is].XU(ip);
            xu[0]      = x;
            xu[1]      = y;
            xu[2]      = z;
            xu[3]      = ux;
            xu[4]      = uy;
            xu[5]      = uz;
            int64_t id = c.lastid + ip;
            std::memcpy(&xu[6], &id, sizeof(id));
          }
        }
      }
  for (int is = 0; is < 2; is++)
    c.up[is].Np = ip;
  c.lastid += mp;
}

void setup(Sim& s)
{
  if (s.p.problem == 1)
    setup_shock_params(s);
  const pno_params& p = s.p;
  s.Ns                = p.Ns;
  s.order             = p.order;
  const bool xper     = p.problem != 1; // set_periodicity(1, 1, 0) for the shock
  const int N[3]      = {p.Nz, p.Ny, p.Nx};
  const int C[3]      = {p.Cz, p.Cy, p.Cx};
  bool      has_dim[3];
  for (int d = 0; d < 3; d++)
    has_dim[d] = !(N[d] == 1 && C[d] == 1); // nix/application.cpp:263-267
  s.dim = has_dim[0] ? 3 : (has_dim[1] ? 2 : 1);

  s.id_of_coord = chunk_map(C[0], C[1], C[2]);
  const int nc  = C[0] * C[1] * C[2];
  s.chunks.resize(nc);
  for (int cz = 0; cz < C[0]; cz++)
    for (int cy = 0; cy < C[1]; cy++)
      for (int cx = 0; cx < C[2]; cx++) {
        int    id  = s.id_of_coord[(cz * C[1] + cy) * C[2] + cx];
        Chunk& c   = s.chunks[id];
        c.id       = id;
        c.coord[0] = cz;
        c.coord[1] = cy;
        c.coord[2] = cx;
      }

  const int Nb = (p.order + 3) / 2; // pic/pic_chunk.cpp:195
  for (auto& c : s.chunks) {
    for (int d = 0; d < 3; d++) {
      c.has_dim[d] = has_dim[d];
      c.dims[d]    = N[d] / C[d];
      c.offset[d]  = c.coord[d] * N[d] / C[d];
      c.del[d]     = p.delh;
      c.Lb[d]      = Nb;
      c.Ub[d]      = has_dim[d] ? Nb + c.dims[d] - 1 : Nb;
      c.n[d]       = c.dims[d] + 2 * Nb;
      // nix/chunk.cpp:210-237 (set_coordinate)
      c.lim[d][0]  = c.offset[d] * c.del[d];
      c.lim[d][1]  = c.offset[d] * c.del[d] + c.dims[d] * c.del[d];
      c.glim[d][0] = 0.0;
      c.glim[d][1] = N[d] * c.del[d];
      c.rng[d][0]  = has_dim[d] ? c.lim[d][0] : -std::numeric_limits<double>::max();
      c.rng[d][1]  = has_dim[d] ? c.lim[d][1] : +std::numeric_limits<double>::max();
    }
    c.Nb = Nb;
    // neighbours, nix/chunkvector.hpp:56-81 with periodic nix/chunkmap.cpp:100-112
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int cc3[3] = {c.coord[0] + dz, c.coord[1] + dy, c.coord[2] + dx};
          if (!xper && (cc3[2] < 0 || cc3[2] >= C[2])) {
            c.nb_id[dir_index(dz + 1, dy + 1, dx + 1)] = -1; // MPI_PROC_NULL
            continue;
          }
          for (int d = 0; d < 3; d++) {
            cc3[d] = cc3[d] >= 0 ? cc3[d] : C[d] - 1;
            cc3[d] = cc3[d] < C[d] ? cc3[d] : 0;
          }
          c.nb_id[dir_index(dz + 1, dy + 1, dx + 1)] =
              s.id_of_coord[(cc3[0] * C[1] + cc3[1]) * C[2] + cc3[2]];
        }
    // pic/pic_chunk.cpp:106-122 (allocate) + example field init
    size_t ncell = (size_t)c.n[0] * c.n[1] * c.n[2];
    c.uf.assign(ncell * 6, 0.0);
    c.uj.assign(ncell * 4, 0.0);
    c.ff.assign(ncell * 9, 0.0);
    c.um.assign(ncell * s.Ns * 14, 0.0);
    c.wall_lo     = !xper && c.coord[2] == 0;
    c.wall_hi     = !xper && c.coord[2] == C[2] - 1;
    c.random_seed = c.id + p.seed_offset; // seed_type = 'fixed' (pic/pic_chunk.cpp:243-254)
    if (p.problem == 1) {
      init_shock_fields(s, c);
      init_friedman(c);
      continue;
    }
    for (int iz = c.Lb[0]; iz <= c.Ub[0]; iz++)
      for (int iy = c.Lb[1]; iy <= c.Ub[1]; iy++)
        for (int ix = c.Lb[2]; ix <= c.Ub[2]; ix++) {
          c.UF(iz, iy, ix, 0) = p.Ex;
          c.UF(iz, iy, ix, 1) = p.Ey;
          c.UF(iz, iy, ix, 2) = p.Ez;
          c.UF(iz, iy, ix, 3) = p.Bx;
          c.UF(iz, iy, ix, 4) = p.By;
          c.UF(iz, iy, ix, 5) = p.Bz;
        }
    init_friedman(c);
  }

  if (p.problem == 1) {
    for_chunks(s, [&](Chunk& c) {
      int Ng = 1;
      for (int d = 0; d < 3; d++)
        Ng *= (c.Ub[d] - c.Lb[d] + 2 * c.Nb + 1);
      c.up.assign(s.Ns, Particles());
      for (int is = 0; is < s.Ns; is++) {
        Particles& P = c.up[is];
        P.m          = is == 0 ? s.sh.me : s.sh.mi;
        P.q          = is == 0 ? s.sh.qe : s.sh.qi;
        P.Ng         = Ng;
        P.pindex.assign(Ng + 1, 0);
        P.pcount.assign((size_t)(Ng + 1) * SIMD_WIDTH, 0);
      }
      if (p.load)
        load_shock_particles(s, c, N);
      for (int is = 0; is < s.Ns; is++) {
        count(c, c.up[is], 0, c.up[is].Np - 1, true, s.order);
        sort(c.up[is]);
      }
    });
    exchange_field(s, 0);
    return;
  }
  // particles: example/beam/main.cpp:67-116 (thermal is the vx=vy=vz=0 case)
  for_chunks(s, [&](Chunk& c) {
    int  nb = c.Nb;
    int  Ng = 1;
    for (int d = 0; d < 3; d++)
      Ng *= (c.Ub[d] - c.Lb[d] + 2 * nb + 1); // nix/particle.hpp:139-146
    int  seed = c.id + p.seed_offset;
    MT64 mtp(seed), mtv(seed);
    Normal normal;
    c.up.assign(s.Ns, Particles());
    for (int is = 0; is < s.Ns; is++) {
      Particles& P = c.up[is];
      int        np = p.np[is];
      int        mp = np * c.dims[0] * c.dims[1] * c.dims[2];
      P.Ng          = Ng;
      P.m           = p.ro[is] / np;
      P.q           = p.qm[is] * P.m;
      P.pindex.assign(Ng + 1, 0);
      P.pcount.assign((size_t)(Ng + 1) * SIMD_WIDTH, 0);
      if (!p.load) {
        P.Np = 0;
        continue;
      }
      P.reserve((int)(mp * (1 + p.buffer_ratio)) + 1);
      P.Np = mp;
      mtp.seed(seed);
      int64_t id = (int64_t)mp * c.id;
      for (int ip = 0; ip < mp; ip++) {
        double* x = P.XU(ip);
        x[0]      = uniform01(mtp) * (c.lim[2][1] - c.lim[2][0]) + c.lim[2][0];
        x[1]      = uniform01(mtp) * (c.lim[1][1] - c.lim[1][0]) + c.lim[1][0];
        x[2]      = uniform01(mtp) * (c.lim[0][1] - c.lim[0][0]) + c.lim[0][0];
        x[3]      = normal(mtv) * p.vt[is] + p.vx[is];
        x[4]      = normal(mtv) * p.vt[is] + p.vy[is];
        x[5]      = normal(mtv) * p.vt[is] + p.vz[is];
        int64_t v = id + ip;
        std::memcpy(&x[6], &v, sizeof(v));
      }
    }
    // initial sort (pic/pic_chunk.cpp:447-453)
    for (int is = 0; is < s.Ns; is++) {
      count(c, c.up[is], 0, c.up[is].Np - 1, true, s.order);
      sort(c.up[is]);
    }
  });
  // initial EMF exchange (pic/pic_application.cpp:106-130)
  exchange_field(s, 0);
}


// --------------------------------------------------------------------------
// Tangent-linear reference of the adjoint form: the 2D periodic order-1 step of
// the hot path (push_bfd_2d / push_efd_2d, pic/engine/maxwell.hpp:231-348; MC
// gather + Boris, pic/engine/velocity.hpp:390-418, nix/primitives.hpp:163-189;
// position push, pic/engine/position.hpp:117-130; Esirkepov 2D,
// nix/esirkepov.hpp:76-143) evaluated in forward-mode dual numbers: every
// state variable carries its derivative along one direction of the initial
// state, and L = 1/2 sum B^2 carries dL along it.
// --------------------------------------------------------------------------
struct TD {
  double v, d;
  TD() : v(0), d(0) {}
  TD(double x) : v(x), d(0) {}
  TD(double x, double dx) : v(x), d(dx) {}
};
inline TD operator+(TD a, TD b) { return TD(a.v + b.v, a.d + b.d); }
inline TD operator-(TD a, TD b) { return TD(a.v - b.v, a.d - b.d); }
inline TD operator-(TD a) { return TD(-a.v, -a.d); }
inline TD operator*(TD a, TD b) { return TD(a.v * b.v, a.d * b.v + a.v * b.d); }
inline TD operator/(TD a, TD b)
{
  double r = a.v / b.v;
  return TD(r, (a.d - r * b.d) / b.v);
}
inline TD& operator+=(TD& a, TD b) { return a = a + b; }
inline TD& operator-=(TD& a, TD b) { return a = a - b; }
inline TD& operator*=(TD& a, TD b) { return a = a * b; }
inline TD tsqrt(TD a)
{
  double r = std::sqrt(a.v);
  return TD(r, a.d * 0.5 / r);
}

struct TLGeom {
  int    nx, ny;
  double dx, dt, cc, q, qmdt, lx, ly;
};

inline size_t tl_cell(const TLGeom& g, int iy, int ix)
{
  iy %= g.ny;
  iy = iy < 0 ? iy + g.ny : iy;
  ix %= g.nx;
  ix = ix < 0 ? ix + g.nx : ix;
  return (size_t)iy * g.nx + ix;
}

// shape_mc1 (nix/primitives.hpp:256-266) in dual numbers
inline void tl_shape1(TD x, double X, double rdx, TD* s)
{
  TD delta = (x - TD(X)) * TD(rdx);
  s[0]     = TD(1.0) - delta;
  s[1]     = delta;
}

// weights2d at order 1: integer grid (centres) and half grid (faces)
inline void tl_weights(const TLGeom& g, TD x, TD wi[2], TD wh[2], int& i0, int& h0)
{
  const double rdx = 1 / g.dx;
  i0               = digitize(x.v, 0.5 * g.dx, rdx);
  h0               = digitize(x.v, 0.0, rdx);
  tl_shape1(x, 0.5 * g.dx + static_cast<double>(i0) * g.dx, rdx, wi);
  tl_shape1(x, 0.0 + static_cast<double>(h0) * g.dx, rdx, wh);
}

inline TD tl_interp(const TLGeom& g, const std::vector<TD>& F, size_t off, int y0, int x0,
                    const TD wy[2], const TD wx[2])
{
  TD ry(0.0);
  for (int jy = 0; jy < 2; jy++) {
    TD rx(0.0);
    for (int jx = 0; jx < 2; jx++)
      rx += F[off + tl_cell(g, y0 + jy, x0 + jx)] * wx[jx];
    ry += rx * wy[jy];
  }
  return ry * TD(g.qmdt);
}

// push_boris (nix/primitives.hpp:163-189)
inline void tl_boris(TD& ux, TD& uy, TD& uz, TD ex, TD ey, TD ez, TD bx, TD by, TD bz, double cc)
{
  ux += ex;
  uy += ey;
  uz += ez;
  TD gm = TD(1.0) / tsqrt(TD(cc * cc) + ux * ux + uy * uy + uz * uz);
  bx *= gm;
  by *= gm;
  bz *= gm;
  TD bb = TD(2.0) / (TD(1.0) + bx * bx + by * by + bz * bz);
  TD vx = ux + (uy * bz - uz * by);
  TD vy = uy + (uz * bx - ux * bz);
  TD vz = uz + (ux * by - uy * bx);
  ux += (vy * bz - vz * by) * bb + ex;
  uy += (vz * bx - vx * bz) * bb + ey;
  uz += (vx * by - vy * bx) * bb + ez;
}

// Esirkepov 2D, order 1, with shift_weights; J[3][ny][nx] accumulated
inline void tl_deposit(const TLGeom& g, TD x0, TD y0, TD z0, TD x1, TD y1, TD z1,
                       std::vector<TD>& J)
{
  const int    N = 4;
  const double A = 1.0 / 2, B = 1.0 / 3, rdx = 1 / g.dx;
  TD           s0x[N], s0y[N], s1x[N], s1y[N];
  const int    ix0 = digitize(x0.v, 0.5 * g.dx, rdx), iy0 = digitize(y0.v, 0.5 * g.dx, rdx);
  const int    ix1 = digitize(x1.v, 0.5 * g.dx, rdx), iy1 = digitize(y1.v, 0.5 * g.dx, rdx);
  tl_shape1(x0, 0.5 * g.dx + static_cast<double>(ix0) * g.dx, rdx, &s0x[1]);
  tl_shape1(y0, 0.5 * g.dx + static_cast<double>(iy0) * g.dx, rdx, &s0y[1]);
  tl_shape1(x1, 0.5 * g.dx + static_cast<double>(ix1) * g.dx, rdx, &s1x[1]);
  tl_shape1(y1, 0.5 * g.dx + static_cast<double>(iy1) * g.dx, rdx, &s1y[1]);
  if (ix1 - ix0 < 0)
    for (int ii = 0; ii < 3; ii++)
      s1x[ii] = s1x[ii + 1];
  else if (ix1 - ix0 > 0)
    for (int ii = 3; ii > 0; ii--)
      s1x[ii] = s1x[ii - 1];
  if (iy1 - iy0 < 0)
    for (int ii = 0; ii < 3; ii++)
      s1y[ii] = s1y[ii + 1];
  else if (iy1 - iy0 > 0)
    for (int ii = 3; ii > 0; ii--)
      s1y[ii] = s1y[ii - 1];
  const int    ox = ix0 - 1, oy = iy0 - 1;
  const TD     vz = (z1 - z0) / TD(g.dt);
  const size_t v  = (size_t)g.nx * g.ny;
  for (int l = 0; l < N; l++) {
    s1x[l] -= s0x[l];
    s1y[l] -= s0y[l];
  }
  const TD qdxdt(g.q * (g.dx / g.dt)), qvz = TD(g.q) * vz;
  for (int jy = 0; jy < N; jy++) {
    TD ww(0.0), wx = -(s0y[jy] + TD(A) * s1y[jy]) * qdxdt;
    for (int jx = 0; jx < N - 1; jx++) {
      ww += s1x[jx] * wx;
      J[tl_cell(g, oy + jy, ox + jx + 1)] += ww;
    }
  }
  for (int jx = 0; jx < N; jx++) {
    TD ww(0.0), wy = -(s0x[jx] + TD(A) * s1x[jx]) * qdxdt;
    for (int jy = 0; jy < N - 1; jy++) {
      ww += s1y[jy] * wy;
      J[v + tl_cell(g, oy + jy + 1, ox + jx)] += ww;
    }
  }
  for (int jy = 0; jy < N; jy++)
    for (int jx = 0; jx < N; jx++)
      J[2 * v + tl_cell(g, oy + jy, ox + jx)] +=
          ((TD(1.0) * s0x[jx] + TD(A) * s1x[jx]) * s0y[jy] +
           (TD(A) * s0x[jx] + TD(B) * s1x[jx]) * s1y[jy]) *
          qvz;
}

void tl_bfd(const TLGeom& g, const std::vector<TD>& E, std::vector<TD>& Bf, double half)
{
  const double cfl = g.cc * (half * g.dt) / g.dx;
  const size_t v   = (size_t)g.nx * g.ny;
  std::vector<TD> B = Bf;
  for (int iy = 0; iy < g.ny; iy++)
    for (int ix = 0; ix < g.nx; ix++) {
      size_t i = tl_cell(g, iy, ix), jx = tl_cell(g, iy, ix - 1), jy = tl_cell(g, iy - 1, ix);
      B[i] += TD(-cfl) * (E[2 * v + i] - E[2 * v + jy]);
      B[v + i] += TD(+cfl) * (E[2 * v + i] - E[2 * v + jx]);
      B[2 * v + i] += TD(-cfl) * (E[v + i] - E[v + jx]) + TD(+cfl) * (E[i] - E[jy]);
    }
  Bf = B;
}

void tl_efd(const TLGeom& g, std::vector<TD>& E, const std::vector<TD>& B, const std::vector<TD>& J)
{
  const double cfl = g.cc * g.dt / g.dx;
  const size_t v   = (size_t)g.nx * g.ny;
  for (int iy = 0; iy < g.ny; iy++)
    for (int ix = 0; ix < g.nx; ix++) {
      size_t i = tl_cell(g, iy, ix), px = tl_cell(g, iy, ix + 1), py = tl_cell(g, iy + 1, ix);
      E[i] += TD(+cfl) * (B[2 * v + py] - B[2 * v + i]) - TD(g.dt) * J[i];
      E[v + i] += TD(-cfl) * (B[2 * v + px] - B[2 * v + i]) - TD(g.dt) * J[v + i];
      E[2 * v + i] += TD(+cfl) * (B[v + px] - B[v + i]) + TD(-cfl) * (B[py] - B[i]) -
                      TD(g.dt) * J[2 * v + i];
    }
}

} // namespace

extern "C" void pno_adj_tangent(int nx, int ny, double delh, double delt, double cc, double q,
                                double m, int64_t np, const double* xu6, const double* dxu6,
                                int nsteps, double* loss, double* dloss)
{
  TLGeom g{nx, ny, delh, delt, cc, q, 0.5 * q / m * delt, nx * delh, ny * delh};
  const size_t    v = (size_t)nx * ny;
  std::vector<TD> P(6 * (size_t)np), E(3 * v), B(3 * v);
  for (int64_t p = 0; p < np; p++)
    for (int c = 0; c < 6; c++)
      P[6 * p + c] = TD(xu6[6 * p + c], dxu6[6 * p + c]);
  const int nthr = g_threads > 0 ? g_threads : omp_get_max_threads();
  for (int st = 0; st < nsteps; st++) {
    std::vector<TD> B1 = B;
    tl_bfd(g, E, B1, 0.5);
    std::vector<std::vector<TD>> Jt(nthr, std::vector<TD>(3 * v));
#pragma omp parallel num_threads(nthr)
    {
      std::vector<TD>& J = Jt[omp_get_thread_num()];
#pragma omp for schedule(static)
      for (int64_t p = 0; p < np; p++) {
        TD* s = &P[6 * p];
        TD  wix[2], whx[2], wiy[2], why[2];
        int ix0, hx0, iy0, hy0;
        tl_weights(g, s[0], wix, whx, ix0, hx0);
        tl_weights(g, s[1], wiy, why, iy0, hy0);
        TD eb[6] = {tl_interp(g, E, 0, iy0, hx0, wiy, whx), tl_interp(g, E, v, hy0, ix0, why, wix),
                    tl_interp(g, E, 2 * v, iy0, ix0, wiy, wix),
                    tl_interp(g, B1, 0, hy0, ix0, why, wix),
                    tl_interp(g, B1, v, iy0, hx0, wiy, whx),
                    tl_interp(g, B1, 2 * v, hy0, hx0, why, whx)};
        tl_boris(s[3], s[4], s[5], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], g.cc);
        TD gm = tsqrt(TD(1.0) + (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]) * TD(1 / g.cc) *
                                    TD(1 / g.cc));
        TD dt = TD(g.dt) / gm;
        TD x1 = s[0] + s[3] * dt, y1 = s[1] + s[4] * dt, z1 = s[2] + s[5] * dt;
        tl_deposit(g, s[0], s[1], s[2], x1, y1, z1, J);
        // set_boundary_periodic: the shift does not change the derivative
        x1.v = x1.v + (x1.v < 0) * g.lx - (x1.v >= g.lx) * g.lx;
        y1.v = y1.v + (y1.v < 0) * g.ly - (y1.v >= g.ly) * g.ly;
        s[0] = x1;
        s[1] = y1;
        s[2] = z1;
      }
    }
    std::vector<TD> J(3 * v);
    for (int t = 0; t < nthr; t++)
      for (size_t i = 0; i < 3 * v; i++)
        J[i] += Jt[t][i];
    B = B1;
    tl_bfd(g, E, B, 0.5);
    tl_efd(g, E, B, J);
  }
  double L = 0, dL = 0;
  for (size_t i = 0; i < v; i++) {
    L += 0.5 * (B[i].v * B[i].v + B[v + i].v * B[v + i].v + B[2 * v + i].v * B[2 * v + i].v);
    dL += B[i].v * B[i].d + B[v + i].v * B[v + i].d + B[2 * v + i].v * B[2 * v + i].d;
  }
  *loss  = L;
  *dloss = dL;
}

namespace {
} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
struct pno_sim {
  Sim s;
};

extern "C" {

pno_sim* pno_create(const pno_params* p)
{
  auto* h = new pno_sim();
  h->s.p  = *p;
  setup(h->s);
  return h;
}

void pno_destroy(pno_sim* h) { delete h; }

int pno_num_chunks(const pno_sim* h) { return (int)h->s.chunks.size(); }

void pno_chunk_info(const pno_sim* h, int id, int* info)
{
  const Chunk& c = h->s.chunks[id];
  for (int d = 0; d < 3; d++) {
    info[d]     = c.coord[d];
    info[3 + d] = c.offset[d];
    info[6 + d] = c.n[d];
  }
  info[9]  = c.Nb;
  info[10] = c.up.empty() ? 0 : c.up[0].Ng;
}

int64_t pno_get_np(const pno_sim* h, int chunk, int species)
{
  return h->s.chunks[chunk].up[species].Np;
}

void pno_get_particles(const pno_sim* h, int chunk, int species, double* xu7)
{
  const Particles& P = h->s.chunks[chunk].up[species];
  std::memcpy(xu7, P.xu.data(), (size_t)P.Np * NC * sizeof(double));
}

void pno_get_particles_old(const pno_sim* h, int chunk, int species, double* xv7)
{
  const Particles& P = h->s.chunks[chunk].up[species];
  std::memcpy(xv7, P.xv.data(), (size_t)P.Np * NC * sizeof(double));
}

void pno_set_particles(pno_sim* h, int chunk, int species, const double* xu7, int64_t np)
{
  Particles& P = h->s.chunks[chunk].up[species];
  P.reserve((int)np + 1);
  std::memcpy(P.xu.data(), xu7, (size_t)np * NC * sizeof(double));
  P.Np = (int)np;
}

void pno_get_gindex(const pno_sim* h, int chunk, int species, int32_t* out)
{
  const Particles& P = h->s.chunks[chunk].up[species];
  std::memcpy(out, P.gindex.data(), (size_t)P.Np * sizeof(int32_t));
}

void pno_get_pindex(const pno_sim* h, int chunk, int species, int32_t* out)
{
  const Particles& P = h->s.chunks[chunk].up[species];
  std::memcpy(out, P.pindex.data(), (size_t)(P.Ng + 1) * sizeof(int32_t));
}

void pno_get_field(const pno_sim* h, int chunk, int which, double* out)
{
  const Chunk& c = h->s.chunks[chunk];
  const auto&  v = which == 0 ? c.uf : (which == 1 ? c.uj : (which == 2 ? c.ff : c.um));
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}

void pno_set_field(pno_sim* h, int chunk, int which, const double* in)
{
  Chunk& c = h->s.chunks[chunk];
  auto&  v = which == 0 ? c.uf : (which == 1 ? c.uj : (which == 2 ? c.ff : c.um));
  std::memcpy(v.data(), in, v.size() * sizeof(double));
}

void pno_push_bfd(pno_sim* h, double delt)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) { push_bfd(c, s.dim, s.p.cc, s.p.friedman, delt); });
}

void pno_push_efd(pno_sim* h, double delt)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) { push_efd(c, s.dim, s.p.cc, s.p.friedman, delt); });
}

void pno_push_velocity(pno_sim* h, double delt)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) {
    dispatch<VelocityOp>(s.dim, s.order, c, s.Ns, s.p.cc, s.p.pusher, s.p.interp, delt);
  });
}

void pno_push_position(pno_sim* h, double delt)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) { push_position_chunk(s, c, s.Ns, s.p.cc, s.order, delt); });
}

void pno_deposit_current(pno_sim* h, double delt)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) { dispatch<CurrentOp>(s.dim, s.order, c, s.Ns, delt); });
}

void pno_deposit_moment(pno_sim* h)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) { dispatch<MomentOp>(s.dim, s.order, c, s.Ns, s.p.cc); });
  exchange_field(s, 2);
}

void pno_sort_particle(pno_sim* h)
{
  Sim& s = h->s;
  for_chunks(s, [&](Chunk& c) {
    for (int is = 0; is < s.Ns; is++) {
      count(c, c.up[is], 0, c.up[is].Np - 1, true, s.order);
      sort(c.up[is]);
    }
  });
}

void pno_exchange(pno_sim* h, int mode)
{
  if (mode == 3)
    exchange_particle(h->s);
  else
    exchange_field(h->s, mode);
}

// pic/pic_application.cpp:219-292 (push_openmp) with MPI replaced by in-memory exchange
void pno_step(pno_sim* h)
{
  Sim&         s    = h->s;
  const double delt = s.p.delt;
  for_chunks(s, [&](Chunk& c) {
    push_bfd(c, s.dim, s.p.cc, s.p.friedman, 0.5 * delt);
    dispatch<VelocityOp>(s.dim, s.order, c, s.Ns, s.p.cc, s.p.pusher, s.p.interp, delt);
    push_position_chunk(s, c, s.Ns, s.p.cc, s.order, delt);
    dispatch<CurrentOp>(s.dim, s.order, c, s.Ns, delt);
    push_bfd(c, s.dim, s.p.cc, s.p.friedman, 0.5 * delt);
  });
  exchange_field(s, 1);
  for_chunks(s, [&](Chunk& c) { push_efd(c, s.dim, s.p.cc, s.p.friedman, delt); });
  exchange_particle(s);
  exchange_field(s, 0);
}

// pic/diag/history.hpp:20-66 (sum over chunks in id order)
void pno_history(pno_sim* h, double* out)
{
  Sim& s = h->s;
  pno_deposit_moment(h);
  for (int i = 0; i < s.Ns + 4; i++)
    out[i] = 0;
  std::vector<double> ep(s.Ns);
  for (auto& c : s.chunks) {
    double de, db, ee, eb;
    get_diverror(c, s.dim, de, db);
    get_energy(c, s.Ns, s.p.cc, ee, eb, ep.data());
    out[0] += de;
    out[1] += db;
    out[2] += ee;
    out[3] += eb;
    for (int is = 0; is < s.Ns; is++)
      out[4 + is] += ep[is];
  }
}

void pno_set_threads(int n) { g_threads = n; }

// ---- primitives ----
void pno_shape_mc(int order, double x, double X, double rdx, double* s)
{
  switch (order) {
  case 1: shape_mc<1>(x, X, rdx, s); break;
  case 2: shape_mc<2>(x, X, rdx, s); break;
  case 3: shape_mc<3>(x, X, rdx, s); break;
  default: shape_mc<4>(x, X, rdx, s); break;
  }
}

void pno_shape_wt(int order, double x, double X, double rdx, double dt, double rdt, double* s)
{
  switch (order) {
  case 1: shape_wt<1>(x, X, rdx, dt, rdt, s); break;
  case 2: shape_wt<2>(x, X, rdx, dt, rdt, s); break;
  case 3: shape_wt<3>(x, X, rdx, dt, rdt, s); break;
  default: shape_wt<4>(x, X, rdx, dt, rdt, s); break;
  }
}

int pno_digitize(double x, double xmin, double rdx) { return digitize(x, xmin, rdx); }

double pno_lorentz_factor(double ux, double uy, double uz, double rc)
{
  return lorentz_factor(ux, uy, uz, rc);
}

void pno_push_boris(double* u, const double* eb, double cc)
{
  push_boris(u[0], u[1], u[2], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
}

void pno_push_vay(double* u, const double* eb, double cc)
{
  push_vay(u[0], u[1], u[2], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
}

void pno_push_higuera_cary(double* u, const double* eb, double cc)
{
  push_higuera_cary(u[0], u[1], u[2], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
}

void pno_init_friedman(pno_sim* h)
{
  for_chunks(h->s, [&](Chunk& c) { init_friedman(c); });
}

void pno_esirkepov3d(int order, const double* xv, const double* xu, double q, double delt,
                     double dx, double dy, double dz, double* tile, int* origin)
{
  Chunk c{};
  for (int d = 0; d < 3; d++) {
    c.Lb[d]     = 0;
    c.lim[d][0] = 0.0;
  }
  c.del[0] = dz;
  c.del[1] = dy;
  c.del[2] = dx;
  Geo g(c);
  int N    = order + 3;
  std::fill(tile, tile + N * N * N * 4, 0.0);
  switch (order) {
  case 1: local_current<3, 1>(g, xv, xu, q, delt, tile, origin); break;
  case 2: local_current<3, 2>(g, xv, xu, q, delt, tile, origin); break;
  case 3: local_current<3, 3>(g, xv, xu, q, delt, tile, origin); break;
  default: local_current<3, 4>(g, xv, xu, q, delt, tile, origin); break;
  }
}

void pno_esirkepov2d(int order, const double* xv, const double* xu, double q, double delt,
                     double dx, double dy, double* tile, int* origin)
{
  Chunk c{};
  for (int d = 0; d < 3; d++) {
    c.Lb[d]     = 0;
    c.lim[d][0] = 0.0;
  }
  c.del[0] = 1.0;
  c.del[1] = dy;
  c.del[2] = dx;
  Geo g(c);
  int N    = order + 3;
  std::fill(tile, tile + N * N * 4, 0.0);
  switch (order) {
  case 1: local_current<2, 1>(g, xv, xu, q, delt, tile, origin); break;
  case 2: local_current<2, 2>(g, xv, xu, q, delt, tile, origin); break;
  case 3: local_current<2, 3>(g, xv, xu, q, delt, tile, origin); break;
  default: local_current<2, 4>(g, xv, xu, q, delt, tile, origin); break;
  }
}

void pno_esirkepov1d(int order, const double* xv, const double* xu, double q, double delt,
                     double dx, double* tile, int* origin)
{
  Chunk c{};
  for (int d = 0; d < 3; d++) {
    c.Lb[d]     = 0;
    c.lim[d][0] = 0.0;
  }
  c.del[0] = 1.0;
  c.del[1] = 1.0;
  c.del[2] = dx;
  Geo g(c);
  int N    = order + 3;
  std::fill(tile, tile + N * 4, 0.0);
  switch (order) {
  case 1: local_current<1, 1>(g, xv, xu, q, delt, tile, origin); break;
  case 2: local_current<1, 2>(g, xv, xu, q, delt, tile, origin); break;
  case 3: local_current<1, 3>(g, xv, xu, q, delt, tile, origin); break;
  default: local_current<1, 4>(g, xv, xu, q, delt, tile, origin); break;
  }
}

uint64_t pno_mt19937_64(uint64_t seed, int64_t skip)
{
  MT64 g(seed);
  uint64_t v = 0;
  for (int64_t i = 0; i <= skip; i++)
    v = g();
  return v;
}

void pno_uniform(uint64_t seed, int64_t n, double* out)
{
  MT64 g(seed);
  for (int64_t i = 0; i < n; i++)
    out[i] = uniform01(g);
}

void pno_normal(uint64_t seed, int64_t n, double* out)
{
  MT64   g(seed);
  Normal nd;
  for (int64_t i = 0; i < n; i++)
    out[i] = nd(g);
}

void pno_gamma(double alpha, double beta, uint64_t seed, int64_t n, double* out)
{
  MT64  g(seed);
  Gamma gd(alpha, beta);
  for (int64_t i = 0; i < n; i++)
    out[i] = gd(g);
}

int pno_poisson(double mean, uint64_t seed, int64_t n, int* out)
{
  if (!(mean < 12))
    return -1;
  MT64         g(seed);
  const double thr = std::exp(-mean);
  for (int64_t i = 0; i < n; i++)
    out[i] = poisson_small(g, thr);
  return 0;
}

void pno_maxwell_juttner(double temperature, double drift, uint64_t seed, int64_t n,
                         double* out3)
{
  MT64           g(seed);
  MaxwellJuttner mj(temperature, drift);
  for (int64_t i = 0; i < n; i++)
    mj(g, out3[3 * i], out3[3 * i + 1], out3[3 * i + 2]);
}

void pno_shock_state(const pno_sim* h, int chunk, int64_t* out6)
{
  const Chunk& c = h->s.chunks[chunk];
  out6[0]        = c.random_seed;
  out6[1]        = c.lastid;
  out6[2]        = c.influx[0];
  out6[3]        = c.influx[1];
  out6[4]        = c.efflux[0];
  out6[5]        = c.efflux[1];
}

void pno_chunk_map(int Cz, int Cy, int Cx, int* id_of_coord)
{
  auto m = chunk_map(Cz, Cy, Cx);
  std::memcpy(id_of_coord, m.data(), m.size() * sizeof(int));
}

} // extern "C"
