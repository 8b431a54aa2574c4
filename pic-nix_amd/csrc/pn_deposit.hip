// pn_deposit.hip -- charge-conserving current deposition (Esirkepov 2001).
//
// Per particle this evaluates exactly the reference's BaseCurrent::local{2,3}d
// (pic/engine/current.hpp:275-395) with nix::esirkepov::deposit{2,3}d
// (nix/esirkepov.hpp:76-237, 287-337): the old and new shape weights, the shift
// by the cell crossing, rho from the new weights and the Jx/Jy/Jz prefix sums.
// Each contribution value equals the reference's per-particle tile entry; only
// the order in which tiles are summed into J differs (the reference adds them
// particle by particle, here they meet in LDS or global atomics), so J agrees to
// round-off.
//
// Slab kernel: one block per (z-plane of bins, segment) for 3D (a y-row for 2D);
// particles of a plane are contiguous after the sort.  The block accumulates the
// plane's (Order+3) node planes x full ghost-inclusive (y, x) extent in LDS, with
// consecutive lanes assigned to particles ~one cell apart so LDS atomics rarely
// collide, then flushes the tile to the global padded J with one atomic pass.
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

struct DepConsts {
  double q[kMaxSpecies];
  double delt;
  double dxdt, dydt, dzdt;
};

// Evaluate one particle's contributions; `add(lz, ly, lx, comp, value)` receives
// chunk-local ghost-inclusive node indices.
template <int Dim, int Order, typename Add>
__device__ __forceinline__ void esirkepov(const Geom& g, double xmin, double ymin, double zmin,
                                          const double xv[3], const double xu[3], double qs,
                                          const DepConsts& dc, Add add)
{
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int N      = Order + 3;
  const double  dx = g.del[2], dy = g.del[1], dz = g.del[0];
  const double  rdx = 1 / dx, rdy = 1 / dy, rdz = 1 / dz;
  const double  ximin = xmin + 0.5 * dx * is_odd;
  const double  yimin = ymin + 0.5 * dy * is_odd;
  const double  zimin = zmin + 0.5 * dz * is_odd;
  const double  xgrid = xmin + 0.5 * dx;
  const double  ygrid = ymin + 0.5 * dy;
  const double  zgrid = zmin + 0.5 * dz;
  const int     lb    = g.Nb;
  const double  A     = 1.0 / 2;
  const double  B     = 1.0 / 3;

  if constexpr (Dim == 1) {
    // local1d (pic/engine/current.hpp:207-258) + deposit1d (nix/esirkepov.hpp:
    // 18-67,277-291): rho, Jx by the prefix sum, Jy and Jz from the velocities
    // (x_new - x_old) / dt times the time-averaged weight
    double s0[N], s1[N];
#pragma unroll
    for (int l = 0; l < N; l++)
      s0[l] = s1[l] = 0.0;
    const int i0 = digitize(xv[0], ximin, rdx);
    const int i1 = digitize(xu[0], ximin, rdx);
    shape_mc<Order>(xv[0], xgrid + static_cast<double>(i0) * dx, rdx, &s0[1]);
    shape_mc<Order>(xu[0], xgrid + static_cast<double>(i1) * dx, rdx, &s1[1]);
    if (i1 - i0 < 0) {
#pragma unroll
      for (int ii = 0; ii < Order + 2; ii++)
        s1[ii] = s1[ii + 1];
    } else if (i1 - i0 > 0) {
#pragma unroll
      for (int ii = Order + 2; ii > 0; ii--)
        s1[ii] = s1[ii - 1];
    }
    const int    ox = i0 + lb - (Order / 2) - 1;
    const double vy = (xu[1] - xv[1]) / dc.delt;
    const double vz = (xu[2] - xv[2]) / dc.delt;
#pragma unroll
    for (int jx = 0; jx < N; jx++)
      add(lb, lb, ox + jx, 0, qs * s1[jx]);
#pragma unroll
    for (int l = 0; l < N; l++)
      s1[l] -= s0[l];
    {
      double       ww = 0;
      const double wx = -(qs * dc.dxdt);
#pragma unroll
      for (int jx = 0; jx < N - 1; jx++) {
        ww += s1[jx] * wx;
        add(lb, lb, ox + jx + 1, 1, ww);
      }
    }
    const double qvy = qs * vy;
    const double qvz = qs * vz;
#pragma unroll
    for (int jx = 0; jx < N; jx++) {
      add(lb, lb, ox + jx, 2, (s0[jx] + A * s1[jx]) * qvy);
      add(lb, lb, ox + jx, 3, (s0[jx] + A * s1[jx]) * qvz);
    }
    (void)yimin;
    (void)zimin;
    (void)ygrid;
    (void)zgrid;
    (void)rdy;
    (void)rdz;
    (void)B;
    return;
  }
  double s0x[N], s0y[N], s0z[N], s1x[N], s1y[N], s1z[N];
#pragma unroll
  for (int l = 0; l < N; l++) {
    s0x[l] = s0y[l] = s0z[l] = 0.0;
    s1x[l] = s1y[l] = s1z[l] = 0.0;
  }
  int ix0 = digitize(xv[0], ximin, rdx);
  int iy0 = digitize(xv[1], yimin, rdy);
  shape_mc<Order>(xv[0], xgrid + static_cast<double>(ix0) * dx, rdx, &s0x[1]);
  shape_mc<Order>(xv[1], ygrid + static_cast<double>(iy0) * dy, rdy, &s0y[1]);
  int ix1 = digitize(xu[0], ximin, rdx);
  int iy1 = digitize(xu[1], yimin, rdy);
  shape_mc<Order>(xu[0], xgrid + static_cast<double>(ix1) * dx, rdx, &s1x[1]);
  shape_mc<Order>(xu[1], ygrid + static_cast<double>(iy1) * dy, rdy, &s1y[1]);
  int iz0 = 0;
  if constexpr (Dim == 3) {
    iz0     = digitize(xv[2], zimin, rdz);
    int iz1 = digitize(xu[2], zimin, rdz);
    shape_mc<Order>(xv[2], zgrid + static_cast<double>(iz0) * dz, rdz, &s0z[1]);
    shape_mc<Order>(xu[2], zgrid + static_cast<double>(iz1) * dz, rdz, &s1z[1]);
    // shift_weights (nix/esirkepov.hpp:239-270), z
    if (iz1 - iz0 < 0) {
#pragma unroll
      for (int ii = 0; ii < Order + 2; ii++)
        s1z[ii] = s1z[ii + 1];
    } else if (iz1 - iz0 > 0) {
#pragma unroll
      for (int ii = Order + 2; ii > 0; ii--)
        s1z[ii] = s1z[ii - 1];
    }
  }
  if (ix1 - ix0 < 0) {
#pragma unroll
    for (int ii = 0; ii < Order + 2; ii++)
      s1x[ii] = s1x[ii + 1];
  } else if (ix1 - ix0 > 0) {
#pragma unroll
    for (int ii = Order + 2; ii > 0; ii--)
      s1x[ii] = s1x[ii - 1];
  }
  if (iy1 - iy0 < 0) {
#pragma unroll
    for (int ii = 0; ii < Order + 2; ii++)
      s1y[ii] = s1y[ii + 1];
  } else if (iy1 - iy0 > 0) {
#pragma unroll
    for (int ii = Order + 2; ii > 0; ii--)
      s1y[ii] = s1y[ii - 1];
  }

  const int ox = ix0 + lb - (Order / 2) - 1;
  const int oy = iy0 + lb - (Order / 2) - 1;

  if constexpr (Dim == 3) {
    const int oz = iz0 + lb - (Order / 2) - 1;
    // ro3d (esirkepov.hpp:145-155): rho from the new weights
#pragma unroll
    for (int jz = 0; jz < N; jz++)
#pragma unroll
      for (int jy = 0; jy < N; jy++)
#pragma unroll
        for (int jx = 0; jx < N; jx++)
          add(oz + jz, oy + jy, ox + jx, 0, qs * s1x[jx] * s1y[jy] * s1z[jz]);
    // ds3d: DS = S1 - S0
#pragma unroll
    for (int l = 0; l < N; l++) {
      s1x[l] -= s0x[l];
      s1y[l] -= s0y[l];
      s1z[l] -= s0z[l];
    }
    const double qdxdt = qs * dc.dxdt;
    const double qdydt = qs * dc.dydt;
    const double qdzdt = qs * dc.dzdt;
    // jx3d (esirkepov.hpp:168-188)
#pragma unroll
    for (int jz = 0; jz < N; jz++)
#pragma unroll
      for (int jy = 0; jy < N; jy++) {
        double ww = 0;
        double wx =
            -((1 * s0y[jy] + A * s1y[jy]) * s0z[jz] + (A * s0y[jy] + B * s1y[jy]) * s1z[jz]) *
            qdxdt;
#pragma unroll
        for (int jx = 0; jx < N - 1; jx++) {
          ww += s1x[jx] * wx;
          add(oz + jz, oy + jy, ox + jx + 1, 1, ww);
        }
      }
    // jy3d (esirkepov.hpp:190-210)
#pragma unroll
    for (int jz = 0; jz < N; jz++)
#pragma unroll
      for (int jx = 0; jx < N; jx++) {
        double ww = 0;
        double wy =
            -((1 * s0z[jz] + A * s1z[jz]) * s0x[jx] + (A * s0z[jz] + B * s1z[jz]) * s1x[jx]) *
            qdydt;
#pragma unroll
        for (int jy = 0; jy < N - 1; jy++) {
          ww += s1y[jy] * wy;
          add(oz + jz, oy + jy + 1, ox + jx, 2, ww);
        }
      }
    // jz3d (esirkepov.hpp:212-232)
#pragma unroll
    for (int jy = 0; jy < N; jy++)
#pragma unroll
      for (int jx = 0; jx < N; jx++) {
        double ww = 0;
        double wz =
            -((1 * s0x[jx] + A * s1x[jx]) * s0y[jy] + (A * s0x[jx] + B * s1x[jx]) * s1y[jy]) *
            qdzdt;
#pragma unroll
        for (int jz = 0; jz < N - 1; jz++) {
          ww += s1z[jz] * wz;
          add(oz + jz + 1, oy + jy, ox + jx, 3, ww);
        }
      }
  } else {
    const int oz = lb;
    // ro2d / ds2d / jx2d / jy2d / jz2d (esirkepov.hpp:76-143); Jz from the
    // velocity (x_new - x_old) / dt (pic/engine/current.hpp:285)
    const double vz = (xu[2] - xv[2]) / dc.delt;
#pragma unroll
    for (int jy = 0; jy < N; jy++)
#pragma unroll
      for (int jx = 0; jx < N; jx++)
        add(oz, oy + jy, ox + jx, 0, qs * s1x[jx] * s1y[jy]);
#pragma unroll
    for (int l = 0; l < N; l++) {
      s1x[l] -= s0x[l];
      s1y[l] -= s0y[l];
    }
    const double qdxdt = qs * dc.dxdt;
    const double qdydt = qs * dc.dydt;
    const double qvz   = qs * vz;
#pragma unroll
    for (int jy = 0; jy < N; jy++) {
      double ww = 0;
      double wx = -(s0y[jy] + A * s1y[jy]) * qdxdt;
#pragma unroll
      for (int jx = 0; jx < N - 1; jx++) {
        ww += s1x[jx] * wx;
        add(oz, oy + jy, ox + jx + 1, 1, ww);
      }
    }
#pragma unroll
    for (int jx = 0; jx < N; jx++) {
      double ww = 0;
      double wy = -(s0x[jx] + A * s1x[jx]) * qdydt;
#pragma unroll
      for (int jy = 0; jy < N - 1; jy++) {
        ww += s1y[jy] * wy;
        add(oz, oy + jy + 1, ox + jx, 2, ww);
      }
    }
#pragma unroll
    for (int jy = 0; jy < N; jy++)
#pragma unroll
      for (int jx = 0; jx < N; jx++) {
        double wz = ((1 * s0x[jx] + A * s1x[jx]) * s0y[jy] + (A * s0x[jx] + B * s1x[jx]) * s1y[jy]) *
                    qvz;
        add(oz, oy + jy, ox + jx, 3, wz);
      }
  }
}

__device__ __forceinline__ long gnode(const Geom& g, const int* __restrict__ off, int c, int lz,
                                      int ly, int lx)
{
  // chunk-local ghost-inclusive node -> global padded index (absent y, z map to 0)
  int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
  int gy = g.has[1] ? off[c * 3 + 1] + ly : 0;
  int gx = off[c * 3 + 2] + lx;
  return ((long)gz * g.P[1] + gy) * g.P[2] + gx;
}

// As gnode, or -1 for a node outside the padded grid (a particle placed outside
// its chunk through pn_set_particles): such contributions are dropped.
__device__ __forceinline__ long gnode_checked(const Geom& g, const int* __restrict__ off, int c,
                                              int lz, int ly, int lx)
{
  int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
  int gy = g.has[1] ? off[c * 3 + 1] + ly : 0;
  int gx = off[c * 3 + 2] + lx;
  if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
      (unsigned)gx >= (unsigned)g.P[2])
    return -1;
  return ((long)gz * g.P[1] + gy) * g.P[2] + gx;
}

// Value of lane (lane ^ M) within each quad of lanes, M = 1 or 2 (DPP
// quad_perm; no LDS traffic).
template <int M>
__device__ __forceinline__ double quad_xor(double x)
{
  constexpr int ctrl = M == 1 ? 0xB1 /* quad_perm [1,0,3,2] */ : 0x4E /* [2,3,0,1] */;
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo     = __builtin_amdgcn_mov_dpp(lo, ctrl, 0xF, 0xF, true);
  hi     = __builtin_amdgcn_mov_dpp(hi, ctrl, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Generic form: one thread per particle, contributions go straight to global J.
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_deposit_global(Geom g, const PArr p, const double* __restrict__ xo,
                     const double* __restrict__ yo, const double* __restrict__ zo,
                     const int32_t* __restrict__ np, const int* __restrict__ off,
                     const double* __restrict__ lim, double* __restrict__ J, DepConsts dc)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int  c  = seg / g.Ns;
  const int  sp = seg - c * g.Ns;
  const long i  = (long)seg * g.cap + ip;
  double     xv[3] = {xo[i], yo[i], zo[i]};
  double     xu[3] = {p.x[i], p.y[i], p.z[i]};
  const long v     = g.vol;
  esirkepov<Dim, Order>(g, lim[c * 6 + 4], lim[c * 6 + 2], lim[c * 6 + 0], xv, xu, dc.q[sp], dc,
                        [&](int lz, int ly, int lx, int k, double val) {
                          // drop nodes outside the padded grid (particles placed
                          // outside their chunk through pn_set_particles)
                          int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
                          int gy = g.has[1] ? off[c * 3 + 1] + ly : 0;
                          int gx = off[c * 3 + 2] + lx;
                          if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
                              (unsigned)gx >= (unsigned)g.P[2])
                            return;
                          unsafeAtomicAdd(&J[k * v + gnode(g, off, c, lz, ly, lx)], val);
                        });
}

// Slab form (see file header).  Dynamic LDS: 4 x TZ x TY x TX doubles.
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_deposit_slab(Geom g, const PArr p, const double* __restrict__ xo,
                   const double* __restrict__ yo, const double* __restrict__ zo,
                   const int32_t* __restrict__ cellstart, const int* __restrict__ off,
                   const double* __restrict__ lim, double* __restrict__ J, DepConsts dc)
{
  extern __shared__ double tile[];
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int N      = Order + 3;
  const int     slab   = blockIdx.x;
  const int     seg    = blockIdx.y;
  const int     c      = seg / g.Ns;
  const int     sp     = seg - c * g.Ns;
  // tile geometry in chunk-local node indices
  const int TX = g.dims[2] + 2 * g.Nb;
  int       TY, TZ, ylo, zlo, b0, b1;
  if constexpr (Dim == 3) {
    TY  = g.dims[1] + 2 * g.Nb;
    TZ  = N;
    zlo = slab - is_odd + g.Nb - (Order / 2) - 1;
    ylo = 0;
    b0  = slab * g.bsz;
    b1  = (slab + 1) * g.bsz;
  } else {
    TY  = N;
    TZ  = 1;
    ylo = slab - is_odd + g.Nb - (Order / 2) - 1;
    zlo = g.Nb;
    b0  = slab * g.bsy;
    b1  = (slab + 1) * g.bsy;
  }
  const int      tsz = TX * TY * TZ;
  const int32_t* cs  = cellstart + (long)seg * (g.nbin + 1);
  const int      p0  = cs[b0];
  const int      p1  = cs[b1 < g.nbin ? b1 : g.nbin];
  if (p1 <= p0)
    return;
  for (int t = threadIdx.x; t < 4 * tsz; t += blockDim.x)
    tile[t] = 0.0;
  __syncthreads();

  const long   v    = g.vol;
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const double qs   = dc.q[sp];
  const int    L    = p1 - p0;
  const int    per  = (L + blockDim.x - 1) / blockDim.x;
  auto         add  = [&](int lz, int ly, int lx, int k, double val) {
    int tz = lz - zlo, ty = ly - ylo;
    if ((unsigned)tz < (unsigned)TZ && (unsigned)ty < (unsigned)TY && (unsigned)lx < (unsigned)TX)
      atomicAdd(&tile[((k * TZ + tz) * TY + ty) * TX + lx], val);
    else
      unsafeAtomicAdd(&J[k * v + gnode(g, off, c, lz, ly, lx)], val);
  };
  const int q0 = p0 + (int)threadIdx.x * per;
  for (int k = 0; k < per; k++) {
    int q = q0 + k;
    if (q >= p1)
      break;
    long   i     = (long)seg * g.cap + q;
    double xv[3] = {xo[i], yo[i], zo[i]};
    double xu[3] = {p.x[i], p.y[i], p.z[i]};
    esirkepov<Dim, Order>(g, xmin, ymin, zmin, xv, xu, qs, dc, add);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 4 * tsz; t += blockDim.x) {
    double val = tile[t];
    if (val == 0.0)
      continue;
    int k  = t / tsz;
    int r  = t - k * tsz;
    int tz = r / (TY * TX);
    int r2 = r - tz * TY * TX;
    int ty = r2 / TX;
    int tx = r2 - ty * TX;
    unsafeAtomicAdd(&J[k * v + gnode(g, off, c, tz + zlo, ty + ylo, tx)], val);
  }
}

// March kernel (3D current): one block per (bin column by, chunk, range of bin
// rows bz).  The block walks its rows in z order; for each row its wavefronts
// take whole cells, of every species of the chunk, from a block-level counter.
// All particles of a bin share the tile origin (the old cell,
// pic/engine/current.hpp:390-392), so the contributions of a cell are summed
// in registers.  Jx, Jy and Jz have one structure (nix/esirkepov.hpp:168-232):
// a plane weight
//   w = -((1 s0F[f] + A dsF[f]) s0S[s] + (A s0F[f] + B dsF[f]) dsS[s]) q d/dt
// with (F, S) = (y, z), (z, x), (x, y), and the running sum along C = x, y, z
//   J(f, s, c + 1) += sum_{c' <= c} dsC[c'] w  =  P[c] w,
// P being the prefix sum of dsC.  The kernel evaluates P[c] w, one fused
// multiply-add per node, and forms w and its factors with fused multiply-adds
// too; the reference adds the products one at a time.  The two agree to
// rounding (a few ulp of each contribution), the tolerance the current is held
// to (the sum over particles runs in another order anyway).
// A lane owns the plane F = f of one component (N rows s of N-1 nodes) for one
// particle of a slot of PPS particles, so all lanes evaluate the same
// expression on their own operands.  The shape weights are computed by one lane
// per (particle, axis) and handed over in LDS; the coordinates of the next batch
// are loaded while the current one is processed.  After a cell, the PPS partial
// planes are summed across lanes (DPP) and added to the block's LDS tile, a ring
// of N z-planes: when the block moves on to the next row, the lowest z-plane
// holds all of this block's contributions and goes to global J (one atomic per
// non-zero node), and its slot becomes the new top plane.
// rho is deposited after the sort (k_rho_march) or by k_rho_global.
template <int Order>
struct MarchPlan {
  static constexpr int N     = Order + 3;
  static constexpr int NPL   = 3 * N; // planes per particle (3 components x N)
  static constexpr int PPS   = 64 / NPL >= 4 ? 4 : (64 / NPL >= 2 ? 2 : 1);
  static constexpr int BATCH = 16;    // phase 1 uses 3 x BATCH lanes
  static constexpr int NS    = N & ~1;   // prefix sums per axis, padded to pairs (>= N-1)
  // Weights live in LDS as pairs of doubles, [axis][row][particle]: rows 0..N-1 hold
  // (s0[l], ds[l]), rows N.. the prefix sums of ds two by two.  The particle pitch
  // PITCH = BATCH + 4 pairs puts the 4 particles of a slot, and the neighbouring
  // rows of one particle, on distinct LDS banks.
  static constexpr int ROWS  = N + NS / 2;
  static constexpr int PITCH = BATCH + 4;
  static constexpr int WPAIR = 3 * ROWS * PITCH; // pairs per wave
  static constexpr int NWAVE = 4;
  static constexpr size_t WLDS =
      (sizeof(double) * 2 * WPAIR + sizeof(int) * 4 * BATCH) * NWAVE;
  // waves per SIMD the register allocation targets.  The kernel is latency
  // bound; at Order 2 three waves (a few spilled registers) measured 19 %
  // faster than two (no spills) and 10 % faster than four (many spills).
  static constexpr int WAVES = Order <= 2 ? 3 : 1;
};

// Dynamic LDS of the march kernels: the ring tile (doubles), then the bin
// offsets of the current row for every species (ints).
inline size_t march_lds(const Geom& g, int ntile)
{
  return sizeof(double) * (size_t)ntile + sizeof(int) * (size_t)g.Ns * (g.bsy + 1);
}

template <int Order, bool Queue>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(MarchPlan<Order>::WAVES, MarchPlan<Order>::WAVES)))
    k_deposit_march(Geom g, const PArr p, const double* __restrict__ xo,
                    const double* __restrict__ yo, const double* __restrict__ zo,
                    const int32_t* __restrict__ cellstart, const int* __restrict__ off,
                    const double* __restrict__ lim, double* __restrict__ J, DepConsts dc,
                    int zrows, int* __restrict__ wq, int gx, int nitems)
{
  using MP             = MarchPlan<Order>;
  constexpr int N      = MP::N;
  constexpr int NPL    = MP::NPL;
  constexpr int PPS    = MP::PPS;
  constexpr int BATCH  = MP::BATCH;
  constexpr int NS     = MP::NS;
  constexpr int ROWS   = MP::ROWS;
  constexpr int PITCH  = MP::PITCH;
  constexpr int is_odd = Order % 2;
  constexpr int NWAVE  = MP::NWAVE;

  __shared__ double2 wts[NWAVE][MP::WPAIR];
  __shared__ int4    forg[NWAVE][BATCH]; // (same origin as the bin, own origin x, y, z)
  __shared__ int     next_cell;
  extern __shared__ double tile[]; // [3][N z-ring][N][TX] (Jx, Jy, Jz), then rcs

  // Queue: persistent blocks take (row group, chunk) items from the work queue
  // wq until it is drained; otherwise the item is the block's own index.
  __shared__ int wq_item;
  for (;;) {
  int item_x = blockIdx.x, item_y = blockIdx.y;
  if constexpr (Queue) {
    const int it = wq_next(wq, &wq_item);
    if (it >= nitems)
      break; // uniform: every thread read the same item
    item_y = it / gx;
    item_x = it - item_y * gx;
  }
  do {
  const int c   = item_y;
  const int ns  = g.Ns;
  const int nbx = g.bsy;
  const int nby = g.bsz / g.bsy;
  const int nbz = g.nbin / g.bsz;
  const int by  = item_x % nby;
  const int bz0 = (item_x / nby) * zrows;
  const int bz1 = min(nbz, bz0 + zrows);

  // nothing to do when every row of the range is empty for every species
  int any = 0;
  for (int q = threadIdx.x; q < ns * (bz1 - bz0); q += blockDim.x) {
    const int      sp = q % ns;
    const int      bz = bz0 + q / ns;
    const int32_t* cs = cellstart + (long)(c * ns + sp) * (g.nbin + 1);
    const int      b0 = bz * g.bsz + by * g.bsy;
    any |= cs[b0 + nbx] > cs[b0] ? 1 : 0;
  }
  if (!__syncthreads_or(any))
    break; // uniform for the block

  const int lb  = g.Nb;
  const int ylo = by - is_odd + lb - (Order / 2) - 1; // tile origin (local node index)
  const int xlo = -is_odd + lb - (Order / 2) - 1;
  const int TX  = nbx + N - 1;
  const int psz = N * TX; // one z-plane of one component
  const int tsz = N * psz;
  int*      rcs = reinterpret_cast<int*>(tile + 3 * tsz); // [ns][nbx + 1]
  for (int q = threadIdx.x; q < 3 * tsz; q += blockDim.x)
    tile[q] = 0.0;
  auto load_row = [&](int bz) {
    for (int q = threadIdx.x; q < ns * (nbx + 1); q += blockDim.x) {
      const int sp = q / (nbx + 1);
      const int i  = q - sp * (nbx + 1);
      rcs[q] = cellstart[(long)(c * ns + sp) * (g.nbin + 1) + bz * g.bsz + by * g.bsy + i];
    }
  };
  load_row(bz0);
  if (threadIdx.x == 0)
    next_cell = 0;
  __syncthreads();

  const int    lane = threadIdx.x & 63;
  const int    wv   = threadIdx.x >> 6;
  const long   v    = g.vol;
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const double A    = 1.0 / 2;
  const double B    = 1.0 / 3;

  // phase-2 lane -> (plane, slot particle k); plane -> (component, index t along F)
  const bool   act   = lane < NPL * PPS;
  const int    plane = act ? lane / PPS : 0;
  const int    k     = lane & (PPS - 1);
  const int    comp  = plane / N; // 0 Jx, 1 Jy, 2 Jz
  const int    t     = plane - comp * N;
  const int    axF   = comp == 0 ? 1 : (comp == 1 ? 2 : 0); // axes: 0 x, 1 y, 2 z
  const int    axS   = comp == 0 ? 2 : (comp == 1 ? 0 : 1);
  const int    axC   = comp;
  const double ddt   = comp == 0 ? dc.dxdt : (comp == 1 ? dc.dydt : dc.dzdt);
  // node offsets (z, y, x) of plane t, row r, prefix j: F = t, S = r, C = j + 1
  // phase-1 lane -> (axis, batch particle)
  const int     a1    = lane / BATCH; // 3: idle
  const int     k1    = lane - a1 * BATCH;
  const double* X0    = a1 == 0 ? xo : (a1 == 1 ? yo : zo);
  const double* X1    = a1 == 0 ? p.x : (a1 == 1 ? p.y : p.z);
  const double  d1    = a1 == 0 ? g.del[2] : (a1 == 1 ? g.del[1] : g.del[0]);
  const double  mn1   = a1 == 0 ? xmin : (a1 == 1 ? ymin : zmin);
  const double  rd1   = 1 / d1;
  const double  imin1 = mn1 + 0.5 * d1 * is_odd;
  const double  grid1 = mn1 + 0.5 * d1;
  double2*      W     = &wts[wv][0];
  int4*         FO    = &forg[wv][0];

  // claim the next cell of the row that holds particles of any species
  // (uniform per wave); false when the row is done
  auto claim = [&](int& bx) {
    for (;;) {
      int q = 0;
      if (lane == 0)
        q = atomicAdd(&next_cell, 1);
      q = __builtin_amdgcn_readfirstlane(q);
      if (q >= nbx)
        return false;
      bx = q;
      for (int sp = 0; sp < ns; sp++)
        if (rcs[sp * (nbx + 1) + bx + 1] > rcs[sp * (nbx + 1) + bx])
          return true;
    }
  };
  // the first species >= sp of cell bx that holds particles, and its range
  auto species_from = [&](int bx, int& sp, int& p0, int& p1) {
    for (; sp < ns; sp++) {
      p0 = rcs[sp * (nbx + 1) + bx];
      p1 = rcs[sp * (nbx + 1) + bx + 1];
      if (p1 > p0)
        return true;
    }
    return false;
  };
  // coordinates of batch particle k1 (axis a1) of [pb, p1) in segment seg
  auto fetch = [&](int seg, int pb, int p1, double& xv, double& xu) {
    if (a1 < 3 && pb + k1 < p1) {
      const long i = (long)seg * g.cap + pb + k1;
      xv           = X0[i];
      xu           = X1[i];
    }
  };

  for (int bz = bz0; bz < bz1; bz++) {
    const int zlo = bz - is_odd + lb - (Order / 2) - 1;
    const int zs  = ((zlo % N) + N) % N; // ring slot of z-plane zlo
    // One claimed cell at a time; every species of the cell accumulates into
    // the same registers (the tile origin is the cell's), flushed once per cell.
    int    bx, sp = 0, p0 = 0, p1 = 0;
    bool   have = claim(bx);
    double nxv = 0, nxu = 0; // prefetched coordinates
    if (have) {
      species_from(bx, sp, p0, p1);
      fetch(c * ns + sp, p0, p1, nxv, nxu);
    }
    while (have) {
      int        bxn, spn = 0, p0n = 0, p1n = 0;
      const bool more = claim(bxn);
      if (more)
        species_from(bxn, spn, p0n, p1n);
      const int ox = bx + xlo;

      double acc[N][N - 1];
#pragma unroll
      for (int r = 0; r < N; r++)
#pragma unroll
        for (int j = 0; j < N - 1; j++)
          acc[r][j] = 0.0;

      for (;;) { // species of cell bx
      const int    seg = c * ns + sp;
      const double qd  = dc.q[sp] * ddt; // qs * dxdt etc. (esirkepov.hpp:172-174)
      int          sp2 = sp + 1, q0 = 0, q1 = 0;
      const bool   next_sp = species_from(bx, sp2, q0, q1);
      for (int pb = p0; pb < p1; pb += BATCH) {
        const int    nbt = min(BATCH, p1 - pb);
        const double xv  = nxv, xu = nxu;
        if (pb + BATCH < p1)
          fetch(seg, pb + BATCH, p1, nxv, nxu);
        else if (next_sp)
          fetch(c * ns + sp2, q0, q1, nxv, nxu);
        else if (more)
          fetch(c * ns + spn, p0n, p1n, nxv, nxu);
        // ---- phase 1: shape weights, one lane per (particle, axis) ----
        const bool on1  = a1 < 3 && k1 < nbt;
        bool       same = false;
        int        org  = 0;
        if (on1) {
          double s0[N], s1[N];
#pragma unroll
          for (int l = 0; l < N; l++)
            s0[l] = s1[l] = 0.0;
          const int i0 = digitize(xv, imin1, rd1);
          const int i1 = digitize(xu, imin1, rd1);
          shape_mc<Order>(xv, grid1 + static_cast<double>(i0) * d1, rd1, &s0[1]);
          shape_mc<Order>(xu, grid1 + static_cast<double>(i1) * d1, rd1, &s1[1]);
          // shift_weights (nix/esirkepov.hpp:239-270)
          if (i1 - i0 < 0) {
#pragma unroll
            for (int ii = 0; ii < Order + 2; ii++)
              s1[ii] = s1[ii + 1];
          } else if (i1 - i0 > 0) {
#pragma unroll
            for (int ii = Order + 2; ii > 0; ii--)
              s1[ii] = s1[ii - 1];
          }
          org            = i0 + lb - (Order / 2) - 1;
          const int want = a1 == 0 ? ox : (a1 == 1 ? ylo : zlo);
          same           = org == want;
          double2* pr    = W + a1 * ROWS * PITCH + k1;
          double   ds[N];
#pragma unroll
          for (int l = 0; l < N; l++) {
            ds[l]           = s1[l] - s0[l]; // DS (esirkepov.hpp:157-166)
            pr[l * PITCH] = make_double2(s0[l], ds[l]);
          }
          double pre[NS];
          pre[0] = ds[0];
#pragma unroll
          for (int l = 1; l < NS; l++)
            pre[l] = l < N - 1 ? pre[l - 1] + ds[l] : 0.0;
#pragma unroll
          for (int l = 0; l < NS / 2; l++)
            pr[(N + l) * PITCH] = make_double2(pre[2 * l], pre[2 * l + 1]);
        }
        // Odd orders only: per particle, are all three axes at the bin's origin, and
        // its own origin.  For even orders the bin (the sort's digitize of the same
        // position with the same offsets) is always the stencil's origin.
        if constexpr (is_odd) {
          const unsigned long long m    = __ballot(on1 && same);
          const int                orgy = __shfl(org, (lane + BATCH) & 63);
          const int                orgz = __shfl(org, (lane + 2 * BATCH) & 63);
          if (lane < nbt) {
            const bool all3 =
                ((m >> lane) & (m >> (lane + BATCH)) & (m >> (lane + 2 * BATCH)) & 1ull) != 0;
            FO[lane] = make_int4(all3 ? 1 : 0, org, orgy, orgz);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // ---- phase 2: each lane accumulates its plane ----
        for (int k0 = 0; k0 < nbt; k0 += PPS) {
          const int kk = k0 + k;
          if (act && kk < nbt) {
            const double2* pp  = W + kk;
            const double2  pf  = pp[(axF * ROWS + t) * PITCH]; // (s0F[t], dsF[t])
            double2        psr[N]; // (s0S[r], dsS[r]), every row read before use
#pragma unroll
            for (int r = 0; r < N; r++)
              psr[r] = pp[(axS * ROWS + r) * PITCH];
            const double   aF  = fma(A, pf.y, pf.x);                 // 1 s0F + A dsF
            const double   bF  = fma(B, pf.y, A * pf.x);             // A s0F + B dsF
            double         P[NS];
#pragma unroll
            for (int j = 0; j < NS / 2; j++) {
              const double2 d2 = pp[(axC * ROWS + N + j) * PITCH];
              P[2 * j]         = d2.x;
              P[2 * j + 1]     = d2.y;
            }
            const int4 fo = is_odd ? FO[kk] : make_int4(1, 0, 0, 0);
            if (fo.x != 0) {
#pragma unroll
              for (int r = 0; r < N; r++) {
                const double2 ps = psr[r];
                const double  w  = -fma(aF, ps.x, bF * ps.y) * qd;
#pragma unroll
                for (int j = 0; j < N - 1; j++)
                  acc[r][j] = fma(P[j], w, acc[r][j]);
              }
            } else {
              // the particle's tile origin differs from its bin's (odd orders: the
              // bin and shape digitize offsets differ by half a cell, and the two
              // roundings can disagree): same values, added at its own nodes
              const int pox = fo.y;
              const int poy = fo.z;
              const int poz = fo.w;
#pragma unroll 1
              for (int r = 0; r < N; r++) {
                const double2 ps = pp[(axS * ROWS + r) * PITCH];
                const double  w  = -fma(aF, ps.x, bF * ps.y) * qd;
#pragma unroll 1
                for (int j = 0; j < N - 1; j++) {
                  const double val = P[j] * w;
                  const int    oz  = comp == 0 ? r : (comp == 1 ? t : j + 1);
                  const int    oy  = comp == 0 ? t : (comp == 1 ? j + 1 : r);
                  const int    oxx = comp == 0 ? j + 1 : (comp == 1 ? r : t);
                  const int    lz = poz + oz, ly = poy + oy, lx = pox + oxx;
                  const int    tz = lz - zlo, ty = ly - ylo, tx = lx - xlo;
                  if ((unsigned)tz < (unsigned)N && (unsigned)ty < (unsigned)N &&
                      (unsigned)tx < (unsigned)TX) {
                    const int sl = zs + tz >= N ? zs + tz - N : zs + tz;
                    atomicAdd(&tile[comp * tsz + sl * psz + ty * TX + tx], val);
                  } else if (long gi = gnode_checked(g, off, c, lz, ly, lx); gi >= 0) {
                    unsafeAtomicAdd(&J[(comp + 1) * v + gi], val);
                  }
                }
              }
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (!next_sp)
        break;
      sp = sp2;
      p0 = q0;
      p1 = q1;
      } // species
      // ---- sum the PPS partial planes, add them to the ring tile ----
#pragma unroll
      for (int r = 0; r < N; r++)
#pragma unroll
        for (int j = 0; j < N - 1; j++) {
          double s = acc[r][j];
          if (PPS >= 2)
            s += quad_xor<1>(s);
          if (PPS >= 4)
            s += quad_xor<2>(s);
          acc[r][j] = s;
        }
      if (act && k == 0) {
        int ro[N], co[N - 1];
        plane_offsets<N>(comp, t, zs, TX, psz, tsz, bx, ro, co);
#pragma unroll
        for (int r = 0; r < N; r++)
#pragma unroll
          for (int j = 0; j < N - 1; j++)
            atomicAdd(&tile[ro[r] + co[j]], acc[r][j]);
      }
      have = more;
      sp   = spn;
      bx   = bxn;
      p0   = p0n;
      p1   = p1n;
    }
    __syncthreads();
    // z-plane zlo is complete for this block: flush and clear its ring slot
    for (int q = threadIdx.x; q < 3 * psz; q += blockDim.x) {
      const int    kc  = q / psz;
      const int    r2  = q - kc * psz;
      const int    idx = kc * tsz + zs * psz + r2;
      const double val = tile[idx];
      if (val != 0.0) {
        const int  ty = r2 / TX;
        const int  tx = r2 - ty * TX;
        const long gi = gnode_checked(g, off, c, zlo, ylo + ty, xlo + tx);
        if (gi >= 0)
          unsafeAtomicAdd(&J[(kc + 1) * v + gi], val);
        tile[idx] = 0.0;
      }
    }
    if (bz + 1 < bz1)
      load_row(bz + 1);
    if (threadIdx.x == 0)
      next_cell = 0;
    __syncthreads();
  }
  // the N-1 planes above the last row's zlo
  const int zl = (bz1 - 1) - is_odd + lb - (Order / 2) - 1;
  const int zs = ((zl % N) + N) % N;
  for (int q = threadIdx.x; q < 3 * (N - 1) * psz; q += blockDim.x) {
    const int    kc  = q / ((N - 1) * psz);
    const int    r   = q - kc * (N - 1) * psz;
    const int    dz  = 1 + r / psz;
    const int    r2  = r - (dz - 1) * psz;
    const int    sl  = zs + dz >= N ? zs + dz - N : zs + dz;
    const double val = tile[kc * tsz + sl * psz + r2];
    if (val == 0.0)
      continue;
    const int  ty = r2 / TX;
    const int  tx = r2 - ty * TX;
    const long gi = gnode_checked(g, off, c, zl + dz, ylo + ty, xlo + tx);
    if (gi >= 0)
      unsafeAtomicAdd(&J[(kc + 1) * v + gi], val);
  }
  } while (false);
  if constexpr (!Queue)
    return;
  __syncthreads(); // the LDS of this item is read out before the next one
  }
  if constexpr (Queue)
    wq_finish(wq);
}

// rho from the new positions (3D), after the sort: the reference's rho term
// qs S1x S1y S1z (ro3d, nix/esirkepov.hpp:145-155) uses the new-position
// weights, which -- undoing the shift -- are the Order+1 unshifted weights around
// the new cell.  Same block walk as k_deposit_march, with a ring of S = Order+1
// z-planes; lanes own (z plane t, slot particle k), PPS particles per slot.
template <int Order>
struct RhoPlan {
  static constexpr int S     = Order + 1;
  static constexpr int PPS   = 64 / S >= 32 ? 32 : (64 / S >= 16 ? 16 : 8);
  static constexpr int BATCH = 64;
  static constexpr int FLAG  = 3 * S;     // 1: origin of the bin, 0: own origin
  static constexpr int ORG   = 3 * S + 1; // own origin (x, y, z), local nodes
  static constexpr int REC   = (3 * S + 5) & ~1;
  static constexpr int NWAVE = 4;
  static constexpr size_t WLDS = sizeof(double) * NWAVE * BATCH * REC;
};

template <int Order>
__global__ void __launch_bounds__(256)
    k_rho_march(Geom g, const PArr p, const int32_t* __restrict__ cellstart,
                const int* __restrict__ off, const double* __restrict__ lim,
                double* __restrict__ J, int zrows, const int32_t* __restrict__ perm)
{
  using RP             = RhoPlan<Order>;
  constexpr int S      = RP::S;
  constexpr int PPS    = RP::PPS;
  constexpr int BATCH  = RP::BATCH;
  constexpr int REC    = RP::REC;
  constexpr int is_odd = Order % 2;
  constexpr int NWAVE  = RP::NWAVE;

  __shared__ double wts[NWAVE][BATCH][REC];
  __shared__ int    next_cell;
  extern __shared__ double tile[]; // [S z-ring][S][TX], then rcs

  const int c   = blockIdx.y;
  const int ns  = g.Ns;
  const int nbx = g.bsy;
  const int nby = g.bsz / g.bsy;
  const int nbz = g.nbin / g.bsz;
  const int by  = blockIdx.x % nby;
  const int bz0 = (blockIdx.x / nby) * zrows;
  const int bz1 = min(nbz, bz0 + zrows);

  int any = 0;
  for (int q = threadIdx.x; q < ns * (bz1 - bz0); q += blockDim.x) {
    const int      sp = q % ns;
    const int      bz = bz0 + q / ns;
    const int32_t* cs = cellstart + (long)(c * ns + sp) * (g.nbin + 1);
    const int      b0 = bz * g.bsz + by * g.bsy;
    any |= cs[b0 + nbx] > cs[b0] ? 1 : 0;
  }
  if (!__syncthreads_or(any))
    return;

  const int lb  = g.Nb;
  const int ylo = by - is_odd + lb - (Order / 2);
  const int xlo = -is_odd + lb - (Order / 2);
  const int TX  = nbx + Order;
  const int psz = S * TX;
  const int tsz = S * psz;
  int*      rcs = reinterpret_cast<int*>(tile + tsz); // [ns][nbx + 1]
  for (int q = threadIdx.x; q < tsz; q += blockDim.x)
    tile[q] = 0.0;
  auto load_row = [&](int bz) {
    for (int q = threadIdx.x; q < ns * (nbx + 1); q += blockDim.x) {
      const int sp = q / (nbx + 1);
      const int i  = q - sp * (nbx + 1);
      rcs[q] = cellstart[(long)(c * ns + sp) * (g.nbin + 1) + bz * g.bsz + by * g.bsy + i];
    }
  };
  load_row(bz0);
  if (threadIdx.x == 0)
    next_cell = 0;
  __syncthreads();

  const int    lane = threadIdx.x & 63;
  const int    wv   = threadIdx.x >> 6;
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const bool   act  = lane < S * PPS;
  const int    t    = act ? lane / PPS : 0; // z plane
  const int    k    = lane & (PPS - 1);
  double*      W    = &wts[wv][0][0];

  auto claim = [&](int& sp, int& bx, int& p0, int& p1) {
    for (;;) {
      int q = 0;
      if (lane == 0)
        q = atomicAdd(&next_cell, 1);
      q = __builtin_amdgcn_readfirstlane(q);
      if (q >= ns * nbx)
        return false;
      sp = q / nbx;
      bx = q - sp * nbx;
      p0 = rcs[sp * (nbx + 1) + bx];
      p1 = rcs[sp * (nbx + 1) + bx + 1];
      if (p1 > p0)
        return true;
    }
  };
  auto fetch = [&](int seg, int pb, int p1, double* x3) {
    if (pb + lane < p1) {
      const long i   = (long)seg * g.cap + pb + lane;
      const long src = perm ? (long)perm[i] : i;
      x3[0]          = p.x[src];
      x3[1]          = p.y[src];
      x3[2]          = p.z[src];
    }
  };

  for (int bz = bz0; bz < bz1; bz++) {
    const int zlo = bz - is_odd + lb - (Order / 2);
    const int zs  = ((zlo % S) + S) % S;
    int       sp, bx, p0, p1;
    bool      have   = claim(sp, bx, p0, p1);
    double    nx3[3] = {0, 0, 0};
    if (have)
      fetch(c * ns + sp, p0, p1, nx3);
    while (have) {
      int          spn, bxn, p0n, p1n;
      const bool   more = claim(spn, bxn, p0n, p1n);
      const int    seg  = c * ns + sp;
      const double qs   = g.q[sp];
      double       acc[S][S];
#pragma unroll
      for (int a = 0; a < S; a++)
#pragma unroll
        for (int e = 0; e < S; e++)
          acc[a][e] = 0.0;
      for (int pb = p0; pb < p1; pb += BATCH) {
        const int    nbt   = min(BATCH, p1 - pb);
        const double xu[3] = {nx3[0], nx3[1], nx3[2]};
        if (pb + BATCH < p1)
          fetch(seg, pb + BATCH, p1, nx3);
        else if (more)
          fetch(c * ns + spn, p0n, p1n, nx3);
        if (lane < nbt) {
          const double d3[3]   = {g.del[2], g.del[1], g.del[0]};
          const double mn[3]   = {xmin, ymin, zmin};
          const int    want[3] = {bx + xlo, ylo, zlo};
          double*      rec     = W + lane * REC;
          bool         same    = true;
#pragma unroll
          for (int a = 0; a < 3; a++) {
            const double rd    = 1 / d3[a];
            const double imin  = mn[a] + 0.5 * d3[a] * is_odd;
            const double grid0 = mn[a] + 0.5 * d3[a];
            const int    cell  = digitize(xu[a], imin, rd);
            shape_mc<Order>(xu[a], grid0 + static_cast<double>(cell) * d3[a], rd, &rec[a * S]);
            const int org    = cell + lb - (Order / 2);
            same             = same && (org == want[a]);
            rec[RP::ORG + a] = static_cast<double>(org);
          }
          rec[RP::FLAG] = same ? 1.0 : 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int k0 = 0; k0 < nbt; k0 += PPS) {
          const int kk = k0 + k;
          if (act && kk < nbt) {
            const double* rec = W + kk * REC;
            const double  sz  = rec[2 * S + t];
            double        sx[S];
#pragma unroll
            for (int jx = 0; jx < S; jx++)
              sx[jx] = rec[jx];
            // even orders: the sort's bin is always the stencil's origin
            if (!is_odd || rec[RP::FLAG] != 0.0) {
#pragma unroll
              for (int jy = 0; jy < S; jy++) {
                const double sy = rec[S + jy];
#pragma unroll
                for (int jx = 0; jx < S; jx++)
                  acc[jy][jx] = fma(qs * sx[jx] * sy, sz, acc[jy][jx]);
              }
            } else {
              // own origin (odd orders, as in k_deposit_march)
              const int pox = (int)rec[RP::ORG + 0];
              const int poy = (int)rec[RP::ORG + 1];
              const int poz = (int)rec[RP::ORG + 2];
              for (int jy = 0; jy < S; jy++) {
                const double sy = rec[S + jy];
                for (int jx = 0; jx < S; jx++) {
                  const double val = qs * sx[jx] * sy * sz;
                  const int    lz = poz + t, ly = poy + jy, lx = pox + jx;
                  const int    tz = lz - zlo, ty = ly - ylo, tx = lx - xlo;
                  if ((unsigned)tz < (unsigned)S && (unsigned)ty < (unsigned)S &&
                      (unsigned)tx < (unsigned)TX) {
                    const int sl = zs + tz >= S ? zs + tz - S : zs + tz;
                    atomicAdd(&tile[sl * psz + ty * TX + tx], val);
                  } else if (long gi = gnode_checked(g, off, c, lz, ly, lx); gi >= 0) {
                    unsafeAtomicAdd(&J[gi], val);
                  }
                }
              }
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
#pragma unroll
      for (int a = 0; a < S; a++)
#pragma unroll
        for (int e = 0; e < S; e++) {
          double x = acc[a][e];
          x += quad_xor<1>(x);
          x += quad_xor<2>(x);
#pragma unroll
          for (int m = 4; m < PPS; m <<= 1)
            x += __shfl_xor(x, m);
          acc[a][e] = x;
        }
      if (act && k == 0) {
        const int sl = zs + t >= S ? zs + t - S : zs + t;
#pragma unroll
        for (int jy = 0; jy < S; jy++)
#pragma unroll
          for (int jx = 0; jx < S; jx++)
            atomicAdd(&tile[sl * psz + jy * TX + bx + jx], acc[jy][jx]);
      }
      have = more;
      sp   = spn;
      bx   = bxn;
      p0   = p0n;
      p1   = p1n;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < psz; q += blockDim.x) {
      const int    idx = zs * psz + q;
      const double val = tile[idx];
      if (val != 0.0) {
        const int  ty = q / TX;
        const int  tx = q - ty * TX;
        const long gi = gnode_checked(g, off, c, zlo, ylo + ty, xlo + tx);
        if (gi >= 0)
          unsafeAtomicAdd(&J[gi], val);
        tile[idx] = 0.0;
      }
    }
    if (bz + 1 < bz1)
      load_row(bz + 1);
    if (threadIdx.x == 0)
      next_cell = 0;
    __syncthreads();
  }
  const int zl = (bz1 - 1) - is_odd + lb - (Order / 2);
  const int zs = ((zl % S) + S) % S;
  for (int q = threadIdx.x; q < (S - 1) * psz; q += blockDim.x) {
    const int    dz  = 1 + q / psz;
    const int    r2  = q - (dz - 1) * psz;
    const int    sl  = zs + dz >= S ? zs + dz - S : zs + dz;
    const double val = tile[sl * psz + r2];
    if (val == 0.0)
      continue;
    const int  ty = r2 / TX;
    const int  tx = r2 - ty * TX;
    const long gi = gnode_checked(g, off, c, zl + dz, ylo + ty, xlo + tx);
    if (gi >= 0)
      unsafeAtomicAdd(&J[gi], val);
  }
}

// rho from the new positions, one thread per particle (any order; 2D or 3D).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_rho_global(Geom g, const PArr p, const int32_t* __restrict__ np,
                 const int* __restrict__ off, const double* __restrict__ lim,
                 double* __restrict__ J, const int32_t* __restrict__ perm)
{
  constexpr int is_odd = Order % 2;
  constexpr int S      = Order + 1;
  const int     seg    = blockIdx.y;
  const long    ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int    c     = seg / g.Ns;
  const int    sp    = seg - c * g.Ns;
  const long   i     = (long)seg * g.cap + ip;
  const long   src   = perm ? (long)perm[i] : i;
  const double xu[3] = {p.x[src], p.y[src], p.z[src]};
  const double d3[3] = {g.del[2], g.del[1], g.del[0]};
  const double mn[3] = {lim[c * 6 + 4], lim[c * 6 + 2], lim[c * 6 + 0]};
  double       s[3][S];
  int          o[3];
  for (int a = 0; a < Dim; a++) {
    const double rd = 1 / d3[a];
    int cell = digitize(xu[a], mn[a] + 0.5 * d3[a] * is_odd, rd);
    shape_mc<Order>(xu[a], mn[a] + 0.5 * d3[a] + static_cast<double>(cell) * d3[a], rd, s[a]);
    o[a] = cell + g.Nb - (Order / 2);
  }
  const double qs = g.q[sp];
  for (int jz = 0; jz < (Dim == 3 ? S : 1); jz++)
    for (int jy = 0; jy < (Dim >= 2 ? S : 1); jy++)
      for (int jx = 0; jx < S; jx++) {
        // ro1d / ro2d / ro3d: the product of the present axes' new weights
        double val = Dim == 3 ? qs * s[0][jx] * s[1][jy] * s[2][jz]
                              : (Dim == 2 ? qs * s[0][jx] * s[1][jy] : qs * s[0][jx]);
        const int  lz = Dim == 3 ? o[2] + jz : g.Nb;
        const int  ly = Dim >= 2 ? o[1] + jy : g.Nb;
        const long gi = gnode_checked(g, off, c, lz, ly, o[0] + jx);
        if (gi >= 0)
          unsafeAtomicAdd(&J[gi], val);
      }
}

// Remove from rho the particles appended by the injection of this step (the
// reference deposits rho before it injects, example/shock/main.cpp:445-535 runs
// from set_boundary_pack(BoundaryParticle)): minus k_rho_global's deposit of
// every injected particle, read from the injection staging buffer (7 doubles per
// particle) of every upper segment k.
template <int Dim, int Order, typename TJ>
__global__ void __launch_bounds__(256)
    k_rho_unject(Geom g, const int32_t* __restrict__ useg, const int32_t* __restrict__ ninj,
                 const int64_t* __restrict__ voff, const double* __restrict__ vals,
                 const int* __restrict__ off, const double* __restrict__ lim,
                 TJ* __restrict__ J)
{
  constexpr int is_odd = Order % 2;
  constexpr int S      = Order + 1;
  const int     k      = blockIdx.y;
  const long    t      = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ninj[k])
    return;
  const int     seg   = useg[k];
  const int     c     = seg / g.Ns;
  const int     sp    = seg - c * g.Ns;
  const double* v     = vals + voff[k] + 7 * t;
  const double  xu[3] = {v[0], v[1], v[2]};
  const double  d3[3] = {g.del[2], g.del[1], g.del[0]};
  const double  mn[3] = {lim[c * 6 + 4], lim[c * 6 + 2], lim[c * 6 + 0]};
  double        s[3][S];
  int           o[3];
  for (int a = 0; a < Dim; a++) {
    const double rd = 1 / d3[a];
    int cell = digitize(xu[a], mn[a] + 0.5 * d3[a] * is_odd, rd);
    shape_mc<Order>(xu[a], mn[a] + 0.5 * d3[a] + static_cast<double>(cell) * d3[a], rd, s[a]);
    o[a] = cell + g.Nb - (Order / 2);
  }
  const double qs = g.q[sp];
  for (int jz = 0; jz < (Dim == 3 ? S : 1); jz++)
    for (int jy = 0; jy < (Dim >= 2 ? S : 1); jy++)
      for (int jx = 0; jx < S; jx++) {
        double val = Dim == 3 ? qs * s[0][jx] * s[1][jy] * s[2][jz]
                              : (Dim == 2 ? qs * s[0][jx] * s[1][jy] : qs * s[0][jx]);
        const int  lz = Dim == 3 ? o[2] + jz : g.Nb;
        const int  ly = Dim >= 2 ? o[1] + jy : g.Nb;
        const long gi = gnode_checked(g, off, c, lz, ly, o[0] + jx);
        if (gi >= 0)
          unsafeAtomicAdd(&J[gi], static_cast<TJ>(-val));
      }
}

// Rows of bins per block of the march kernels: whole bin columns unless the
// grid would then hold fewer than ~4096 blocks (16 per CU), in which case each
// column is cut into nzr ranges of zrows rows (at 64^3 in 16^3 chunks: 4 ranges
// of 5 rows; measured 2.32 -> 2.26 ms FP64 and 1.44 -> 1.28 ms FP32 against 2
// ranges of 9 rows).  `fixed` > 0 sets zrows.
} // namespace

void march_split(const Geom& g, int fixed, int& zrows, int& nzr)
{
  const int  nby  = g.bsz / g.bsy;
  const int  nbz  = g.nbin / g.bsz;
  const long cols = (long)nby * g.nchunk;
  const int  split = (int)std::min<long>(nbz, std::max<long>(1, (4096 + cols - 1) / cols));
  zrows            = fixed > 0 ? std::min(fixed, nbz) : (nbz + split - 1) / split;
  nzr              = (nbz + zrows - 1) / zrows;
}

namespace {

// The 3D march kernel runs (sorted particles, its LDS fits); it deposits J only.
template <int Order>
bool march_fits(const Sim& s)
{
  const Geom& g = s.g;
  const int   N = Order + 3;
  return g.dim == 3 && s.sorted && !s.generic_kernels &&
         march_lds(g, 3 * N * N * (g.bsy + N - 1)) + MarchPlan<Order>::WLDS + 64 <= 65536;
}

bool deposit_marches(const Sim& s)
{
  switch (s.g.order) {
  case 1: return march_fits<1>(s);
  case 2: return march_fits<2>(s);
  case 3: return march_fits<3>(s);
  default: return march_fits<4>(s);
  }
}

// Returns true when rho still has to be deposited (3D march path).
template <int Dim, int Order>
bool deposit_impl(Sim& s, const DepConsts& dc)
{
  const Geom& g = s.g;
  const int   N = Order + 3;
  const int   TX = g.dims[2] + 2 * g.Nb;
  size_t      lds;
  int         nslab;
  if (Dim == 3) {
    lds   = sizeof(double) * 4 * (size_t)N * (g.dims[1] + 2 * g.Nb) * TX;
    nslab = g.dims[0] + 1;
  } else {
    lds   = sizeof(double) * 4 * (size_t)N * TX;
    nslab = g.dims[1] + 1;
  }
  if (Dim == 3 && march_fits<Order>(s)) {
    const size_t mlds = march_lds(g, 3 * N * N * (g.bsy + N - 1));
    int zrows, nzr;
    march_split(g, s.dep_rows ? s.dep_rows : s.march_rows, zrows, nzr);
    const int gx     = nzr * (g.bsz / g.bsy);
    const int nitems = gx * g.nchunk;
    int*      wq     = s.d_wq + 2 * kWqDeposit;
    if (s.dep_queue) {
      const int nb = queue_blocks(s, reinterpret_cast<const void*>(&k_deposit_march<Order, true>),
                                  mlds, nitems);
      hipLaunchKernelGGL((k_deposit_march<Order, true>), dim3(nb), dim3(256), mlds, s.stream, g,
                         s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.ss.cellstart, s.d_off,
                         s.d_lim, s.d_J, dc, zrows, wq, gx, nitems);
    } else {
      hipLaunchKernelGGL((k_deposit_march<Order, false>), dim3(gx, g.nchunk), dim3(256), mlds,
                         s.stream, g, s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.ss.cellstart,
                         s.d_off, s.d_lim, s.d_J, dc, zrows, wq, gx, nitems);
    }
    return true;
  }
  if (Dim >= 2 && s.sorted && lds <= 65536 && !s.generic_kernels) {
    hipLaunchKernelGGL((k_deposit_slab<Dim, Order>), dim3(nslab, g.nseg), dim3(256), lds,
                       s.stream, g, s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.ss.cellstart,
                       s.d_off, s.d_lim, s.d_J, dc);
  } else {
    hipLaunchKernelGGL((k_deposit_global<Dim, Order>), dim3(nblk(g.cap, 256), g.nseg), dim3(256),
                       0, s.stream, g, s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.d_np, s.d_off,
                       s.d_lim, s.d_J, dc);
  }
  return false;
}

// A sort left pending (perm) is read through: sorted slot i holds record perm[i].
template <int Dim, int Order>
void rho_impl(Sim& s, bool sorted_by_new_cell)
{
  const Geom&    g    = s.g;
  const int32_t* perm = s.perm_pending ? s.ss.perm : nullptr;
  const int   S = Order + 1;
  const size_t mlds = march_lds(g, S * S * (g.bsy + Order));
  if (Dim == 3 && sorted_by_new_cell && !s.generic_kernels &&
      mlds + RhoPlan<Order>::WLDS + 64 <= 65536) {
    int zrows, nzr;
    march_split(g, s.rho_rows ? s.rho_rows : s.march_rows, zrows, nzr);
    hipLaunchKernelGGL((k_rho_march<Order>), dim3(nzr * (g.bsz / g.bsy), g.nchunk), dim3(256),
                       mlds, s.stream, g, s.pa, s.ss.cellstart, s.d_off, s.d_lim, s.d_J, zrows,
                       perm);
  } else {
    hipLaunchKernelGGL((k_rho_global<Dim, Order>), dim3(nblk(g.cap, 256), g.nseg), dim3(256), 0,
                       s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_J, perm);
  }
}

} // namespace

RhoDone launch_deposit(Sim& s, double delt, RhoMode mode)
{
  materialize(s);
  const Geom& g = s.g;
  DepConsts   dc{};
  for (int sp = 0; sp < g.Ns; sp++)
    dc.q[sp] = g.q[sp];
  dc.delt = delt;
  dc.dxdt = g.del[2] / delt;
  dc.dydt = g.del[1] / delt;
  dc.dzdt = g.del[0] / delt;
  // the 3D march kernel deposits J only; the others deposit rho with J
  const bool keep = mode == RhoMode::Keep && deposit_marches(s);
  if (!s.dep_accumulate) {
    PhaseTimer pz(s, 4);
    if (keep)
      launch_zero(s, s.d_J + g.vol, 3 * g.vol);
    else
      launch_zero(s, s.d_J, 4 * g.vol);
  }
  bool deferred;
  {
    PhaseTimer pt(s, 3);
#define PN_DEP(D, O) deferred = deposit_impl<D, O>(s, dc)
    if (g.dim == 3) {
      switch (g.order) {
      case 1: PN_DEP(3, 1); break;
      case 2: PN_DEP(3, 2); break;
      case 3: PN_DEP(3, 3); break;
      default: PN_DEP(3, 4); break;
      }
    } else if (g.dim == 1) {
      switch (g.order) {
      case 1: PN_DEP(1, 1); break;
      case 2: PN_DEP(1, 2); break;
      case 3: PN_DEP(1, 3); break;
      default: PN_DEP(1, 4); break;
      }
    } else {
      switch (g.order) {
      case 1: PN_DEP(2, 1); break;
      case 2: PN_DEP(2, 2); break;
      case 3: PN_DEP(2, 3); break;
      default: PN_DEP(2, 4); break;
      }
    }
#undef PN_DEP
  }
  if (!deferred)
    return RhoDone::Deposited;
  if (keep)
    return RhoDone::Kept;
  if (mode == RhoMode::Now) {
    launch_rho(s, false);
    return RhoDone::Deposited;
  }
  return RhoDone::Deferred;
}

void launch_rho_unject(Sim& s, int nmax)
{
  const Geom& g  = s.g;
  const int   nk = (int)s.shk.useg.size();
  if (nk == 0 || nmax <= 0)
    return;
  PhaseTimer pt(s, 7);
  dim3       grid(nblk(nmax, 256), nk);
#define PN_UNJ(D, O)                                                                        \
  do {                                                                                        \
    if (s.fp32)                                                                               \
      hipLaunchKernelGGL((k_rho_unject<D, O, float>), grid, dim3(256), 0, s.stream, g,        \
                         s.shk.d_useg, s.shk.d_count, s.shk.d_voff, s.shk.d_vals, s.d_off,    \
                         s.d_lim, s.f_J);                                                     \
    else                                                                                      \
      hipLaunchKernelGGL((k_rho_unject<D, O, double>), grid, dim3(256), 0, s.stream, g,       \
                         s.shk.d_useg, s.shk.d_count, s.shk.d_voff, s.shk.d_vals, s.d_off,    \
                         s.d_lim, s.d_J);                                                     \
  } while (0)
  if (g.dim == 3) {
    switch (g.order) {
    case 1: PN_UNJ(3, 1); break;
    case 2: PN_UNJ(3, 2); break;
    case 3: PN_UNJ(3, 3); break;
    default: PN_UNJ(3, 4); break;
    }
  } else if (g.dim == 1) {
    switch (g.order) {
    case 1: PN_UNJ(1, 1); break;
    case 2: PN_UNJ(1, 2); break;
    case 3: PN_UNJ(1, 3); break;
    default: PN_UNJ(1, 4); break;
    }
  } else {
    switch (g.order) {
    case 1: PN_UNJ(2, 1); break;
    case 2: PN_UNJ(2, 2); break;
    case 3: PN_UNJ(2, 3); break;
    default: PN_UNJ(2, 4); break;
    }
  }
#undef PN_UNJ
}

void launch_rho(Sim& s, bool sorted_by_new_cell)
{
  PhaseTimer  pt(s, 7);
  PhaseTimer  pr(s, kPhaseRhoDeposit);
  const Geom& g = s.g;
#define PN_RHO(D, O) rho_impl<D, O>(s, sorted_by_new_cell)
  if (g.dim == 3) {
    switch (g.order) {
    case 1: PN_RHO(3, 1); break;
    case 2: PN_RHO(3, 2); break;
    case 3: PN_RHO(3, 3); break;
    default: PN_RHO(3, 4); break;
    }
  } else if (g.dim == 1) {
    switch (g.order) {
    case 1: PN_RHO(1, 1); break;
    case 2: PN_RHO(1, 2); break;
    case 3: PN_RHO(1, 3); break;
    default: PN_RHO(1, 4); break;
    }
  } else {
    switch (g.order) {
    case 1: PN_RHO(2, 1); break;
    case 2: PN_RHO(2, 2); break;
    case 3: PN_RHO(2, 3); break;
    default: PN_RHO(2, 4); break;
    }
  }
#undef PN_RHO
}

} // namespace picnix
