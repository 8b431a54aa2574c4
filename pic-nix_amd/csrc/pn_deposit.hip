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
  // chunk-local ghost-inclusive node -> global padded index (absent z maps to 0)
  int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
  int gy = off[c * 3 + 1] + ly;
  int gx = off[c * 3 + 2] + lx;
  return ((long)gz * g.P[1] + gy) * g.P[2] + gx;
}

// As gnode, or -1 for a node outside the padded grid (a particle placed outside
// its chunk through pn_set_particles): such contributions are dropped.
__device__ __forceinline__ long gnode_checked(const Geom& g, const int* __restrict__ off, int c,
                                              int lz, int ly, int lx)
{
  int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
  int gy = off[c * 3 + 1] + ly;
  int gx = off[c * 3 + 2] + lx;
  if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
      (unsigned)gx >= (unsigned)g.P[2])
    return -1;
  return ((long)gz * g.P[1] + gy) * g.P[2] + gx;
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
                          int gy = off[c * 3 + 1] + ly;
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

// Plane kernel (3D current): one block per (bin row (bz, by), segment); each
// wavefront takes whole cells of the row from a block-level counter.  All
// particles of a bin share the tile origin (the old cell, pic/engine/
// current.hpp:390-392), so a cell's contributions are summed in registers.
// Jx, Jy and Jz have the same structure: a plane weight
//   w = -((s0F[f] + A dsF[f]) s0S[s] + (A s0F[f] + B dsF[f]) dsS[s]) q d/dt
// with (F, S) = (y, z), (z, x), (x, y), followed by a prefix sum along x, y, z
// (nix/esirkepov.hpp:168-232).  A lane owns one plane of one component -- N rows
// of N values, the rows running across the plane and the prefix along each row
// -- for one particle of a slot of PPS particles, so every active lane evaluates
// the same expression on its own operands.  Shape weights are computed
// lane-parallel per batch and read back from LDS.  Each value added is the
// reference's per-particle tile entry; per lane the additions follow the sorted
// particle order.  After the cell, the PPS partial planes are summed with xor
// shuffles and added to the block's LDS tile; the tile goes to global J once per
// block.  rho is deposited after the sort (k_rho_cells) or by k_rho_global.
template <int Order>
struct PlanePlan {
  static constexpr int N     = Order + 3;
  static constexpr int NPL   = 3 * N; // planes per particle (3 components x N)
  static constexpr int PPS   = 64 / NPL >= 4 ? 4 : (64 / NPL >= 2 ? 2 : 1);
  static constexpr int BATCH = 16;
  static constexpr int PAIRS = 0;                   // (s0[j], ds[j]) per axis: [3][N][2]
  static constexpr int DSC   = 6 * N;               // ds per axis, contiguous: [3][N]
  static constexpr int FLAG  = 9 * N;               // 1: origin of the bin, 0: own origin
  static constexpr int ORG   = 9 * N + 1;           // own tile origin (x, y, z), local nodes
  static constexpr int REC   = (9 * N + 5) & ~1;    // doubles per particle record
};

template <int Order>
__global__ void __launch_bounds__(256)
    k_deposit_planes(Geom g, const PArr p, const double* __restrict__ xo,
                     const double* __restrict__ yo, const double* __restrict__ zo,
                     const int32_t* __restrict__ cellstart, const int* __restrict__ off,
                     const double* __restrict__ lim, double* __restrict__ J, DepConsts dc)
{
  using PP             = PlanePlan<Order>;
  constexpr int N      = PP::N;
  constexpr int NPL    = PP::NPL;
  constexpr int PPS    = PP::PPS;
  constexpr int BATCH  = PP::BATCH;
  constexpr int REC    = PP::REC;
  constexpr int is_odd = Order % 2;
  constexpr int NWAVE  = 4;

  __shared__ double wts[NWAVE][BATCH][REC];
  __shared__ int    next_cell;
  extern __shared__ double tile[]; // [3][N][N][TX] (Jx, Jy, Jz)

  const int seg = blockIdx.y;
  const int c   = seg / g.Ns;
  const int sp  = seg - c * g.Ns;
  const int nbx = g.bsy;
  const int nby = g.bsz / g.bsy;
  const int bz  = blockIdx.x / nby;
  const int by  = blockIdx.x - bz * nby;
  const int32_t* cs = cellstart + (long)seg * (g.nbin + 1);
  const int      b0 = bz * g.bsz + by * g.bsy;
  if (cs[b0 + nbx] <= cs[b0])
    return; // empty row (uniform for the block)

  const int lb  = g.Nb;
  const int zlo = bz - is_odd + lb - (Order / 2) - 1; // tile origin (local node index)
  const int ylo = by - is_odd + lb - (Order / 2) - 1;
  const int xlo = -is_odd + lb - (Order / 2) - 1;
  const int TX  = nbx + N - 1;
  const int tsz = N * N * TX;
  for (int t = threadIdx.x; t < 3 * tsz; t += blockDim.x)
    tile[t] = 0.0;
  if (threadIdx.x == 0)
    next_cell = 0;
  __syncthreads();

  const int    lane = threadIdx.x & 63;
  const int    wv   = threadIdx.x >> 6;
  const long   v    = g.vol;
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const double qs   = dc.q[sp];
  const double A    = 1.0 / 2;
  const double B    = 1.0 / 3;

  // lane -> (plane, slot particle k); plane -> (component, plane index t)
  const bool   act   = lane < NPL * PPS;
  const int    plane = act ? lane / PPS : 0;
  const int    k     = lane - (lane / PPS) * PPS;
  const int    comp  = plane / N;            // 0 Jx, 1 Jy, 2 Jz
  const int    t     = plane - comp * N;
  // axes (0 x, 1 y, 2 z): F, S of the plane weight, C the prefix direction
  const int    axF   = comp == 0 ? 1 : (comp == 1 ? 2 : 0);
  const int    axS   = comp == 0 ? 2 : (comp == 1 ? 0 : 1);
  const int    axC   = comp;
  const bool   pIsS  = comp != 1;            // the plane index runs along S (else F)
  const int    axP   = pIsS ? axS : axF;
  const int    axR   = pIsS ? axF : axS;
  const double qd    = qs * (comp == 0 ? dc.dxdt : (comp == 1 ? dc.dydt : dc.dzdt));

  double* W = &wts[wv][0][0];

  for (;;) {
    int bx = 0;
    if (lane == 0)
      bx = atomicAdd(&next_cell, 1);
    bx = __shfl(bx, 0);
    if (bx >= nbx)
      break;
    const int p0 = cs[b0 + bx], p1 = cs[b0 + bx + 1];
    if (p0 >= p1)
      continue;
    const int ox = bx + xlo;

    double acc[N][N];
#pragma unroll
    for (int r = 0; r < N; r++)
#pragma unroll
      for (int j = 0; j < N; j++)
        acc[r][j] = 0.0;

    for (int pb = p0; pb < p1; pb += BATCH) {
      const int nbt = min(BATCH, p1 - pb);
      // ---- phase 1: shape weights, lane-parallel ----
      if (lane < nbt) {
        const long   i       = (long)seg * g.cap + pb + lane;
        double       xv[3]   = {xo[i], yo[i], zo[i]};
        double       xu[3]   = {p.x[i], p.y[i], p.z[i]};
        const double d3[3]   = {g.del[2], g.del[1], g.del[0]};
        const double mn[3]   = {xmin, ymin, zmin};
        const int    want[3] = {ox, ylo, zlo};
        double*      rec     = W + lane * REC;
        bool         same    = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const double rd    = 1 / d3[a];
          const double imin  = mn[a] + 0.5 * d3[a] * is_odd;
          const double grid0 = mn[a] + 0.5 * d3[a];
          double       s0[N], s1[N];
#pragma unroll
          for (int l = 0; l < N; l++)
            s0[l] = s1[l] = 0.0;
          int i0 = digitize(xv[a], imin, rd);
          int i1 = digitize(xu[a], imin, rd);
          shape_mc<Order>(xv[a], grid0 + static_cast<double>(i0) * d3[a], rd, &s0[1]);
          shape_mc<Order>(xu[a], grid0 + static_cast<double>(i1) * d3[a], rd, &s1[1]);
          if (i1 - i0 < 0) {
#pragma unroll
            for (int ii = 0; ii < Order + 2; ii++)
              s1[ii] = s1[ii + 1];
          } else if (i1 - i0 > 0) {
#pragma unroll
            for (int ii = Order + 2; ii > 0; ii--)
              s1[ii] = s1[ii - 1];
          }
          const int org  = i0 + lb - (Order / 2) - 1;
          same           = same && (org == want[a]);
          rec[PP::ORG + a] = static_cast<double>(org);
#pragma unroll
          for (int l = 0; l < N; l++) {
            const double ds = s1[l] - s0[l]; // DS (esirkepov.hpp:157-166)
            rec[PP::PAIRS + (a * N + l) * 2]     = s0[l];
            rec[PP::PAIRS + (a * N + l) * 2 + 1] = ds;
            rec[PP::DSC + a * N + l]             = ds;
          }
        }
        rec[PP::FLAG] = same ? 1.0 : 0.0;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // ---- phase 2: each lane accumulates its plane ----
      for (int k0 = 0; k0 < nbt; k0 += PPS) {
        const int kk = k0 + k;
        if (act && kk < nbt) {
          const double* rec = W + kk * REC;
          const double  pa0 = rec[PP::PAIRS + (axP * N + t) * 2];     // s0 along P at t
          const double  pa1 = rec[PP::PAIRS + (axP * N + t) * 2 + 1]; // ds along P at t
          double        dsC[N - 1];
#pragma unroll
          for (int j = 0; j < N - 1; j++)
            dsC[j] = rec[PP::DSC + axC * N + j];
          if (rec[PP::FLAG] != 0.0) {
#pragma unroll
            for (int r = 0; r < N; r++) {
              const double ra0 = rec[PP::PAIRS + (axR * N + r) * 2];
              const double ra1 = rec[PP::PAIRS + (axR * N + r) * 2 + 1];
              const double s0F = pIsS ? ra0 : pa0, dsF = pIsS ? ra1 : pa1;
              const double s0S = pIsS ? pa0 : ra0, dsS = pIsS ? pa1 : ra1;
              const double w   = -((1 * s0F + A * dsF) * s0S + (A * s0F + B * dsF) * dsS) * qd;
              double       ww  = 0;
#pragma unroll
              for (int j = 0; j < N - 1; j++) {
                ww += dsC[j] * w;
                acc[r][j + 1] += ww;
              }
            }
          } else {
            // the particle's tile origin differs from its bin's (odd orders: the
            // bin and shape digitize offsets differ by half a cell, and the two
            // roundings can disagree): same values, added at its own nodes
            const int pox = (int)rec[PP::ORG + 0];
            const int poy = (int)rec[PP::ORG + 1];
            const int poz = (int)rec[PP::ORG + 2];
            for (int r = 0; r < N; r++) {
              const double ra0 = rec[PP::PAIRS + (axR * N + r) * 2];
              const double ra1 = rec[PP::PAIRS + (axR * N + r) * 2 + 1];
              const double s0F = pIsS ? ra0 : pa0, dsF = pIsS ? ra1 : pa1;
              const double s0S = pIsS ? pa0 : ra0, dsS = pIsS ? pa1 : ra1;
              const double w   = -((1 * s0F + A * dsF) * s0S + (A * s0F + B * dsF) * dsS) * qd;
              double       ww  = 0;
              for (int j = 0; j < N - 1; j++) {
                ww += dsC[j] * w;
                int lz, ly, lx;
                if (comp == 0) {
                  lz = poz + t, ly = poy + r, lx = pox + j + 1;
                } else if (comp == 1) {
                  lz = poz + t, ly = poy + j + 1, lx = pox + r;
                } else {
                  lz = poz + j + 1, ly = poy + t, lx = pox + r;
                }
                const int tz = lz - zlo, ty = ly - ylo, tx = lx - xlo;
                if ((unsigned)tz < (unsigned)N && (unsigned)ty < (unsigned)N &&
                    (unsigned)tx < (unsigned)TX)
                  atomicAdd(&tile[((comp * N + tz) * N + ty) * TX + tx], ww);
                else if (long gi = gnode_checked(g, off, c, lz, ly, lx); gi >= 0)
                  unsafeAtomicAdd(&J[(comp + 1) * v + gi], ww);
              }
            }
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // ---- reduce the PPS partial planes and add them to the block tile ----
#pragma unroll
    for (int r = 0; r < N; r++)
#pragma unroll
      for (int j = 1; j < N; j++) {
        double a = acc[r][j];
        if (PPS >= 2)
          a += __shfl_xor(a, 1);
        if (PPS >= 4)
          a += __shfl_xor(a, 2);
        acc[r][j] = a;
      }
    if (act && k == 0) {
      const int tx0 = ox - xlo;
#pragma unroll
      for (int r = 0; r < N; r++)
#pragma unroll
        for (int j = 1; j < N; j++) {
          // node: P-axis = t, R-axis = r, C-axis = j
          int idx;
          if (comp == 0) // Jx: (z = t, y = r, x = j)
            idx = ((0 * N + t) * N + r) * TX + tx0 + j;
          else if (comp == 1) // Jy: (z = t, y = j, x = r)
            idx = ((1 * N + t) * N + j) * TX + tx0 + r;
          else // Jz: (z = j, y = t, x = r)
            idx = ((2 * N + j) * N + t) * TX + tx0 + r;
          atomicAdd(&tile[idx], acc[r][j]);
        }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * tsz; q += blockDim.x) {
    double val = tile[q];
    if (val == 0.0)
      continue;
    int kc = q / tsz;
    int r2 = q - kc * tsz;
    int tz = r2 / (N * TX);
    int r3 = r2 - tz * N * TX;
    int ty = r3 / TX;
    int tx = r3 - ty * TX;
    const long gi = gnode_checked(g, off, c, tz + zlo, ty + ylo, tx + xlo);
    if (gi >= 0)
      unsafeAtomicAdd(&J[(kc + 1) * v + gi], val);
  }
}

// rho from the new positions (3D), one block per (bin row, segment) after the
// sort: the reference's rho term qs S1x S1y S1z (ro3d, nix/esirkepov.hpp:145-155)
// uses the new-position weights, which -- once shifted back -- are the (Order+1)
// unshifted weights around the new cell.  Lanes own (jz plane, slot particle);
// PPS_R particles per slot; partials are xor-shuffle reduced per cell.
template <int Order>
struct RhoPlan {
  static constexpr int S     = Order + 1;
  static constexpr int PPS   = 64 / S >= 32 ? 32 : (64 / S >= 16 ? 16 : 8);
  static constexpr int BATCH = 64;
  static constexpr int FLAG  = 3 * S;      // 1: origin of the bin, 0: own origin
  static constexpr int ORG   = 3 * S + 1;  // own origin (x, y, z), local nodes
  static constexpr int REC   = (3 * S + 5) & ~1;
};

template <int Order>
__global__ void __launch_bounds__(256)
    k_rho_cells(Geom g, const PArr p, const int32_t* __restrict__ cellstart,
                const int* __restrict__ off, const double* __restrict__ lim,
                double* __restrict__ J)
{
  using RP             = RhoPlan<Order>;
  constexpr int S      = RP::S;
  constexpr int PPS    = RP::PPS;
  constexpr int BATCH  = RP::BATCH;
  constexpr int REC    = RP::REC;
  constexpr int is_odd = Order % 2;
  constexpr int NWAVE  = 4;

  __shared__ double wts[NWAVE][BATCH][REC];
  __shared__ int    next_cell;
  extern __shared__ double tile[]; // [S][S][TX]

  const int seg = blockIdx.y;
  const int c   = seg / g.Ns;
  const int sp  = seg - c * g.Ns;
  const int nbx = g.bsy;
  const int nby = g.bsz / g.bsy;
  const int bz  = blockIdx.x / nby;
  const int by  = blockIdx.x - bz * nby;
  const int32_t* cs = cellstart + (long)seg * (g.nbin + 1);
  const int      b0 = bz * g.bsz + by * g.bsy;
  if (cs[b0 + nbx] <= cs[b0])
    return;

  const int lb  = g.Nb;
  const int zlo = bz - is_odd + lb - (Order / 2);
  const int ylo = by - is_odd + lb - (Order / 2);
  const int xlo = -is_odd + lb - (Order / 2);
  const int TX  = nbx + Order;
  const int tsz = S * S * TX;
  for (int q = threadIdx.x; q < tsz; q += blockDim.x)
    tile[q] = 0.0;
  if (threadIdx.x == 0)
    next_cell = 0;
  __syncthreads();

  const int    lane = threadIdx.x & 63;
  const int    wv   = threadIdx.x >> 6;
  const long   v    = g.vol;
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const double qs   = g.q[sp];
  const bool   act  = lane < S * PPS;
  const int    t    = act ? lane / PPS : 0; // jz plane
  const int    k    = lane - (lane / PPS) * PPS;
  double*      W    = &wts[wv][0][0];

  for (;;) {
    int bx = 0;
    if (lane == 0)
      bx = atomicAdd(&next_cell, 1);
    bx = __shfl(bx, 0);
    if (bx >= nbx)
      break;
    const int p0 = cs[b0 + bx], p1 = cs[b0 + bx + 1];
    if (p0 >= p1)
      continue;
    double acc[S][S];
#pragma unroll
    for (int a = 0; a < S; a++)
#pragma unroll
      for (int b = 0; b < S; b++)
        acc[a][b] = 0.0;
    for (int pb = p0; pb < p1; pb += BATCH) {
      const int nbt = min(BATCH, p1 - pb);
      if (lane < nbt) {
        const long   i       = (long)seg * g.cap + pb + lane;
        double       xu[3]   = {p.x[i], p.y[i], p.z[i]};
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
          if (rec[RP::FLAG] != 0.0) {
#pragma unroll
            for (int jy = 0; jy < S; jy++) {
              const double sy = rec[S + jy];
#pragma unroll
              for (int jx = 0; jx < S; jx++)
                acc[jy][jx] += qs * sx[jx] * sy * sz;
            }
          } else {
            // own origin (odd orders, see k_deposit_planes)
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
                    (unsigned)tx < (unsigned)TX)
                  atomicAdd(&tile[(tz * S + ty) * TX + tx], val);
                else if (long gi = gnode_checked(g, off, c, lz, ly, lx); gi >= 0)
                  unsafeAtomicAdd(&J[gi], val);
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
      for (int b = 0; b < S; b++) {
        double x = acc[a][b];
#pragma unroll
        for (int m = 1; m < PPS; m <<= 1)
          x += __shfl_xor(x, m);
        acc[a][b] = x;
      }
    if (act && k == 0) {
#pragma unroll
      for (int jy = 0; jy < S; jy++)
#pragma unroll
        for (int jx = 0; jx < S; jx++)
          atomicAdd(&tile[(t * S + jy) * TX + bx + jx], acc[jy][jx]);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < tsz; q += blockDim.x) {
    double val = tile[q];
    if (val == 0.0)
      continue;
    int tz = q / (S * TX);
    int r3 = q - tz * S * TX;
    int ty = r3 / TX;
    int tx = r3 - ty * TX;
    const long gi = gnode_checked(g, off, c, tz + zlo, ty + ylo, tx + xlo);
    if (gi >= 0)
      unsafeAtomicAdd(&J[gi], val);
  }
}

// rho from the new positions, one thread per particle (any order; 2D or 3D).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_rho_global(Geom g, const PArr p, const int32_t* __restrict__ np,
                 const int* __restrict__ off, const double* __restrict__ lim,
                 double* __restrict__ J)
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
  const double xu[3] = {p.x[i], p.y[i], p.z[i]};
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
    for (int jy = 0; jy < S; jy++)
      for (int jx = 0; jx < S; jx++) {
        double    val = Dim == 3 ? qs * s[0][jx] * s[1][jy] * s[2][jz] : qs * s[0][jx] * s[1][jy];
        const int  lz = Dim == 3 ? o[2] + jz : g.Nb;
        const long gi = gnode_checked(g, off, c, lz, o[1] + jy, o[0] + jx);
        if (gi >= 0)
          unsafeAtomicAdd(&J[gi], val);
      }
}

// Returns true when rho still has to be deposited (3D plane path).
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
  const size_t tlds = sizeof(double) * 3 * (size_t)N * N * (g.bsy + N - 1);
  const size_t wlds = sizeof(double) * 4 * PlanePlan<Order>::BATCH * PlanePlan<Order>::REC;
  if (Dim == 3 && s.sorted && !s.generic_kernels && tlds + wlds + 64 <= 65536) {
    const int nrow = g.nbin / g.bsy;
    hipLaunchKernelGGL((k_deposit_planes<Order>), dim3(nrow, g.nseg), dim3(256), tlds, s.stream,
                       g, s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.ss.cellstart, s.d_off,
                       s.d_lim, s.d_J, dc);
    return true;
  }
  if (s.sorted && lds <= 65536 && !s.generic_kernels) {
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

template <int Dim, int Order>
void rho_impl(Sim& s, bool sorted_by_new_cell)
{
  const Geom& g = s.g;
  const int   S = Order + 1;
  const size_t tlds = sizeof(double) * (size_t)S * S * (g.bsy + Order);
  if (Dim == 3 && sorted_by_new_cell && !s.generic_kernels && tlds <= 32768) {
    const int nrow = g.nbin / g.bsy;
    hipLaunchKernelGGL((k_rho_cells<Order>), dim3(nrow, g.nseg), dim3(256), tlds, s.stream, g,
                       s.pa, s.ss.cellstart, s.d_off, s.d_lim, s.d_J);
  } else {
    hipLaunchKernelGGL((k_rho_global<Dim, Order>), dim3(nblk(g.cap, 256), g.nseg), dim3(256), 0,
                       s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_J);
  }
}

} // namespace

bool launch_deposit(Sim& s, double delt, bool defer_rho)
{
  const Geom& g = s.g;
  DepConsts   dc{};
  for (int sp = 0; sp < g.Ns; sp++)
    dc.q[sp] = g.q[sp];
  dc.delt = delt;
  dc.dxdt = g.del[2] / delt;
  dc.dydt = g.del[1] / delt;
  dc.dzdt = g.del[0] / delt;
  {
    PhaseTimer pz(s, 4);
    hipMemsetAsync(s.d_J, 0, sizeof(double) * 4 * (size_t)g.vol, s.stream);
  }
  bool need_rho;
  {
    PhaseTimer pt(s, 3);
#define PN_DEP(D, O) need_rho = deposit_impl<D, O>(s, dc)
    if (g.dim == 3) {
      switch (g.order) {
      case 1: PN_DEP(3, 1); break;
      case 2: PN_DEP(3, 2); break;
      case 3: PN_DEP(3, 3); break;
      default: PN_DEP(3, 4); break;
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
  if (need_rho && !defer_rho) {
    launch_rho(s, false);
    need_rho = false;
  }
  return need_rho;
}

void launch_rho(Sim& s, bool sorted_by_new_cell)
{
  PhaseTimer  pt(s, 7);
  const Geom& g = s.g;
#define PN_RHO(D, O) rho_impl<D, O>(s, sorted_by_new_cell)
  if (g.dim == 3) {
    switch (g.order) {
    case 1: PN_RHO(3, 1); break;
    case 2: PN_RHO(3, 2); break;
    case 3: PN_RHO(3, 3); break;
    default: PN_RHO(3, 4); break;
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
