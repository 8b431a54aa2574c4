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

template <int Dim, int Order>
void deposit_impl(Sim& s, const DepConsts& dc)
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
  if (s.sorted && lds <= 65536) {
    hipLaunchKernelGGL((k_deposit_slab<Dim, Order>), dim3(nslab, g.nseg), dim3(256), lds,
                       s.stream, g, s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.ss.cellstart,
                       s.d_off, s.d_lim, s.d_J, dc);
  } else {
    hipLaunchKernelGGL((k_deposit_global<Dim, Order>), dim3(nblk(g.cap, 256), g.nseg), dim3(256),
                       0, s.stream, g, s.pa, s.d_xo[0], s.d_xo[1], s.d_xo[2], s.d_np, s.d_off,
                       s.d_lim, s.d_J, dc);
  }
}

} // namespace

void launch_deposit(Sim& s, double delt)
{
  const Geom& g = s.g;
  DepConsts   dc{};
  for (int sp = 0; sp < g.Ns; sp++)
    dc.q[sp] = g.q[sp];
  dc.delt = delt;
  dc.dxdt = g.del[2] / delt;
  dc.dydt = g.del[1] / delt;
  dc.dzdt = g.del[0] / delt;
  hipMemsetAsync(s.d_J, 0, sizeof(double) * 4 * (size_t)g.vol, s.stream);
  if (g.dim == 3) {
    switch (g.order) {
    case 1: deposit_impl<3, 1>(s, dc); break;
    case 2: deposit_impl<3, 2>(s, dc); break;
    case 3: deposit_impl<3, 3>(s, dc); break;
    default: deposit_impl<3, 4>(s, dc); break;
    }
  } else {
    switch (g.order) {
    case 1: deposit_impl<2, 1>(s, dc); break;
    case 2: deposit_impl<2, 2>(s, dc); break;
    case 3: deposit_impl<2, 3>(s, dc); break;
    default: deposit_impl<2, 4>(s, dc); break;
    }
  }
}

} // namespace picnix
