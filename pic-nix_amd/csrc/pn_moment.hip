// pn_moment.hip -- moment deposit (PicChunk::deposit_moment, pic/pic_chunk.cpp:525-534).
//
// Fourteen moments per species at the cell centres, weighted by the Order+1
// unshifted shape weights of the current position (pic/engine/moment.hpp:74-98
// scalar path, local2d/local3d :271-394):
//   t  = m w             x,y,z    = m w u/gamma        tt = m w gamma c
//   xx,yy,zz = m w u u/gamma     tx,ty,tz = m w u     xy,yz,zx = m w u u'/gamma
// in the reference's component order (moment.hpp:19-33).  This is a diagnostic
// (history and moment output), outside the timed step.  Sorted particles take
// one block per (bin row, segment) that sums into an LDS tile (per-bin register
// sums, or LDS atomics per particle for order-4 3D stencils) and adds the tile
// to global memory once; otherwise one thread per particle adds straight to
// global memory.  The Mom exchange (pn_exchange) folds the
// padding as the Cur exchange does for J.
#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// the 14 moment values of one particle for node weight ww (moment.hpp:307-320)
__device__ __forceinline__ void moments14(double ww, double ux, double uy, double uz, double gm,
                                          double cc, double m[14])
{
  m[0]  = ww;
  m[1]  = ww * ux / gm;
  m[2]  = ww * uy / gm;
  m[3]  = ww * uz / gm;
  m[8]  = ww * ux;
  m[9]  = ww * uy;
  m[10] = ww * uz;
  m[4]  = ww * gm * cc;
  m[5]  = ww * ux * ux / gm;
  m[6]  = ww * uy * uy / gm;
  m[7]  = ww * uz * uz / gm;
  m[11] = ww * ux * uy / gm;
  m[12] = ww * uy * uz / gm;
  m[13] = ww * uz * ux / gm;
}

// weights and first node (local, ghost-inclusive) of one particle
template <int Dim, int Order>
__device__ __forceinline__ void moment_weights(const Geom& g, const double* lim, int c,
                                               const double xu[3], double w[3][Order + 1],
                                               int o[3])
{
  constexpr int is_odd = Order % 2;
  const double  d3[3]  = {g.del[2], g.del[1], g.del[0]};
  const double  mn[3]  = {lim[c * 6 + 4], lim[c * 6 + 2], lim[c * 6 + 0]};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (a < Dim) {
      const double rd   = 1 / d3[a];
      const int    cell = digitize(xu[a], mn[a] + 0.5 * d3[a] * is_odd, rd);
      shape_mc<Order>(xu[a], (mn[a] + 0.5 * d3[a]) + static_cast<double>(cell) * d3[a], rd, w[a]);
      o[a] = cell + g.Nb - (Order / 2);
    } else {
      w[a][0] = 1.0;
      o[a]    = g.Nb;
    }
  }
}

template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_moment_global(Geom g, const PArr p, const int32_t* __restrict__ np,
                    const int* __restrict__ off, const double* __restrict__ lim,
                    double* __restrict__ M)
{
  constexpr int S   = Order + 1;
  constexpr int Sz  = Dim == 3 ? S : 1;
  constexpr int Sy  = Dim >= 2 ? S : 1;
  const int     seg = blockIdx.y;
  const long    ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int    c     = seg / g.Ns;
  const int    sp    = seg - c * g.Ns;
  const long   i     = (long)seg * g.cap + ip;
  const double xu[3] = {p.x[i], p.y[i], p.z[i]};
  const double ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
  const double gm = lorentz_factor(ux, uy, uz, 1 / g.cc);
  double       w[3][S];
  int          o[3];
  moment_weights<Dim, Order>(g, lim, c, xu, w, o);
  const double ms   = g.m[sp];
  double*      Msp  = M + (long)sp * 14 * g.vol;
  for (int jz = 0; jz < Sz; jz++)
    for (int jy = 0; jy < Sy; jy++)
      for (int jx = 0; jx < S; jx++) {
        // moment.hpp local1d/2d/3d: the product of the present axes' weights
        const double ww = Dim == 3 ? ms * w[0][jx] * w[1][jy] * w[2][jz]
                                   : (Dim == 2 ? ms * w[0][jx] * w[1][jy] : ms * w[0][jx]);
        const int    gz = g.has[0] ? off[c * 3 + 0] + o[2] + jz : 0;
        const int    gy = g.has[1] ? off[c * 3 + 1] + o[1] + jy : 0;
        const int    gx = off[c * 3 + 2] + o[0] + jx;
        if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
            (unsigned)gx >= (unsigned)g.P[2])
          continue; // a particle placed outside its chunk
        const long gi = ((long)gz * g.P[1] + gy) * g.P[2] + gx;
        double     m[14];
        moments14(ww, ux, uy, uz, gm, g.cc, m);
#pragma unroll
        for (int k = 0; k < 14; k++)
          unsafeAtomicAdd(&Msp[k * g.vol + gi], m[k]);
      }
}

// One block per (bin row, segment) of sorted particles: the row's particles
// touch an Sz x S x TX block of nodes, summed in LDS with one LDS atomic per
// particle, node and component (the form for stencils too large for registers).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_moment_rows_atomic(Geom g, const PArr p, const int32_t* __restrict__ cellstart,
                  const int* __restrict__ off, const double* __restrict__ lim,
                  double* __restrict__ M)
{
  constexpr int S      = Order + 1;
  constexpr int Sz     = Dim == 3 ? S : 1;
  constexpr int is_odd = Order % 2;
  extern __shared__ double tile[]; // [14][Sz][S][TX]
  const int seg = blockIdx.y;
  const int c   = seg / g.Ns;
  const int sp  = seg - c * g.Ns;
  const int nbx = g.bsy;
  int       bz = 0, by;
  if constexpr (Dim == 3) {
    const int nby = g.bsz / g.bsy;
    bz            = blockIdx.x / nby;
    by            = blockIdx.x - bz * nby;
  } else {
    by = blockIdx.x;
  }
  const int32_t* cs = cellstart + (long)seg * (g.nbin + 1);
  const int      b0 = bz * g.bsz + by * g.bsy;
  const int      p0 = cs[b0], p1 = cs[b0 + nbx];
  if (p1 <= p0)
    return;
  const int lb  = g.Nb;
  const int zlo = Dim == 3 ? bz - is_odd + lb - (Order / 2) : lb;
  const int ylo = by - is_odd + lb - (Order / 2);
  const int xlo = -is_odd + lb - (Order / 2);
  const int TX  = nbx + Order;
  const int csz = Sz * S * TX; // one component
  for (int q = threadIdx.x; q < 14 * csz; q += blockDim.x)
    tile[q] = 0.0;
  __syncthreads();

  const double ms  = g.m[sp];
  double*      Msp = M + (long)sp * 14 * g.vol;
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const long   i     = (long)seg * g.cap + q;
    const double xu[3] = {p.x[i], p.y[i], p.z[i]};
    const double ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
    const double gm = lorentz_factor(ux, uy, uz, 1 / g.cc);
    double       w[3][S];
    int          o[3];
    moment_weights<Dim, Order>(g, lim, c, xu, w, o);
    for (int jz = 0; jz < Sz; jz++)
      for (int jy = 0; jy < S; jy++)
        for (int jx = 0; jx < S; jx++) {
          const double ww =
              Dim == 3 ? ms * w[0][jx] * w[1][jy] * w[2][jz] : ms * w[0][jx] * w[1][jy];
          double m[14];
          moments14(ww, ux, uy, uz, gm, g.cc, m);
          const int lz = o[2] + jz, ly = o[1] + jy, lx = o[0] + jx;
          const int tz = lz - zlo, ty = ly - ylo, tx = lx - xlo;
          if ((unsigned)tz < (unsigned)Sz && (unsigned)ty < (unsigned)S &&
              (unsigned)tx < (unsigned)TX) {
            const int t = (tz * S + ty) * TX + tx;
#pragma unroll
            for (int k = 0; k < 14; k++)
              atomicAdd(&tile[k * csz + t], m[k]);
          } else {
            // off the row's block (a rounding edge between the bin and shape offsets)
            const int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
            const int gy = off[c * 3 + 1] + ly;
            const int gx = off[c * 3 + 2] + lx;
            if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
                (unsigned)gx >= (unsigned)g.P[2])
              continue;
            const long gi = ((long)gz * g.P[1] + gy) * g.P[2] + gx;
#pragma unroll
            for (int k = 0; k < 14; k++)
              unsafeAtomicAdd(&Msp[k * g.vol + gi], m[k]);
          }
        }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 14 * csz; q += blockDim.x) {
    const double val = tile[q];
    if (val == 0.0)
      continue;
    const int k  = q / csz;
    const int r  = q - k * csz;
    const int tz = r / (S * TX);
    const int r2 = r - tz * S * TX;
    const int ty = r2 / TX;
    const int tx = r2 - ty * TX;
    const int gz = g.has[0] ? off[c * 3 + 0] + zlo + tz : 0;
    const int gy = off[c * 3 + 1] + ylo + ty;
    const int gx = off[c * 3 + 2] + xlo + tx;
    if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
        (unsigned)gx >= (unsigned)g.P[2])
      continue;
    unsafeAtomicAdd(&Msp[k * g.vol + ((long)gz * g.P[1] + gy) * g.P[2] + gx], val);
  }
}

// component k of moments14 divided by the node weight
__device__ __forceinline__ double moment_factor(int k, double ux, double uy, double uz, double gm,
                                                double cc)
{
  switch (k) {
  case 0: return 1.0;
  case 1: return ux / gm;
  case 2: return uy / gm;
  case 3: return uz / gm;
  case 4: return gm * cc;
  case 5: return ux * ux / gm;
  case 6: return uy * uy / gm;
  case 7: return uz * uz / gm;
  case 8: return ux;
  case 9: return uy;
  case 10: return uz;
  case 11: return ux * uy / gm;
  case 12: return uy * uz / gm;
  default: return uz * ux / gm;
  }
}

// One block per (bin row, segment) of sorted particles, one work item per
// (component, bin of the row): the item sums its component over the bin's
// particles in registers on the Sz x S x S nodes every particle of the bin
// touches (the bin fixes the first node: bins are offset by half a cell for odd
// orders, as the shape is), then adds the sums to the row's LDS tile once.  A
// particle whose first node differs from its bin's (a rounding edge) adds
// straight to the tile or, off the tile, to global memory.
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_moment_rows(Geom g, const PArr p, const int32_t* __restrict__ cellstart,
                  const int* __restrict__ off, const double* __restrict__ lim,
                  double* __restrict__ M)
{
  constexpr int S      = Order + 1;
  constexpr int Sz     = Dim == 3 ? S : 1;
  constexpr int is_odd = Order % 2;
  extern __shared__ double tile[]; // [14][Sz][S][TX]
  const int seg = blockIdx.y;
  const int c   = seg / g.Ns;
  const int sp  = seg - c * g.Ns;
  const int nbx = g.bsy;
  int       bz = 0, by;
  if constexpr (Dim == 3) {
    const int nby = g.bsz / g.bsy;
    bz            = blockIdx.x / nby;
    by            = blockIdx.x - bz * nby;
  } else {
    by = blockIdx.x;
  }
  const int32_t* cs = cellstart + (long)seg * (g.nbin + 1);
  const int      b0 = bz * g.bsz + by * g.bsy;
  if (cs[b0 + nbx] <= cs[b0])
    return;
  const int lb  = g.Nb;
  const int zlo = Dim == 3 ? bz - is_odd + lb - (Order / 2) : lb;
  const int ylo = by - is_odd + lb - (Order / 2);
  const int xlo = -is_odd + lb - (Order / 2);
  const int TX  = nbx + Order;
  const int csz = Sz * S * TX; // one component
  for (int q = threadIdx.x; q < 14 * csz; q += blockDim.x)
    tile[q] = 0.0;
  __syncthreads();

  const double ms  = g.m[sp];
  double*      Msp = M + (long)sp * 14 * g.vol;
  // consecutive items: consecutive bins of one component (few distinct k per wave)
  for (int w = threadIdx.x; w < 14 * nbx; w += blockDim.x) {
    const int k  = w / nbx;
    const int bx = w - k * nbx;
    double    acc[Sz][S][S];
#pragma unroll
    for (int jz = 0; jz < Sz; jz++)
#pragma unroll
      for (int jy = 0; jy < S; jy++)
#pragma unroll
        for (int jx = 0; jx < S; jx++)
          acc[jz][jy][jx] = 0.0;
    const int q1 = cs[b0 + bx + 1];
    for (int q = cs[b0 + bx]; q < q1; q++) {
      const long   i     = (long)seg * g.cap + q;
      const double xu[3] = {p.x[i], p.y[i], p.z[i]};
      const double ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
      const double gm = lorentz_factor(ux, uy, uz, 1 / g.cc);
      double       wt[3][S];
      int          o[3];
      moment_weights<Dim, Order>(g, lim, c, xu, wt, o);
      const double f = moment_factor(k, ux, uy, uz, gm, g.cc);
      if (o[0] == xlo + bx && o[1] == ylo && o[2] == zlo) {
#pragma unroll
        for (int jz = 0; jz < Sz; jz++)
#pragma unroll
          for (int jy = 0; jy < S; jy++) {
            const double wyz = Dim == 3 ? ms * wt[1][jy] * wt[2][jz] : ms * wt[1][jy];
#pragma unroll
            for (int jx = 0; jx < S; jx++)
              acc[jz][jy][jx] += wyz * wt[0][jx] * f;
          }
        continue;
      }
      for (int jz = 0; jz < Sz; jz++)
        for (int jy = 0; jy < S; jy++)
          for (int jx = 0; jx < S; jx++) {
            const double val = (Dim == 3 ? ms * wt[1][jy] * wt[2][jz] : ms * wt[1][jy]) *
                               wt[0][jx] * f;
            const int    lz = o[2] + jz, ly = o[1] + jy, lx = o[0] + jx;
            const int    tz = lz - zlo, ty = ly - ylo, tx = lx - xlo;
            if ((unsigned)tz < (unsigned)Sz && (unsigned)ty < (unsigned)S &&
                (unsigned)tx < (unsigned)TX) {
              atomicAdd(&tile[k * csz + (tz * S + ty) * TX + tx], val);
            } else {
              const int gz = g.has[0] ? off[c * 3 + 0] + lz : 0;
              const int gy = off[c * 3 + 1] + ly;
              const int gx = off[c * 3 + 2] + lx;
              if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
                  (unsigned)gx >= (unsigned)g.P[2])
                continue;
              unsafeAtomicAdd(&Msp[k * g.vol + ((long)gz * g.P[1] + gy) * g.P[2] + gx], val);
            }
          }
    }
#pragma unroll
    for (int jz = 0; jz < Sz; jz++)
#pragma unroll
      for (int jy = 0; jy < S; jy++)
#pragma unroll
        for (int jx = 0; jx < S; jx++)
          if (acc[jz][jy][jx] != 0.0)
            atomicAdd(&tile[k * csz + (jz * S + jy) * TX + bx + jx], acc[jz][jy][jx]);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 14 * csz; q += blockDim.x) {
    const double val = tile[q];
    if (val == 0.0)
      continue;
    const int k  = q / csz;
    const int r  = q - k * csz;
    const int tz = r / (S * TX);
    const int r2 = r - tz * S * TX;
    const int ty = r2 / TX;
    const int tx = r2 - ty * TX;
    const int gz = g.has[0] ? off[c * 3 + 0] + zlo + tz : 0;
    const int gy = off[c * 3 + 1] + ylo + ty;
    const int gx = off[c * 3 + 2] + xlo + tx;
    if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
        (unsigned)gx >= (unsigned)g.P[2])
      continue;
    unsafeAtomicAdd(&Msp[k * g.vol + ((long)gz * g.P[1] + gy) * g.P[2] + gx], val);
  }
}

template <int Dim, int Order>
void moment_impl(Sim& s)
{
  const Geom&  g    = s.g;
  const int    S    = Order + 1;
  const int    Sz   = Dim == 3 ? S : 1;
  const size_t tlds = sizeof(double) * 14 * (size_t)Sz * S * (g.bsy + Order);
  const int    nrow = g.nbin / g.bsy;
  constexpr bool regs = (Dim == 3 ? (Order + 1) : 1) * (Order + 1) * (Order + 1) <= 64;
  if constexpr (Dim >= 2 && regs) {
    if (s.sorted && !s.generic_kernels && tlds <= 65536) {
      hipLaunchKernelGGL((k_moment_rows<Dim, Order>), dim3(nrow, g.nseg), dim3(256), tlds,
                         s.stream, g, s.pa, s.ss.cellstart, s.d_off, s.d_lim, s.d_M);
      return;
    }
  }
  if (Dim >= 2 && s.sorted && !s.generic_kernels && tlds <= 65536) {
    hipLaunchKernelGGL((k_moment_rows_atomic<Dim, Order>), dim3(nrow, g.nseg), dim3(256), tlds,
                       s.stream, g, s.pa, s.ss.cellstart, s.d_off, s.d_lim, s.d_M);
  } else {
    hipLaunchKernelGGL((k_moment_global<Dim, Order>), dim3(nblk(g.cap, 256), g.nseg), dim3(256),
                       0, s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_M);
  }
}

} // namespace

void launch_deposit_moment(Sim& s)
{
  const Geom&  g     = s.g;
  const size_t bytes = sizeof(double) * 14 * (size_t)g.Ns * g.vol;
  // fill_all(um, 0) (moment.hpp:80); d_M is allocated by the caller
  (void)hipMemsetAsync(s.d_M, 0, bytes, s.stream);
#define PN_MOM(D, O) moment_impl<D, O>(s)
  if (g.dim == 3) {
    switch (g.order) {
    case 1: PN_MOM(3, 1); break;
    case 2: PN_MOM(3, 2); break;
    case 3: PN_MOM(3, 3); break;
    default: PN_MOM(3, 4); break;
    }
  } else if (g.dim == 1) {
    switch (g.order) {
    case 1: PN_MOM(1, 1); break;
    case 2: PN_MOM(1, 2); break;
    case 3: PN_MOM(1, 3); break;
    default: PN_MOM(1, 4); break;
    }
  } else {
    switch (g.order) {
    case 1: PN_MOM(2, 1); break;
    case 2: PN_MOM(2, 2); break;
    case 3: PN_MOM(2, 3); break;
    default: PN_MOM(2, 4); break;
    }
  }
#undef PN_MOM
}

} // namespace picnix
