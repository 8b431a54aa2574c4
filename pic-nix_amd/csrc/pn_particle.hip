// pn_particle.hip -- field gather + Boris push, position push + count, and the
// AoS <-> SoA transfers at the C-ABI boundary (FP64 plain path).
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

struct PushConsts {
  double qmdt[kMaxSpecies]; // 0.5 * q / m * delt  (velocity.hpp:130)
  double cdt[3], rcdt[3];   // c dt / d(x, y, z) and inverses, for WT (velocity.hpp:104-107)
  int    pusher;            // 0 Boris, 1 Vay, 2 Higuera-Cary
  int    wt;                // integer-grid weights: 0 MC, 1 WT (velocity.hpp:240-250)
};

// Integer-grid weights of one axis: shape_mc, or shape_wt with c dt / d
// (BaseVelocity::weights{2,3}d, pic/engine/velocity.hpp:262-359).
template <int Order>
__device__ __forceinline__ void int_weights(const PushConsts& pc, int axis, double x, double X,
                                            double rd, double* w)
{
  if (pc.wt)
    shape_wt<Order>(x, X, rd, pc.cdt[axis], pc.rcdt[axis], w);
  else
    shape_mc<Order>(x, X, rd, w);
}

// One thread per particle.  Weights follow BaseVelocity::weights{2,3}d
// (pic/engine/velocity.hpp:262-359, MC interpolation) and the gather follows
// push_scalar{2,3}d (:390-450) with nix::interp::interp{2,3}d_impl_sorted
// (nix/interp.hpp:50-110): (Order+1) points per direction, x innermost.  Chunk-
// local ghost-inclusive indices map onto the global padded grid by adding the
// chunk's cell offset (the padding equals the ghost margin Nb).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k_push_velocity(Geom g, PArr p, const int32_t* __restrict__ np, const int* __restrict__ off,
                    const double* __restrict__ lim, const double* __restrict__ E,
                    const double* __restrict__ B, PushConsts pc)
{
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int S      = Order + 1;
  const int     seg    = blockIdx.y;
  const long    ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int  c    = seg / g.Ns;
  const int  sp   = seg - c * g.Ns;
  const long i    = (long)seg * g.cap + ip;
  const double qmdt = pc.qmdt[sp];

  const double dx = g.del[2], dy = g.del[1], dz = g.del[0];
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const int    lb   = g.Nb;

  double xu[6] = {p.x[i], p.y[i], p.z[i], p.ux[i], p.uy[i], p.uz[i]};

  if constexpr (Dim == 1) {
    // weights1d + push_scalar1d (pic/engine/velocity.hpp:223-251,362-388): the
    // x line of the grid (no y, z padding)
    const double xim = xmin + 0.5 * dx * is_odd;
    const double xhm = xmin + 0.5 * dx * is_odd - 0.5 * dx;
    const double rd  = 1 / dx;
    double       wi[S], wh[S];
    int          i0 = digitize(xu[0], xim, rd);
    int          h0 = digitize(xu[0], xhm, rd);
    int_weights<Order>(pc, 0, xu[0], (xmin + 0.5 * dx) + static_cast<double>(i0) * dx, rd, wi);
    shape_mc<Order>(xu[0], xmin + static_cast<double>(h0) * dx, rd, wh);
    i0 += lb - (Order / 2) + off[c * 3 + 2];
    h0 += lb - (Order / 2) + off[c * 3 + 2];
    if (min(i0, h0) < 0 || max(i0, h0) > g.P[2] - S)
      return;
    const long v      = g.vol;
    auto       interp = [&](const double* F, int x0, const double* w) {
      double r = 0;
#pragma unroll
      for (int jx = 0; jx < S; jx++)
        r += F[x0 + jx] * w[jx];
      return r * qmdt;
    };
    double eb[6];
    eb[0] = interp(E, h0, wh);
    eb[1] = interp(E + v, i0, wi);
    eb[2] = interp(E + 2 * v, i0, wi);
    eb[3] = interp(B, i0, wi);
    eb[4] = interp(B + v, h0, wh);
    eb[5] = interp(B + 2 * v, h0, wh);
    double ux = xu[3], uy = xu[4], uz = xu[5];
    push_momentum(pc.pusher, ux, uy, uz, eb, g.cc);
    p.ux[i] = ux;
    p.uy[i] = uy;
    p.uz[i] = uz;
    return;
  }

  double ximin  = xmin + 0.5 * dx * is_odd;
  double xhmin  = xmin + 0.5 * dx * is_odd - 0.5 * dx;
  double yimin  = ymin + 0.5 * dy * is_odd;
  double yhmin  = ymin + 0.5 * dy * is_odd - 0.5 * dy;
  double xigrid = xmin + 0.5 * dx, xhgrid = xmin;
  double yigrid = ymin + 0.5 * dy, yhgrid = ymin;
  double rdx = 1 / dx, rdy = 1 / dy;

  double wix[S], wiy[S], wiz[S], whx[S], why[S], whz[S];
  int    ix0 = digitize(xu[0], ximin, rdx);
  int    iy0 = digitize(xu[1], yimin, rdy);
  int    hx0 = digitize(xu[0], xhmin, rdx);
  int    hy0 = digitize(xu[1], yhmin, rdy);
  int_weights<Order>(pc, 0, xu[0], xigrid + static_cast<double>(ix0) * dx, rdx, wix);
  int_weights<Order>(pc, 1, xu[1], yigrid + static_cast<double>(iy0) * dy, rdy, wiy);
  shape_mc<Order>(xu[0], xhgrid + static_cast<double>(hx0) * dx, rdx, whx);
  shape_mc<Order>(xu[1], yhgrid + static_cast<double>(hy0) * dy, rdy, why);
  // local ghost-inclusive start index -> global padded index
  ix0 += lb - (Order / 2) + off[c * 3 + 2];
  iy0 += lb - (Order / 2) + off[c * 3 + 1];
  hx0 += lb - (Order / 2) + off[c * 3 + 2];
  hy0 += lb - (Order / 2) + off[c * 3 + 1];

  const long   sy = g.P[2], sz = g.plane, v = g.vol;
  // stencils must lie inside the padded grid; a particle outside its chunk
  // (possible only through pn_set_particles) is left untouched
  if (min(ix0, hx0) < 0 || max(ix0, hx0) > g.P[2] - S || min(iy0, hy0) < 0 ||
      max(iy0, hy0) > g.P[1] - S)
    return;
  double       eb[6];
  if constexpr (Dim == 3) {
    double zimin  = zmin + 0.5 * dz * is_odd;
    double zhmin  = zmin + 0.5 * dz * is_odd - 0.5 * dz;
    double zigrid = zmin + 0.5 * dz, zhgrid = zmin;
    double rdz    = 1 / dz;
    int    iz0    = digitize(xu[2], zimin, rdz);
    int    hz0    = digitize(xu[2], zhmin, rdz);
    int_weights<Order>(pc, 2, xu[2], zigrid + static_cast<double>(iz0) * dz, rdz, wiz);
    shape_mc<Order>(xu[2], zhgrid + static_cast<double>(hz0) * dz, rdz, whz);
    iz0 += lb - (Order / 2) + off[c * 3 + 0];
    hz0 += lb - (Order / 2) + off[c * 3 + 0];
    if (min(iz0, hz0) < 0 || max(iz0, hz0) > g.P[0] - S)
      return;
    auto interp = [&](const double* F, int z0, int y0, int x0, const double* wz,
                      const double* wy, const double* wx) {
      double rz = 0;
#pragma unroll
      for (int jz = 0; jz < S; jz++) {
        double ry = 0;
#pragma unroll
        for (int jy = 0; jy < S; jy++) {
          const double* row = F + (z0 + jz) * sz + (y0 + jy) * sy + x0;
          double        rx  = 0;
#pragma unroll
          for (int jx = 0; jx < S; jx++)
            rx += row[jx] * wx[jx];
          ry += rx * wy[jy];
        }
        rz += ry * wz[jz];
      }
      return rz * qmdt;
    };
    eb[0] = interp(E, iz0, iy0, hx0, wiz, wiy, whx);
    eb[1] = interp(E + v, iz0, hy0, ix0, wiz, why, wix);
    eb[2] = interp(E + 2 * v, hz0, iy0, ix0, whz, wiy, wix);
    eb[3] = interp(B, hz0, hy0, ix0, whz, why, wix);
    eb[4] = interp(B + v, hz0, iy0, hx0, whz, wiy, whx);
    eb[5] = interp(B + 2 * v, iz0, hy0, hx0, wiz, why, whx);
  } else {
    (void)wiz;
    (void)whz;
    auto interp = [&](const double* F, int y0, int x0, const double* wy, const double* wx) {
      double ry = 0;
#pragma unroll
      for (int jy = 0; jy < S; jy++) {
        const double* row = F + (y0 + jy) * sy + x0;
        double        rx  = 0;
#pragma unroll
        for (int jx = 0; jx < S; jx++)
          rx += row[jx] * wx[jx];
        ry += rx * wy[jy];
      }
      return ry * qmdt;
    };
    eb[0] = interp(E, iy0, hx0, wiy, whx);
    eb[1] = interp(E + v, hy0, ix0, why, wix);
    eb[2] = interp(E + 2 * v, iy0, ix0, wiy, wix);
    eb[3] = interp(B, hy0, ix0, why, wix);
    eb[4] = interp(B + v, iy0, hx0, wiy, whx);
    eb[5] = interp(B + 2 * v, hy0, hx0, why, whx);
  }
  push_momentum(pc.pusher, xu[3], xu[4], xu[5], eb, g.cc);
  p.ux[i] = xu[3];
  p.uy[i] = xu[4];
  p.uz[i] = xu[5];
}

// Arguments of the position push (BasePosition::push, pic/engine/position.hpp:
// 117-130, and count(reset), pic/pic_engine.hpp:292-303) when the row gather
// runs it for the particles it holds.
struct Pos64 {
  double*       xo;
  double*       yo;
  double*       zo;
  const double* rng;
  double        delt;
  double        rc;
  int32_t*      cellcnt; // bin counters, nbin per segment
  int32_t*      key;
  int32_t*      slot;
  int32_t*      tilecnt = nullptr; // leavers per (256-particle tile, direction): fused gather
  int           ntile   = 0;
};

// The position push of particle i of chunk c with momentum (ux, uy, uz): writes
// the old and the new position, and the momentum when WriteU or when the
// reflecting wall flips it.  Returns the new bin when the particle stays in the
// chunk; otherwise -1, with key[i] set (-1 - d for a leaver in direction d, also
// returned in dir, or kKeyReplace past the open end).
template <bool WriteU>
__device__ __forceinline__ int position64(const Geom& g, const PArr& p, const Pos64& ps,
                                          const double* __restrict__ lim, int c, long i,
                                          double x, double y, double z, double ux, double uy,
                                          double uz, int& dir)
{
  dir       = -1;
  ps.xo[i]  = x;
  ps.yo[i]  = y;
  ps.zo[i]  = z;
  double gm = lorentz_factor(ux, uy, uz, ps.rc);
  double dt = ps.delt / gm;
  x += ux * dt;
  y += uy * dt;
  z += uz * dt;
  if (g.xwall && x < g.glim[2][0]) {
    // reflecting wall, MainChunk::set_boundary_particle (example/shock/main.cpp:402-416)
    x  = -x + 2 * g.glim[2][0];
    ux = -ux;
    uy = -uy;
    uz = -uz;
    p.ux[i] = ux;
    p.uy[i] = uy;
    p.uz[i] = uz;
  } else if (WriteU) {
    p.ux[i] = ux;
    p.uy[i] = uy;
    p.uz[i] = uz;
  }
  p.x[i] = x;
  p.y[i] = y;
  p.z[i] = z;
  if (g.xwall && x >= g.glim[2][1]) {
    // past the open end: replaced by an injected particle (k_wall_apply)
    ps.key[i] = kKeyReplace;
    return -1;
  }
  const int b = bin_of(g, lim, ps.rng, c, x, y, z);
  if (b < 0) {
    dir       = dir_of(ps.rng, c, x, y, z);
    ps.key[i] = -1 - dir;
  }
  return b;
}

// Row form of the gather: one block per (bin row, segment).  The E and B
// values every particle of the row can touch -- (Order+2) nodes around the
// row's (z, y) bin and its whole x extent -- are staged in LDS once; each
// particle then gathers exactly as k_push_velocity does (same weights, same
// summation order), reading LDS instead of global memory.  A particle whose
// stencil leaves the staged region (a rounding edge of the bin test) reads
// global memory.
// Fuse: each particle's position push follows its velocity push here (the
// step's next operator), so its position and momentum are read once for both;
// the bin counters then count it and key holds its bin or exchange direction,
// as after k_push_position, and the per-tile leaver counts are added here too.
// perm (a sort left pending by pn_step): sorted slot i holds record perm[i] of
// p; the kernel reads it there and writes the whole record, pushed, to slot i
// of dst.  Without perm, dst is p.
template <int Dim, int Order, bool Plain, bool Queue, bool Fuse>
__global__ void __launch_bounds__(256)
    k_push_velocity_rows(Geom g, PArr p, const int32_t* __restrict__ cellstart,
                         const int* __restrict__ off, const double* __restrict__ lim,
                         const double* __restrict__ E, const double* __restrict__ B,
                         PushConsts pc, int yrows, int nsplit,
                         int* __restrict__ wq, int gx, int nitems, Pos64 ps,
                         const int32_t* __restrict__ perm, PArr dst)
{
  constexpr int is_odd = Order % 2 == 0 ? 0 : 1;
  constexpr int S      = Order + 1;
  constexpr int T      = Order + 2;
  extern __shared__ double ft[]; // [6][TZ][T][TXv]
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
  // nsplit = 1: one block serves every species of the chunk's row (the staged
  // tile is loaded once); nsplit = Ns: one block per species
  const int c     = item_y / nsplit;
  const int sp_lo = nsplit == 1 ? 0 : item_y - c * nsplit;
  const int sp_hi = nsplit == 1 ? g.Ns : sp_lo + 1;
  const int nbx = g.bsy;
  // yrows consecutive bin rows of one z plane per block (the last block of a
  // plane may hold fewer): they are contiguous in the sorted order and share
  // one staged tile of TY = T + nr - 1 rows
  const int nby = Dim == 3 ? g.bsz / g.bsy : g.nbin / g.bsy;
  const int nyb = (nby + yrows - 1) / yrows;
  int       bz = 0, by, b0;
  if constexpr (Dim == 3) {
    bz = item_x / nyb;
    by = (item_x - bz * nyb) * yrows;
    b0 = bz * g.bsz + by * g.bsy;
  } else {
    by = item_x * yrows;
    b0 = by * g.bsy;
  }
  const int nr = min(yrows, nby - by);
  const int TY = T + nr - 1;
  bool any = false;
  for (int sp = sp_lo; sp < sp_hi; sp++) {
    const int32_t* cs = cellstart + (long)(c * g.Ns + sp) * (g.nbin + 1);
    any               = any || cs[b0 + nr * nbx] > cs[b0];
  }
  if (!any)
    break; // uniform for the block
  const int  lb  = g.Nb;
  const int  TZ  = Dim == 3 ? T : 1;
  const int  tz0 = Dim == 3 ? bz - is_odd + lb - (Order / 2) : lb;
  const int  ty0 = by - is_odd + lb - (Order / 2);
  const int  tx0 = -is_odd + lb - (Order / 2);
  const int  TXv = nbx + Order + 1;
  const int  pc1 = TZ * TY * TXv; // per component
  const long sy = g.P[2], sz = g.plane, v = g.vol;
  const int  offz = off[c * 3 + 0], offy = off[c * 3 + 1], offx = off[c * 3 + 2];
  for (int t = threadIdx.x; t < 6 * pc1; t += blockDim.x) {
    int k  = t / pc1;
    int r  = t - k * pc1;
    int tz = r / (TY * TXv);
    int r2 = r - tz * TY * TXv;
    int ty = r2 / TXv;
    int tx = r2 - ty * TXv;
    int gz = Dim == 3 ? offz + tz0 + tz : 0;
    const double* F = k < 3 ? E + k * v : B + (k - 3) * v;
    ft[t] = F[gz * sz + (offy + ty0 + ty) * sy + offx + tx0 + tx];
  }
  __syncthreads();

  double qmdt = 0; // the species' (q/m) dt/2, set per species below
  const double dx = g.del[2], dy = g.del[1], dz = g.del[0];
  const double xmin = lim[c * 6 + 4], ymin = lim[c * 6 + 2], zmin = lim[c * 6 + 0];
  const double ximin  = xmin + 0.5 * dx * is_odd;
  const double xhmin  = xmin + 0.5 * dx * is_odd - 0.5 * dx;
  const double yimin  = ymin + 0.5 * dy * is_odd;
  const double yhmin  = ymin + 0.5 * dy * is_odd - 0.5 * dy;
  const double xigrid = xmin + 0.5 * dx, xhgrid = xmin;
  const double yigrid = ymin + 0.5 * dy, yhgrid = ymin;
  const double rdx = 1 / dx, rdy = 1 / dy;

  // (Order+1)^Dim-point sum of nix::interp::interp{2,3}d_impl_sorted; base[z*bs_z + y*bs_y + x]
  auto interp3 = [&](const double* base, long bs_z, long bs_y, int z0, int y0, int x0,
                     const double* wz, const double* wy, const double* wx) {
    double rz = 0;
#pragma unroll
    for (int jz = 0; jz < S; jz++) {
      double ry = 0;
#pragma unroll
      for (int jy = 0; jy < S; jy++) {
        const double* row = base + (z0 + jz) * bs_z + (y0 + jy) * bs_y + x0;
        double        rx  = 0;
#pragma unroll
        for (int jx = 0; jx < S; jx++)
          rx += row[jx] * wx[jx];
        ry += rx * wy[jy];
      }
      rz += ry * wz[jz];
    }
    return rz * qmdt;
  };
  auto interp2 = [&](const double* base, long bs_y, int y0, int x0, const double* wy,
                     const double* wx) {
    double ry = 0;
#pragma unroll
    for (int jy = 0; jy < S; jy++) {
      const double* row = base + (y0 + jy) * bs_y + x0;
      double        rx  = 0;
#pragma unroll
      for (int jx = 0; jx < S; jx++)
        rx += row[jx] * wx[jx];
      ry += rx * wy[jy];
    }
    return ry * qmdt;
  };

  for (int sp = sp_lo; sp < sp_hi; sp++) {
  const int      seg = c * g.Ns + sp;
  const int32_t* cs  = cellstart + (long)seg * (g.nbin + 1);
  const int      p0  = cs[b0], p1 = cs[b0 + nr * nbx];
  qmdt               = pc.qmdt[sp];
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const long i     = (long)seg * g.cap + q;
    const long src   = perm ? (long)perm[i] : i;
    double     xu[6] = {p.x[src], p.y[src], p.z[src], p.ux[src], p.uy[src], p.uz[src]};
    double     wix[S], wiy[S], wiz[S], whx[S], why[S], whz[S];
    int        ix0 = digitize(xu[0], ximin, rdx);
    int        iy0 = digitize(xu[1], yimin, rdy);
    int        hx0 = digitize(xu[0], xhmin, rdx);
    int        hy0 = digitize(xu[1], yhmin, rdy);
    if constexpr (Plain) {
      shape_mc<Order>(xu[0], xigrid + static_cast<double>(ix0) * dx, rdx, wix);
      shape_mc<Order>(xu[1], yigrid + static_cast<double>(iy0) * dy, rdy, wiy);
    } else {
      int_weights<Order>(pc, 0, xu[0], xigrid + static_cast<double>(ix0) * dx, rdx, wix);
      int_weights<Order>(pc, 1, xu[1], yigrid + static_cast<double>(iy0) * dy, rdy, wiy);
    }
    shape_mc<Order>(xu[0], xhgrid + static_cast<double>(hx0) * dx, rdx, whx);
    shape_mc<Order>(xu[1], yhgrid + static_cast<double>(hy0) * dy, rdy, why);
    ix0 += lb - (Order / 2);
    iy0 += lb - (Order / 2);
    hx0 += lb - (Order / 2);
    hy0 += lb - (Order / 2);
    bool in = min(ix0, hx0) >= tx0 && max(ix0, hx0) + S <= tx0 + TXv && min(iy0, hy0) >= ty0 &&
              max(iy0, hy0) + S <= ty0 + TY;
    bool   ok = true; // false: the stencil leaves the padded arrays, no velocity push
    double eb[6];
    if constexpr (Dim == 3) {
      const double zimin  = zmin + 0.5 * dz * is_odd;
      const double zhmin  = zmin + 0.5 * dz * is_odd - 0.5 * dz;
      const double zigrid = zmin + 0.5 * dz, zhgrid = zmin;
      const double rdz    = 1 / dz;
      int          iz0    = digitize(xu[2], zimin, rdz);
      int          hz0    = digitize(xu[2], zhmin, rdz);
      if constexpr (Plain)
        shape_mc<Order>(xu[2], zigrid + static_cast<double>(iz0) * dz, rdz, wiz);
      else
        int_weights<Order>(pc, 2, xu[2], zigrid + static_cast<double>(iz0) * dz, rdz, wiz);
      shape_mc<Order>(xu[2], zhgrid + static_cast<double>(hz0) * dz, rdz, whz);
      iz0 += lb - (Order / 2);
      hz0 += lb - (Order / 2);
      in = in && min(iz0, hz0) >= tz0 && max(iz0, hz0) + S <= tz0 + T;
      if (in) {
        const long bsz = (long)TY * TXv, bsy = TXv;
        const int  Z = iz0 - tz0, HZ = hz0 - tz0, Y = iy0 - ty0, HY = hy0 - ty0;
        const int  X = ix0 - tx0, HX = hx0 - tx0;
        eb[0] = interp3(ft + 0 * pc1, bsz, bsy, Z, Y, HX, wiz, wiy, whx);
        eb[1] = interp3(ft + 1 * pc1, bsz, bsy, Z, HY, X, wiz, why, wix);
        eb[2] = interp3(ft + 2 * pc1, bsz, bsy, HZ, Y, X, whz, wiy, wix);
        eb[3] = interp3(ft + 3 * pc1, bsz, bsy, HZ, HY, X, whz, why, wix);
        eb[4] = interp3(ft + 4 * pc1, bsz, bsy, HZ, Y, HX, whz, wiy, whx);
        eb[5] = interp3(ft + 5 * pc1, bsz, bsy, Z, HY, HX, wiz, why, whx);
      } else {
        const int Z = iz0 + offz, HZ = hz0 + offz, Y = iy0 + offy, HY = hy0 + offy;
        const int X = ix0 + offx, HX = hx0 + offx;
        if (min(Z, HZ) < 0 || max(Z, HZ) > g.P[0] - S || min(Y, HY) < 0 ||
            max(Y, HY) > g.P[1] - S || min(X, HX) < 0 || max(X, HX) > g.P[2] - S) {
          ok = false; // not pushed (its record and position push still go on below)
        } else {
          eb[0] = interp3(E, sz, sy, Z, Y, HX, wiz, wiy, whx);
          eb[1] = interp3(E + v, sz, sy, Z, HY, X, wiz, why, wix);
          eb[2] = interp3(E + 2 * v, sz, sy, HZ, Y, X, whz, wiy, wix);
          eb[3] = interp3(B, sz, sy, HZ, HY, X, whz, why, wix);
          eb[4] = interp3(B + v, sz, sy, HZ, Y, HX, whz, wiy, whx);
          eb[5] = interp3(B + 2 * v, sz, sy, Z, HY, HX, wiz, why, whx);
        }
      }
    } else {
      (void)wiz;
      (void)whz;
      if (in) {
        const long bsy = TXv;
        const int  Y = iy0 - ty0, HY = hy0 - ty0, X = ix0 - tx0, HX = hx0 - tx0;
        eb[0] = interp2(ft + 0 * pc1, bsy, Y, HX, wiy, whx);
        eb[1] = interp2(ft + 1 * pc1, bsy, HY, X, why, wix);
        eb[2] = interp2(ft + 2 * pc1, bsy, Y, X, wiy, wix);
        eb[3] = interp2(ft + 3 * pc1, bsy, HY, X, why, wix);
        eb[4] = interp2(ft + 4 * pc1, bsy, Y, HX, wiy, whx);
        eb[5] = interp2(ft + 5 * pc1, bsy, HY, HX, why, whx);
      } else {
        const int Y = iy0 + offy, HY = hy0 + offy, X = ix0 + offx, HX = hx0 + offx;
        if (min(Y, HY) < 0 || max(Y, HY) > g.P[1] - S || min(X, HX) < 0 ||
            max(X, HX) > g.P[2] - S) {
          ok = false;
        } else {
          eb[0] = interp2(E, sy, Y, HX, wiy, whx);
          eb[1] = interp2(E + v, sy, HY, X, why, wix);
          eb[2] = interp2(E + 2 * v, sy, Y, X, wiy, wix);
          eb[3] = interp2(B, sy, HY, X, why, wix);
          eb[4] = interp2(B + v, sy, Y, HX, wiy, whx);
          eb[5] = interp2(B + 2 * v, sy, HY, HX, why, whx);
        }
      }
    }
    if (ok) {
      if constexpr (Plain)
        push_boris(xu[3], xu[4], xu[5], eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], g.cc);
      else
        push_momentum(pc.pusher, xu[3], xu[4], xu[5], eb, g.cc);
    }
    if (perm)
      dst.id[i] = p.id[src];
    if constexpr (Fuse) {
      int       dir;
      const int b    = position64<true>(g, dst, ps, lim, c, i, xu[0], xu[1], xu[2], xu[3],
                                        xu[4], xu[5], dir);
      const bool stay = b >= 0;
      const int  sl   = wave_slot(ps.cellcnt + (long)seg * g.nbin, b, stay);
      // the leavers of each 256-particle tile per direction, which the migration
      // ranks by (one atomic per direction and tile present in the wavefront)
      (void)wave_slot(ps.tilecnt, (seg * ps.ntile + q / 256) * 27 + dir, dir >= 0);
      if (stay) {
        ps.key[i]  = b;
        ps.slot[i] = sl;
      }
    } else {
      if (ok || perm) {
        dst.ux[i] = xu[3];
        dst.uy[i] = xu[4];
        dst.uz[i] = xu[5];
      }
      if (perm) {
        dst.x[i] = xu[0];
        dst.y[i] = xu[1];
        dst.z[i] = xu[2];
      }
    }
  }
  } // species
  } while (false);
  if constexpr (!Queue)
    return;
  __syncthreads(); // the LDS of this item is read out before the next one
  }
  if constexpr (Queue)
    wq_finish(wq);
}

// BasePosition::push (pic/engine/position.hpp:117-130) followed by count(reset)
// (Position::set_boundary, pic/pic_engine.hpp:292-303).  Stayers take a slot in
// their bin (one atomic per distinct bin in the wave); leavers record their
// exchange direction and the block's per-direction totals (one 256-particle tile
// per block).
__global__ void __launch_bounds__(256)
    k_push_position(Geom g, PArr p, double* __restrict__ xo, double* __restrict__ yo,
                    double* __restrict__ zo, const int32_t* __restrict__ np,
                    const double* __restrict__ lim, const double* __restrict__ rng, double delt,
                    double rc, int32_t* __restrict__ cellcnt, int32_t* __restrict__ key,
                    int32_t* __restrict__ slot, int32_t* __restrict__ tilecnt, int ntile,
                    int seg0)
{
  __shared__ int cnt[27];
  if (threadIdx.x < 27)
    cnt[threadIdx.x] = 0;
  __syncthreads();
  const int  seg    = seg0 + blockIdx.y;
  const long ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = ip < np[seg];
  const int  c      = seg / g.Ns;
  const long i      = (long)seg * g.cap + ip;
  int        b      = -1;
  if (active) {
    const Pos64 ps{xo, yo, zo, rng, delt, rc, cellcnt, key, slot, nullptr, 0};
    int         d;
    b = position64<false>(g, p, ps, lim, c, i, p.x[i], p.y[i], p.z[i], p.ux[i], p.uy[i],
                          p.uz[i], d);
    if (d >= 0)
      atomicAdd(&cnt[d], 1);
  }
  const bool stay = active && b >= 0;
  int        sl   = wave_slot(cellcnt + (long)seg * g.nbin, b, stay);
  if (stay) {
    slot[i] = sl;
    key[i]  = b;
  }
  __syncthreads();
  if (threadIdx.x < 27)
    tilecnt[((long)seg * ntile + blockIdx.x) * 27 + threadIdx.x] = cnt[threadIdx.x];
}

// count(reset) alone, for PicChunk::sort_particle: out-of-chunk particles are
// dropped by the sort that follows (nix/xtensor_particle.hpp:317-320).
__global__ void __launch_bounds__(256)
    k_count(Geom g, PArr p, const int32_t* __restrict__ np, const double* __restrict__ lim,
            const double* __restrict__ rng, int32_t* __restrict__ cellcnt,
            int32_t* __restrict__ key, int32_t* __restrict__ slot, int32_t* __restrict__ dseg,
            int32_t* __restrict__ didx)
{
  const int  seg    = blockIdx.y;
  const long ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = ip < np[seg];
  const int  c      = seg / g.Ns;
  const long i      = (long)seg * g.cap + ip;
  const int  b      = active ? bin_of(g, lim, rng, c, p.x[i], p.y[i], p.z[i]) : -1;
  const bool stay   = active && b >= 0;
  int        sl     = wave_slot(cellcnt + (long)seg * g.nbin, b, stay);
  if (!active)
    return;
  if (stay) {
    slot[i] = sl;
    key[i]  = b;
    dseg[i] = seg;
    didx[i] = (int32_t)ip;
  } else {
    key[i]  = -1;
    dseg[i] = -1;
  }
}

__global__ void __launch_bounds__(256) k_aos_to_soa(PArr p, long base, long n,
                                                    const double* __restrict__ aos)
{
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n)
    return;
  const double* r = aos + t * 7;
  long          i = base + t;
  p.x[i]          = r[0];
  p.y[i]          = r[1];
  p.z[i]          = r[2];
  p.ux[i]         = r[3];
  p.uy[i]         = r[4];
  p.uz[i]         = r[5];
  p.id[i]         = __double_as_longlong(r[6]);
}

__global__ void __launch_bounds__(256) k_soa_to_aos(PArr p, long base, long n,
                                                    double* __restrict__ aos)
{
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n)
    return;
  double* r = aos + t * 7;
  long    i = base + t;
  r[0]      = p.x[i];
  r[1]      = p.y[i];
  r[2]      = p.z[i];
  r[3]      = p.ux[i];
  r[4]      = p.uy[i];
  r[5]      = p.uz[i];
  r[6]      = __longlong_as_double(p.id[i]);
}

} // namespace

// The row gather's launch shape: yr bin rows per block and the LDS they stage;
// false when the row form does not apply (1D, unsorted particles, the generic
// kernels, or a stencil tile beyond the LDS of a block).
static bool vel_rows_shape(const Sim& s, int& yr, size_t& lds)
{
  const Geom& g      = s.g;
  const int   T      = g.order + 2;
  auto        lds_of = [&](int r) {
    return sizeof(double) * 6 * (g.dim == 3 ? T : 1) * (T + r - 1) * (g.bsy + g.order + 1);
  };
  yr = s.vel_rows > 0 ? s.vel_rows : 4; // 0: four rows, or fewer where LDS is short
  while (yr > 1 && lds_of(yr) > 65536)
    yr = (yr + 1) / 2;
  lds = lds_of(yr);
  return g.dim >= 2 && s.sorted && !s.generic_kernels && !s.vel_generic && lds <= 65536;
}

bool gather_takes_perm(const Sim& s)
{
  if (s.fp32)
    return gather32_takes_perm(s);
  int    yr;
  size_t lds;
  return vel_rows_shape(s, yr, lds);
}

// with_position: the row gather also runs the position push (+ count) when the
// row form applies; returns whether it did (then launch_push_position is not
// needed).
bool launch_push_velocity(Sim& s, double delt, bool with_position)
{
  PhaseTimer  pt(s, 1);
  const Geom& g = s.g;
  PushConsts  pc{};
  for (int sp = 0; sp < g.Ns; sp++)
    pc.qmdt[sp] = 0.5 * g.q[sp] / g.m[sp] * delt;
  pc.pusher = s.cfg.pusher;
  pc.wt     = s.cfg.interp;
  for (int a = 0; a < 3; a++) {
    pc.cdt[a]  = g.cc * delt / g.del[2 - a];
    pc.rcdt[a] = 1 / pc.cdt[a];
  }
  dim3         grid(nblk(g.cap, 256), g.nseg);
  const int    T     = g.order + 2;
  const int    nby   = g.dim == 3 ? g.bsz / g.bsy : g.nbin / g.bsy;
  int          yr;
  size_t       lds;
  const bool   rows = vel_rows_shape(s, yr, lds);
  const bool   fuse = rows && with_position;
  // a sort left pending by the previous step: the row gather applies it;
  // any other form needs the particles in place first
  if (s.perm_pending && !rows)
    materialize(s);
  const bool     use_perm = s.perm_pending;
  const int32_t* perm     = use_perm ? s.ss.perm : nullptr;
  const PArr     dst      = use_perm ? s.pb : s.pa;
  const int    nrow  = (g.nbin / g.bsy / nby) * ((nby + yr - 1) / yr);
  const int    nsplit = s.vel_split ? g.Ns : 1;
  const bool   plain = pc.pusher == 0 && pc.wt == 0; // Boris + MC: no option branches
  const int    vgy    = g.nchunk * nsplit, vitems = nrow * vgy;
  int*         vwq    = s.d_wq + 2 * kWqGather;
  const Pos64  ps{s.d_xo[0], s.d_xo[1], s.d_xo[2], s.d_rng,     delt,
                 1 / g.cc,  s.ss.cellcnt, s.ss.key, s.ss.slot, s.ss.tilecnt, s.ss.ntile};
  if (fuse) {
    launch_zero(s, s.ss.cellcnt, (long)g.nseg * g.nbin);
    launch_zero(s, s.ss.tilecnt, (long)g.nseg * s.ss.ntile * 27);
  }
#define PN_VEL_ROWS(D, O, PL, FU)                                                             \
  do {                                                                                        \
    if (s.vel_queue)                                                                          \
      hipLaunchKernelGGL(                                                                     \
          (k_push_velocity_rows<D, O, PL, true, FU>),                                         \
          dim3(queue_blocks(                                                                  \
              s, reinterpret_cast<const void*>(&k_push_velocity_rows<D, O, PL, true, FU>),    \
              lds, vitems)),                                                                  \
          dim3(256), lds, s.stream, g, s.pa, s.ss.cellstart, s.d_off, s.d_lim, s.d_E, s.d_B,  \
          pc, yr, nsplit, vwq, nrow, vitems, ps, perm, dst);                                  \
    else                                                                                      \
      hipLaunchKernelGGL((k_push_velocity_rows<D, O, PL, false, FU>), dim3(nrow, vgy),        \
                         dim3(256), lds, s.stream, g, s.pa, s.ss.cellstart, s.d_off, s.d_lim, \
                         s.d_E, s.d_B, pc, yr, nsplit, vwq, nrow, vitems, ps, perm, dst);     \
  } while (0)
#define PN_VEL(D, O)                                                                          \
  do {                                                                                        \
    if (rows && plain && fuse)                                                                \
      PN_VEL_ROWS(D, O, true, true);                                                          \
    else if (rows && plain)                                                                   \
      PN_VEL_ROWS(D, O, true, false);                                                         \
    else if (rows && fuse)                                                                    \
      PN_VEL_ROWS(D, O, false, true);                                                         \
    else if (rows)                                                                            \
      PN_VEL_ROWS(D, O, false, false);                                                        \
    else                                                                                      \
      hipLaunchKernelGGL((k_push_velocity<D, O>), grid, dim3(256), 0, s.stream, g, s.pa,      \
                         s.d_np, s.d_off, s.d_lim, s.d_E, s.d_B, pc);                         \
  } while (0)
  if (g.dim == 3) {
    switch (g.order) {
    case 1: PN_VEL(3, 1); break;
    case 2: PN_VEL(3, 2); break;
    case 3: PN_VEL(3, 3); break;
    default: PN_VEL(3, 4); break;
    }
  } else if (g.dim == 1) {
    switch (g.order) {
    case 1: hipLaunchKernelGGL((k_push_velocity<1, 1>), grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_E, s.d_B, pc); break;
    case 2: hipLaunchKernelGGL((k_push_velocity<1, 2>), grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_E, s.d_B, pc); break;
    case 3: hipLaunchKernelGGL((k_push_velocity<1, 3>), grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_E, s.d_B, pc); break;
    default: hipLaunchKernelGGL((k_push_velocity<1, 4>), grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_off, s.d_lim, s.d_E, s.d_B, pc); break;
    }
  } else {
    switch (g.order) {
    case 1: PN_VEL(2, 1); break;
    case 2: PN_VEL(2, 2); break;
    case 3: PN_VEL(2, 3); break;
    default: PN_VEL(2, 4); break;
    }
  }
#undef PN_VEL
#undef PN_VEL_ROWS
  if (use_perm) {
    std::swap(s.pa, s.pb);
    s.perm_pending = false;
  }
  if (fuse) {
    s.counted = true;
    s.sorted  = false;
  }
  return fuse;
}

// chunk >= 0: the segments of that chunk only (their bin counters, keys, slots
// and tile counts; the other segments' keep what their own push left)
void launch_push_position(Sim& s, double delt, int chunk)
{
  materialize(s);
  PhaseTimer  pt(s, 2);
  const Geom& g    = s.g;
  const int   seg0 = chunk < 0 ? 0 : chunk * g.Ns;
  const int   nsg  = chunk < 0 ? g.nseg : g.Ns;
  launch_zero(s, s.ss.cellcnt + (long)seg0 * g.nbin, (long)nsg * g.nbin);
  dim3 grid(s.ss.ntile, nsg);
  hipLaunchKernelGGL(k_push_position, grid, dim3(256), 0, s.stream, g, s.pa, s.d_xo[0],
                     s.d_xo[1], s.d_xo[2], s.d_np, s.d_lim, s.d_rng, delt, 1 / g.cc,
                     s.ss.cellcnt, s.ss.key, s.ss.slot, s.ss.tilecnt, s.ss.ntile, seg0);
  s.counted = true;
  s.sorted  = false;
}

void launch_count(Sim& s)
{
  const Geom& g = s.g;
  launch_zero(s, s.ss.cellcnt, (long)g.nseg * g.nbin);
  dim3 grid(s.ss.ntile, g.nseg);
  hipLaunchKernelGGL(k_count, grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_lim, s.d_rng,
                     s.ss.cellcnt, s.ss.key, s.ss.slot, s.ss.dseg, s.ss.didx);
}

void launch_aos_to_soa(Sim& s, int seg, int64_t np)
{
  if (np <= 0)
    return;
  hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(np, 256)), dim3(256), 0, s.stream, s.pa,
                     (long)seg * s.g.cap, (long)np, s.d_stage);
}

void launch_soa_to_aos(Sim& s, int seg, int64_t np)
{
  if (np <= 0)
    return;
  hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(np, 256)), dim3(256), 0, s.stream, s.pa,
                     (long)seg * s.g.cap, (long)np, s.d_stage);
}

} // namespace picnix
