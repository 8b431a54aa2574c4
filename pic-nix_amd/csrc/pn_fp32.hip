// pn_fp32.hip -- the FP32 form of the step (precision = "fp32").
//
// Particles are stored cell-relative: the cell (the chunk-local sort bin) and
// the offset inside it in cell units, r in [0, 1) per axis after every sort
// (struct P32).  Fields E, B, J are FP32 on the same padded global grid as the
// FP64 form.  The arithmetic follows the FP64 kernels operation by operation
// (gather + pusher, position push, Esirkepov, rho), in FP32, with three
// differences that the representation brings:
//   * shape weights take delta = r - offset directly (no digitize of an
//     absolute position: the particle's cell is known), so weights keep their
//     full FP32 precision however large the domain is;
//   * the position push adds u dt / (gamma dx) to r; a particle that leaves its
//     cell changes its cell index, and one that leaves its chunk keeps r and
//     gets the neighbour's cell index (the periodic wrap is exact);
//   * per-cell current sums run in FP32 registers and the block's ring tile
//     accumulates them as compensated FP32 pairs in LDS (hi + lo, every add's
//     rounding error kept in lo) before the FP32 global add.
// FP64 stays for the chunk origins (positions are rebuilt as
// xmin + (cell + r) dx in FP64 at the C-ABI boundary) and for the energy and
// charge reductions, which read FP64 mirrors of the FP32 fields.
#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

struct Cell3 {
  int x, y, z;
};

__device__ __forceinline__ Cell3 cell_of(const Geom& g, int bin)
{
  Cell3 c;
  c.z       = bin / g.bsz;
  const int r = bin - c.z * g.bsz;
  c.y       = r / g.bsy;
  c.x       = r - c.y * g.bsy;
  return c;
}

__device__ __forceinline__ int bin_of3(const Geom& g, int x, int y, int z)
{
  return z * g.bsz + y * g.bsy + x;
}

// Integer-grid (cell-centred) and half-grid weights of one axis from the
// cell-relative offset r; `first` is the stencil's first cell relative to the
// particle's cell.  The same deltas as digitize + shape_mc on the absolute
// position (pic/engine/velocity.hpp:223-359): integer grid
// delta = x/dx - 1/2 - floor(x/dx - odd/2), half grid delta = x/dx - floor(x/dx + 1/2 - odd/2).
template <int Order>
__device__ __forceinline__ float int_delta(float r, int& first)
{
  constexpr float ho = Order % 2 == 1 ? 0.5f : 0.0f;
  const float     f  = floorf(r - ho);
  first              = (int)f;
  return r - 0.5f - f;
}

template <int Order>
__device__ __forceinline__ float half_delta(float r, int& first)
{
  constexpr float ho = Order % 2 == 1 ? 0.5f : 0.0f;
  const float     f  = floorf(r + 0.5f - ho);
  first              = (int)f;
  return r - f;
}

// ---------------------------------------------------------------------------
// conversions at the C-ABI boundary (FP64 absolute <-> FP32 cell-relative)
// ---------------------------------------------------------------------------

// FP64 SoA records of segments -> FP32 cell-relative (every particle of every
// segment below np).  A particle outside its chunk gets cell -1 (dropped by
// the next count, as the reference's sort drops bin Ng).
__global__ void __launch_bounds__(256)
    k32_from_f64(Geom g, PArr p, P32 q, const int32_t* __restrict__ np,
                 const double* __restrict__ lim, int seg0, int nseg)
{
  const int  seg = seg0 + blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.y >= (unsigned)nseg || ip >= np[seg])
    return;
  const int    c    = seg / g.Ns;
  const long   i    = (long)seg * g.cap + ip;
  const double x[3] = {p.x[i], p.y[i], p.z[i]};
  float        r[3];
  int          cl[3] = {0, 0, 0};
  bool         in    = true;
  for (int a = 0; a < 3; a++) {
    const int ax = 2 - a; // Geom tables are (z, y, x)
    if (!g.has[ax]) {
      r[a] = (float)x[a]; // an absent axis keeps its absolute coordinate
      continue;
    }
    const double u  = (x[a] - lim[c * 6 + 2 * ax]) / g.del[ax];
    const double fl = floor(u);
    float        rf = (float)(u - fl);
    const int    ci = (int)fl;
    // u - floor(u) < 1 can round up to 1.0f: keep the FP64 cell (the reference's
    // bin) with the largest offset below 1, as the sort does
    if (rf >= 1.0f)
      rf = 0.99999994f;
    cl[a] = ci;
    r[a]  = rf;
    in    = in && ci >= 0 && ci < g.dims[ax];
  }
  q.rx[i]   = r[0];
  q.ry[i]   = r[1];
  q.rz[i]   = r[2];
  q.ux[i]   = (float)p.ux[i];
  q.uy[i]   = (float)p.uy[i];
  q.uz[i]   = (float)p.uz[i];
  q.id[i]   = p.id[i];
  q.cell[i] = in ? bin_of3(g, cl[0], cl[1], cl[2]) : -1;
}

// FP32 -> FP64 SoA (x = xmin + (cell + r) dx in FP64); r may be unwrapped.
__global__ void __launch_bounds__(256)
    k32_to_f64(Geom g, P32 q, PArr p, const float* __restrict__ ro0,
               const float* __restrict__ ro1, const float* __restrict__ ro2,
               double* __restrict__ xo0, double* __restrict__ xo1, double* __restrict__ xo2,
               const int32_t* __restrict__ np, const double* __restrict__ lim, int seg0)
{
  const int  seg = seg0 + blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int    c      = seg / g.Ns;
  const long   i      = (long)seg * g.cap + ip;
  const int    bin    = q.cell[i];
  const Cell3  cl     = bin >= 0 ? cell_of(g, bin) : Cell3{0, 0, 0};
  const int    ci[3]  = {cl.x, cl.y, cl.z};
  const float  r[3]   = {q.rx[i], q.ry[i], q.rz[i]};
  const float* ro[3]  = {ro0, ro1, ro2};
  double*      out[3] = {p.x, p.y, p.z};
  double*      old[3] = {xo0, xo1, xo2};
  for (int a = 0; a < 3; a++) {
    const int ax = 2 - a;
    if (!g.has[ax]) {
      out[a][i] = (double)r[a];
      if (old[a])
        old[a][i] = (double)ro[a][i];
      continue;
    }
    const double x0 = lim[c * 6 + 2 * ax];
    out[a][i]       = x0 + ((double)ci[a] + (double)r[a]) * g.del[ax];
    if (old[a])
      old[a][i] = x0 + ((double)ci[a] + (double)ro[a][i]) * g.del[ax];
  }
  p.ux[i] = (double)q.ux[i];
  p.uy[i] = (double)q.uy[i];
  p.uz[i] = (double)q.uz[i];
  p.id[i] = q.id[i];
}

template <typename S, typename D>
__global__ void __launch_bounds__(256) k32_convert(const S* __restrict__ a, D* __restrict__ b,
                                                   long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    b[i] = (D)a[i];
}

// ---------------------------------------------------------------------------
// gather + pusher (pic/engine/velocity.hpp:101-135,223-450) in FP32
// ---------------------------------------------------------------------------

struct Push32 {
  float qmdt[kMaxSpecies];
  float cdt[3], rcdt[3]; // c dt / d(x, y, z), for WT
  int   pusher;
  int   wt;
};

template <int Order>
__device__ __forceinline__ void int_w(const Push32& pc, int axis, float r, int& first, float* w)
{
  const float d = int_delta<Order>(r, first);
  if (pc.wt)
    shape_wt<Order, float>(d, 0.0f, 1.0f, pc.cdt[axis], pc.rcdt[axis], w);
  else
    shape_mc<Order, float>(d, 0.0f, 1.0f, w);
}

// Destination of a particle of segment seg whose cell-relative position is r in
// its old cell `bin` (see k32_push_position): segment ds, bin b, key k; ds < 0
// when it is dropped.
__device__ __forceinline__ void dest32(const Geom& g, int seg, int bin, float rx, float ry, float rz,
                                       const int* __restrict__ nb, int& ds, int& b, int& k)
{
  ds = -1;
  b  = -1;
  k  = -1;
  if (bin < 0)
    return;
  const Cell3 cl = cell_of(g, bin);
  int         cx = cl.x + (g.has[2] ? (int)floorf(rx) : 0);
  int         cy = cl.y + (g.has[1] ? (int)floorf(ry) : 0);
  int         cz = cl.z + (g.has[0] ? (int)floorf(rz) : 0);
  const int   dx = (cx >= g.dims[2]) - (cx < 0) + 1;
  const int   dy = (cy >= g.dims[1]) - (cy < 0) + 1;
  const int   dz = (cz >= g.dims[0]) - (cz < 0) + 1;
  const int   d  = 9 * dz + 3 * dy + dx;
  int         R  = seg / g.Ns;
  if (d != 13) {
    R = nb[R * 27 + d];
    cx += (cx < 0) * g.dims[2] - (cx >= g.dims[2]) * g.dims[2];
    cy += (cy < 0) * g.dims[1] - (cy >= g.dims[1]) * g.dims[1];
    cz += (cz < 0) * g.dims[0] - (cz >= g.dims[0]) * g.dims[0];
  }
  if (R >= 0 && cx >= 0 && cx < g.dims[2] && cy >= 0 && cy < g.dims[1] && cz >= 0 &&
      cz < g.dims[0]) {
    ds = R * g.Ns + (seg - (seg / g.Ns) * g.Ns);
    b  = bin_of3(g, cx, cy, cz);
    k  = d == 13 ? b : -2 - (d * g.nbin + b);
  }
}

// Arguments of the FP32 position push (k32_push_position, and the row gather
// when it runs the position push of each particle it has pushed).
struct Pos32 {
  float*         ro0;
  float*         ro1;
  float*         ro2;
  float          delt, rc, rdx, rdy, rdz;
  const int*     nb;
  const int*     off;
  int32_t*       cellcnt;
  int32_t*       key;
  int32_t*       slot;
  int            on; // the row gather: 1 runs the position push after the momentum push
};

// r_old -> ro; r += u dt / (gamma d) (unwrapped, relative to the old cell).
// Then the particle's destination for the sort: its own segment and new bin, or
// for a leaver the neighbour chunk in its direction (chunk table nb) and the bin
// there -- the periodic wrap is a shift of the cell index, so r stays as it is
// (set_boundary_periodic, nix/xtensor_particle.hpp:359-376, and the migration,
// nix/xtensor_halo3d.hpp:192-499).  Every particle takes a slot in its
// destination bin here, so the sort needs no migration pass.  key: the new bin
// (>= 0) for a particle that stays in its chunk, -2 - (d * nbin + bin) for one
// that moves to the neighbour in direction d, -1 for one that is dropped.
// With a non-periodic x (the shock problem) a particle that crossed x = 0 is
// reflected (MainChunk::set_boundary_particle, example/shock/main.cpp:402-416)
// and one past x = Lx is marked kKeyReplace for its redraw.
// One particle i of segment seg with position r in its old cell `bin` and the
// momentum u of this step; WriteU: u is written here (else only when reflected).
// Called by every active lane of the wave that holds a particle (wave_slot).
template <bool WriteU>
__device__ __forceinline__ void position32(const Geom& g, const P32& p, const Pos32& ps, long i,
                                           int seg, float rx, float ry, float rz, float ux,
                                           float uy, float uz, int bin)
{
  ps.ro0[i]      = rx;
  ps.ro1[i]      = ry;
  ps.ro2[i]      = rz;
  const float gm = lorentz_factor(ux, uy, uz, ps.rc);
  const float dt = ps.delt / gm;
  rx += g.has[2] ? ux * dt * ps.rdx : ux * dt;
  ry += g.has[1] ? uy * dt * ps.rdy : uy * dt;
  rz += g.has[0] ? uz * dt * ps.rdz : uz * dt;
  bool replace = false;
  bool flip    = false;
  if (g.xwall && bin >= 0) {
    // global x cell of the old cell; |dx| < 1 cell, so only the end cells can cross
    const int gx = ps.off[(seg / g.Ns) * 3 + 2] + cell_of(g, bin).x;
    if (gx == 0 && rx < 0.0f) {
      rx   = -rx;
      flip = true;
    } else if (gx == g.N[2] - 1 && rx >= 1.0f) {
      replace = true;
    }
  }
  if (WriteU || flip) {
    p.ux[i] = flip ? -ux : ux;
    p.uy[i] = flip ? -uy : uy;
    p.uz[i] = flip ? -uz : uz;
  }
  p.rx[i] = rx;
  p.ry[i] = ry;
  p.rz[i] = rz;
  int ds = -1, b = -1, k = -1;
  if (replace)
    k = kKeyReplace;
  else
    dest32(g, seg, bin, rx, ry, rz, ps.nb, ds, b, k);
  const bool want = ds >= 0;
  const int  sl   = wave_slot(ps.cellcnt, want ? ds * g.nbin + b : 0, want);
  ps.key[i]       = k;
  if (want)
    ps.slot[i] = sl;
}

// One block per (bin row, segment) of sorted particles: the row's E and B
// stencil region is staged in LDS, as in the FP64 row kernel; one thread per
// particle, (Order+1)^Dim-point gathers with x innermost.
// perm (a sort left pending by pn_step): sorted slot i holds record perm[i] of
// p, with its offset relative to its old cell; the kernel reads it there, finds
// its bin (as the position push did, dest32) and its offset in [0, 1), and
// writes the whole record, pushed, to slot i of dst.  Without perm, dst is p.
template <int Dim, int Order, bool Queue, bool Fuse>
__global__ void __launch_bounds__(256)
    k32_push_velocity_rows(Geom g, P32 p, const int32_t* __restrict__ cellstart,
                           const int* __restrict__ off, const float* __restrict__ E,
                           const float* __restrict__ B, Push32 pc, Pos32 ps, int yrows,
                           int nsplit, int* __restrict__ wq, int gx, int nitems,
                           const int32_t* __restrict__ perm, P32 dst)
{
  constexpr int is_odd = Order % 2;
  constexpr int S      = Order + 1;
  constexpr int T      = Order + 2;
  extern __shared__ float ft32[]; // [6][TZ][T][TXv]
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
  const int  pc1 = TZ * TY * TXv;
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
    const float* F = k < 3 ? E + k * v : B + (k - 3) * v;
    ft32[t] = F[gz * sz + (offy + ty0 + ty) * sy + offx + tx0 + tx];
  }
  __syncthreads();

  float qmdt = 0; // the species' (q/m) dt/2, set per species below
  auto interp = [&](const float* base, long bs_z, long bs_y, int z0, int y0, int x0,
                    const float* wz, const float* wy, const float* wx, int nz) {
    float rz = 0;
    for (int jz = 0; jz < nz; jz++) {
      float ry = 0;
#pragma unroll
      for (int jy = 0; jy < S; jy++) {
        const float* row = base + (z0 + jz) * bs_z + (y0 + jy) * bs_y + x0;
        float        rx  = 0;
#pragma unroll
        for (int jx = 0; jx < S; jx++)
          rx += row[jx] * wx[jx];
        ry += rx * wy[jy];
      }
      rz += ry * wz[jz];
    }
    return rz * qmdt;
  };

  for (int sp = sp_lo; sp < sp_hi; sp++) {
  const int      seg = c * g.Ns + sp;
  const int32_t* cs  = cellstart + (long)seg * (g.nbin + 1);
  const int      p0  = cs[b0], p1 = cs[b0 + nr * nbx];
  qmdt               = pc.qmdt[sp];
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const long i   = (long)seg * g.cap + q;
    long       src = i;
    int        bin;
    float      r[3];
    if (perm) {
      src           = perm[i];
      const float x = p.rx[src], y = p.ry[src], z = p.rz[src];
      int         ds, b, k;
      dest32(g, (int)(src / g.cap), p.cell[src], x, y, z, ps.nb, ds, b, k);
      bin  = b;
      r[0] = g.has[2] ? frac01(x) : x;
      r[1] = g.has[1] ? frac01(y) : y;
      r[2] = g.has[0] ? frac01(z) : z;
    } else {
      bin  = p.cell[i];
      r[0] = p.rx[i];
      r[1] = p.ry[i];
      r[2] = p.rz[i];
    }
    const Cell3 cl = cell_of(g, bin);
    float       wix[S], wiy[S], wiz[S], whx[S], why[S], whz[S];
    int         ix0, iy0, iz0 = 0, hx0, hy0, hz0 = 0;
    int_w<Order>(pc, 0, r[0], ix0, wix);
    int_w<Order>(pc, 1, r[1], iy0, wiy);
    shape_mc<Order, float>(half_delta<Order>(r[0], hx0), 0.0f, 1.0f, whx);
    shape_mc<Order, float>(half_delta<Order>(r[1], hy0), 0.0f, 1.0f, why);
    // first stencil node relative to the staged tile
    const int X  = cl.x + ix0 + lb - (Order / 2) - tx0;
    const int HX = cl.x + hx0 + lb - (Order / 2) - tx0;
    const int Y  = cl.y + iy0 + lb - (Order / 2) - ty0;
    const int HY = cl.y + hy0 + lb - (Order / 2) - ty0;
    int       Z = 0, HZ = 0;
    if constexpr (Dim == 3) {
      int_w<Order>(pc, 2, r[2], iz0, wiz);
      shape_mc<Order, float>(half_delta<Order>(r[2], hz0), 0.0f, 1.0f, whz);
      Z  = cl.z + iz0 + lb - (Order / 2) - tz0;
      HZ = cl.z + hz0 + lb - (Order / 2) - tz0;
    } else {
      wiz[0] = whz[0] = 1.0f;
    }
    // sorted particles sit in the row's bins, so the stencils are inside the tile
    const bool in = min(X, HX) >= 0 && max(X, HX) + S <= TXv && min(Y, HY) >= 0 &&
                    max(Y, HY) + S <= TY && min(Z, HZ) >= 0 && max(Z, HZ) + (Dim == 3 ? S : 1) <= TZ;
    float ux = p.ux[src], uy = p.uy[src], uz = p.uz[src];
    if (in) {
      const long bsz = (long)TY * TXv, bsy = TXv;
      const int  nz  = Dim == 3 ? S : 1;
      float      eb[6];
      eb[0] = interp(ft32 + 0 * pc1, bsz, bsy, Z, Y, HX, wiz, wiy, whx, nz);
      eb[1] = interp(ft32 + 1 * pc1, bsz, bsy, Z, HY, X, wiz, why, wix, nz);
      eb[2] = interp(ft32 + 2 * pc1, bsz, bsy, HZ, Y, X, whz, wiy, wix, nz);
      eb[3] = interp(ft32 + 3 * pc1, bsz, bsy, HZ, HY, X, whz, why, wix, nz);
      eb[4] = interp(ft32 + 4 * pc1, bsz, bsy, HZ, Y, HX, whz, wiy, whx, nz);
      eb[5] = interp(ft32 + 5 * pc1, bsz, bsy, Z, HY, HX, wiz, why, whx, nz);
      push_momentum<float>(pc.pusher, ux, uy, uz, eb, (float)g.cc);
    }
    if (perm) {
      dst.id[i]   = p.id[src];
      dst.cell[i] = bin;
    }
    if constexpr (Fuse) {
      // the position push of this particle with its new momentum (the step's
      // next operator), so r and u are read once for both
      position32<true>(g, dst, ps, i, seg, r[0], r[1], r[2], ux, uy, uz, bin);
    } else {
      if (in || perm) {
        dst.ux[i] = ux;
        dst.uy[i] = uy;
        dst.uz[i] = uz;
      }
      if (perm) {
        dst.rx[i] = r[0];
        dst.ry[i] = r[1];
        dst.rz[i] = r[2];
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


// One thread per particle, straight from global memory (unsorted particles).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k32_push_velocity(Geom g, P32 p, const int32_t* __restrict__ np, const int* __restrict__ off,
                      const float* __restrict__ E, const float* __restrict__ B, Push32 pc)
{
  constexpr int S   = Order + 1;
  const int     seg = blockIdx.y;
  const long    ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int  c   = seg / g.Ns;
  const int  sp  = seg - c * g.Ns;
  const long i   = (long)seg * g.cap + ip;
  const int  bin = p.cell[i];
  if (bin < 0)
    return;
  const Cell3 cl   = cell_of(g, bin);
  const float r[3] = {p.rx[i], p.ry[i], p.rz[i]};
  float       wix[S], wiy[S], wiz[S], whx[S], why[S], whz[S];
  int         ix0, iy0, iz0 = 0, hx0, hy0, hz0 = 0;
  int_w<Order>(pc, 0, r[0], ix0, wix);
  if constexpr (Dim == 1) {
    // the x line of the grid (weights1d + push_scalar1d, pic/engine/velocity.hpp:
    // 223-251,362-388): E_x, B_y, B_z on the half-integer nodes, the rest on the
    // integer ones, as the FP64 1D gather
    shape_mc<Order, float>(half_delta<Order>(r[0], hx0), 0.0f, 1.0f, whx);
    const int X  = off[c * 3 + 2] + cl.x + ix0 + g.Nb - (Order / 2);
    const int HX = off[c * 3 + 2] + cl.x + hx0 + g.Nb - (Order / 2);
    if (min(X, HX) < 0 || max(X, HX) > g.P[2] - S)
      return;
    const long  v    = g.vol;
    const float qmdt = pc.qmdt[sp];
    auto        line = [&](const float* F, int x0, const float* w) {
      float rr = 0;
      for (int jx = 0; jx < S; jx++)
        rr += F[x0 + jx] * w[jx];
      return rr * qmdt;
    };
    float eb[6];
    eb[0] = line(E, HX, whx);
    eb[1] = line(E + v, X, wix);
    eb[2] = line(E + 2 * v, X, wix);
    eb[3] = line(B, X, wix);
    eb[4] = line(B + v, HX, whx);
    eb[5] = line(B + 2 * v, HX, whx);
    float ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
    push_momentum<float>(pc.pusher, ux, uy, uz, eb, (float)g.cc);
    p.ux[i] = ux;
    p.uy[i] = uy;
    p.uz[i] = uz;
    (void)wiy;
    (void)wiz;
    (void)why;
    (void)whz;
    (void)iy0;
    (void)iz0;
    (void)hy0;
    (void)hz0;
    return;
  }
  int_w<Order>(pc, 1, r[1], iy0, wiy);
  shape_mc<Order, float>(half_delta<Order>(r[0], hx0), 0.0f, 1.0f, whx);
  shape_mc<Order, float>(half_delta<Order>(r[1], hy0), 0.0f, 1.0f, why);
  const int lb = g.Nb;
  const int X  = off[c * 3 + 2] + cl.x + ix0 + lb - (Order / 2);
  const int HX = off[c * 3 + 2] + cl.x + hx0 + lb - (Order / 2);
  const int Y  = off[c * 3 + 1] + cl.y + iy0 + lb - (Order / 2);
  const int HY = off[c * 3 + 1] + cl.y + hy0 + lb - (Order / 2);
  int       Z = 0, HZ = 0;
  if constexpr (Dim == 3) {
    int_w<Order>(pc, 2, r[2], iz0, wiz);
    shape_mc<Order, float>(half_delta<Order>(r[2], hz0), 0.0f, 1.0f, whz);
    Z  = off[c * 3 + 0] + cl.z + iz0 + lb - (Order / 2);
    HZ = off[c * 3 + 0] + cl.z + hz0 + lb - (Order / 2);
    if (min(Z, HZ) < 0 || max(Z, HZ) > g.P[0] - S)
      return;
  } else {
    wiz[0] = whz[0] = 1.0f;
  }
  if (min(Y, HY) < 0 || max(Y, HY) > g.P[1] - S || min(X, HX) < 0 || max(X, HX) > g.P[2] - S)
    return;
  const long  sy = g.P[2], sz = g.plane, v = g.vol;
  const int   nz   = Dim == 3 ? S : 1;
  const float qmdt = pc.qmdt[sp];
  auto interp = [&](const float* base, int z0, int y0, int x0, const float* wz, const float* wy,
                    const float* wx) {
    float rz = 0;
    for (int jz = 0; jz < nz; jz++) {
      float ry = 0;
      for (int jy = 0; jy < S; jy++) {
        const float* row = base + (z0 + jz) * sz + (y0 + jy) * sy + x0;
        float        rx  = 0;
        for (int jx = 0; jx < S; jx++)
          rx += row[jx] * wx[jx];
        ry += rx * wy[jy];
      }
      rz += ry * wz[jz];
    }
    return rz * qmdt;
  };
  float eb[6];
  eb[0] = interp(E, Z, Y, HX, wiz, wiy, whx);
  eb[1] = interp(E + v, Z, HY, X, wiz, why, wix);
  eb[2] = interp(E + 2 * v, HZ, Y, X, whz, wiy, wix);
  eb[3] = interp(B, HZ, HY, X, whz, why, wix);
  eb[4] = interp(B + v, HZ, Y, HX, whz, wiy, whx);
  eb[5] = interp(B + 2 * v, Z, HY, HX, wiz, why, whx);
  float ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
  push_momentum<float>(pc.pusher, ux, uy, uz, eb, (float)g.cc);
  p.ux[i] = ux;
  p.uy[i] = uy;
  p.uz[i] = uz;
}

// ---------------------------------------------------------------------------
// position push + count (pic/engine/position.hpp:117-130,
// nix/xtensor_particle.hpp:324-357) in FP32
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
    k32_push_position(Geom g, P32 p, Pos32 ps, const int32_t* __restrict__ np)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long i = (long)seg * g.cap + ip;
  position32<false>(g, p, ps, i, seg, p.rx[i], p.ry[i], p.rz[i], p.ux[i], p.uy[i], p.uz[i],
                    p.cell[i]);
}

// The shock problem's replacements in the FP32 form: x, ux, uy, uz of every
// particle listed by k_wall_list, from the host's draws (absolute x), written
// relative to the particle's old cell; then its destination and slot.
__global__ void __launch_bounds__(256) k32_wall_apply(Geom g, P32 p, const int32_t* __restrict__ useg,
                                                      const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ count,
                                                      const int64_t* __restrict__ voff,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ lim,
                                                      const int* __restrict__ nb,
                                                      int32_t* __restrict__ cellcnt,
                                                      int32_t* __restrict__ key,
                                                      int32_t* __restrict__ slot)
{
  const int k   = blockIdx.x;
  const int seg = useg[k];
  const int c   = seg / g.Ns;
  const int n   = count[k];
  for (int t = (int)threadIdx.x; t < n; t += blockDim.x) {
    const long    i   = (long)seg * g.cap + list[(long)k * g.cap + t];
    const double* v   = vals + voff[k] + 4 * (long)t;
    const int     bin = p.cell[i];
    const Cell3   cl  = cell_of(g, bin);
    const float   rx  = (float)((v[0] - lim[c * 6 + 4]) / g.del[2] - cl.x);
    p.rx[i]           = rx;
    p.ux[i]           = (float)v[1];
    p.uy[i]           = (float)v[2];
    p.uz[i]           = (float)v[3];
    int ds, b, kk;
    dest32(g, seg, bin, rx, p.ry[i], p.rz[i], nb, ds, b, kk);
    key[i] = kk;
    if (ds >= 0)
      slot[i] = atomicAdd(&cellcnt[(long)ds * g.nbin + b], 1);
  }
}

// The shock problem's injected particles in the FP32 form: appended after each
// upper segment's particles as cell + offset in [0, 1) (old = new), each with its
// destination and slot; the count grows, or the capacity flag is raised.
__global__ void __launch_bounds__(256) k32_inject_append(Geom g, P32 p, float* __restrict__ ro0,
                                                         float* __restrict__ ro1,
                                                         float* __restrict__ ro2,
                                                         const int32_t* __restrict__ useg,
                                                         const int32_t* __restrict__ ninj,
                                                         const int64_t* __restrict__ voff,
                                                         const double* __restrict__ vals,
                                                         const double* __restrict__ lim,
                                                         const int* __restrict__ nb,
                                                         int32_t* __restrict__ np,
                                                         int32_t* __restrict__ cellcnt,
                                                         int32_t* __restrict__ key,
                                                         int32_t* __restrict__ slot,
                                                         int32_t* __restrict__ errflag)
{
  __shared__ int n0s;
  const int k   = blockIdx.x;
  const int seg = useg[k];
  const int c   = seg / g.Ns;
  const int n   = ninj[k];
  if (threadIdx.x == 0)
    n0s = np[seg];
  __syncthreads();
  const int n0 = n0s;
  if ((long)n0 + n > g.cap) {
    if (threadIdx.x == 0)
      atomicOr(&errflag[0], 1);
    return;
  }
  for (int t = (int)threadIdx.x; t < n; t += blockDim.x) {
    const long    i = (long)seg * g.cap + n0 + t;
    const double* v = vals + voff[k] + 7 * (long)t;
    const double  X = (v[0] - lim[c * 6 + 4]) / g.del[2];
    const double  Y = g.has[1] ? (v[1] - lim[c * 6 + 2]) / g.del[1] : 0.0;
    const double  Z = g.has[0] ? (v[2] - lim[c * 6 + 0]) / g.del[0] : 0.0;
    int cx = g.has[2] ? (int)floor(X) : 0, cy = g.has[1] ? (int)floor(Y) : 0,
        cz = g.has[0] ? (int)floor(Z) : 0;
    cx = min(max(cx, 0), g.dims[2] - 1);
    cy = min(max(cy, 0), g.dims[1] - 1);
    cz = min(max(cz, 0), g.dims[0] - 1);
    const float rx = (float)(X - cx);
    const float ry = g.has[1] ? (float)(Y - cy) : (float)v[1];
    const float rz = g.has[0] ? (float)(Z - cz) : (float)v[2];
    const int   bin = bin_of3(g, cx, cy, cz);
    p.rx[i]   = rx;
    p.ry[i]   = ry;
    p.rz[i]   = rz;
    ro0[i]    = rx;
    ro1[i]    = ry;
    ro2[i]    = rz;
    p.ux[i]   = (float)v[3];
    p.uy[i]   = (float)v[4];
    p.uz[i]   = (float)v[5];
    p.id[i]   = __double_as_longlong(v[6]);
    p.cell[i] = bin;
    int ds, b, kk;
    dest32(g, seg, bin, rx, ry, rz, nb, ds, b, kk);
    key[i] = kk;
    if (ds >= 0)
      slot[i] = atomicAdd(&cellcnt[(long)ds * g.nbin + b], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0)
    np[seg] = n0 + n;
}

// ---------------------------------------------------------------------------
// Esirkepov deposit (nix/esirkepov.hpp:76-337) in FP32
// ---------------------------------------------------------------------------

struct Dep32 {
  float q[kMaxSpecies];
  float delt, dxdt, dydt, dzdt;
};

// Old and new weights of one axis in the old cell's frame (even orders:
// i0 = the cell, shift = floor(r_new); odd orders: i0 = cell + floor(r_old - 1/2)).
// s0, ds have Order+3 entries; returns the first node relative to the cell.
template <int Order>
__device__ __forceinline__ int esk_axis(float ro, float rn, float* s0, float* ds)
{
  constexpr int   N  = Order + 3;
  constexpr float ho = Order % 2 == 1 ? 0.5f : 0.0f;
  float           s1[N];
#pragma unroll
  for (int l = 0; l < N; l++)
    s0[l] = s1[l] = 0.0f;
  const float f0 = floorf(ro - ho);
  const float f1 = floorf(rn - ho);
  shape_mc<Order, float>(ro - 0.5f - f0, 0.0f, 1.0f, &s0[1]);
  shape_mc<Order, float>(rn - 0.5f - f1, 0.0f, 1.0f, &s1[1]);
  const int sh = (int)(f1 - f0);
  if (sh < 0) {
#pragma unroll
    for (int ii = 0; ii < Order + 2; ii++)
      s1[ii] = s1[ii + 1];
  } else if (sh > 0) {
#pragma unroll
    for (int ii = Order + 2; ii > 0; ii--)
      s1[ii] = s1[ii - 1];
  }
#pragma unroll
  for (int l = 0; l < N; l++)
    ds[l] = s1[l] - s0[l];
  return (int)f0 - (Order / 2) - 1;
}

// One thread per particle (odd orders, 2D, unsorted particles): the full
// (Order+3)^3 (3D) or (Order+3)^2 (2D) tile of J with global FP32 atomics.
// rho is deposited separately (k32_rho_global / k32_rho_march).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k32_deposit_global(Geom g, P32 p, const float* __restrict__ ro0,
                       const float* __restrict__ ro1, const float* __restrict__ ro2,
                       const int32_t* __restrict__ np, const int* __restrict__ off,
                       float* __restrict__ J, Dep32 dc)
{
  constexpr int N   = Order + 3;
  const int     seg = blockIdx.y;
  const long    ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int  c   = seg / g.Ns;
  const int  sp  = seg - c * g.Ns;
  const long i   = (long)seg * g.cap + ip;
  const int  bin = p.cell[i];
  if (bin < 0)
    return;
  const Cell3 cl = cell_of(g, bin);
  float       s0x[N], s0y[N], s0z[N], dsx[N], dsy[N], dsz[N];
  const int   fx = esk_axis<Order>(ro0[i], p.rx[i], s0x, dsx);
  if constexpr (Dim == 1) {
    // local1d + deposit1d (pic/engine/current.hpp:207-258, nix/esirkepov.hpp:18-67):
    // Jx by the prefix sum, Jy and Jz from the velocities (y, z keep absolute
    // coordinates in 1D) times the time-averaged weight
    const int   gx  = off[c * 3 + 2] + cl.x + fx + g.Nb;
    const long  v   = g.vol;
    const float qs  = dc.q[sp];
    const float vy  = (p.ry[i] - ro1[i]) / dc.delt;
    const float vz  = (p.rz[i] - ro2[i]) / dc.delt;
    auto        add = [&](int x, int k, float val) {
      if ((unsigned)x < (unsigned)g.P[2])
        atomicAdd(&J[k * v + x], val);
    };
    const float wx = -(qs * dc.dxdt);
    float       ww = 0;
    for (int jx = 0; jx < N - 1; jx++) {
      ww += dsx[jx] * wx;
      add(gx + jx + 1, 1, ww);
    }
    for (int jx = 0; jx < N; jx++) {
      const float wt = s0x[jx] + 0.5f * dsx[jx];
      add(gx + jx, 2, wt * (qs * vy));
      add(gx + jx, 3, wt * (qs * vz));
    }
    (void)s0y;
    (void)s0z;
    (void)dsy;
    (void)dsz;
    return;
  }
  const int   fy = esk_axis<Order>(ro1[i], p.ry[i], s0y, dsy);
  int         fz = 0;
  if constexpr (Dim == 3)
    fz = esk_axis<Order>(ro2[i], p.rz[i], s0z, dsz);
  const int   lb = g.Nb;
  const int   gx = off[c * 3 + 2] + cl.x + fx + lb;
  const int   gy = off[c * 3 + 1] + cl.y + fy + lb;
  const int   gz = Dim == 3 ? off[c * 3 + 0] + cl.z + fz + lb : 0;
  const long  v  = g.vol;
  const float qs = dc.q[sp];
  const float A = 0.5f, B = 1.0f / 3.0f;
  auto add = [&](int z, int y, int x, int k, float val) {
    if ((unsigned)z >= (unsigned)g.P[0] || (unsigned)y >= (unsigned)g.P[1] ||
        (unsigned)x >= (unsigned)g.P[2])
      return;
    atomicAdd(&J[k * v + ((long)z * g.P[1] + y) * g.P[2] + x], val);
  };
  if constexpr (Dim == 3) {
    const float qx = qs * dc.dxdt, qy = qs * dc.dydt, qz = qs * dc.dzdt;
    for (int jz = 0; jz < N; jz++)
      for (int jy = 0; jy < N; jy++) {
        const float w  = -((s0y[jy] + A * dsy[jy]) * s0z[jz] + (A * s0y[jy] + B * dsy[jy]) * dsz[jz]) * qx;
        float       ww = 0;
        for (int jx = 0; jx < N - 1; jx++) {
          ww += dsx[jx] * w;
          add(gz + jz, gy + jy, gx + jx + 1, 1, ww);
        }
      }
    for (int jz = 0; jz < N; jz++)
      for (int jx = 0; jx < N; jx++) {
        const float w  = -((s0z[jz] + A * dsz[jz]) * s0x[jx] + (A * s0z[jz] + B * dsz[jz]) * dsx[jx]) * qy;
        float       ww = 0;
        for (int jy = 0; jy < N - 1; jy++) {
          ww += dsy[jy] * w;
          add(gz + jz, gy + jy + 1, gx + jx, 2, ww);
        }
      }
    for (int jy = 0; jy < N; jy++)
      for (int jx = 0; jx < N; jx++) {
        const float w  = -((s0x[jx] + A * dsx[jx]) * s0y[jy] + (A * s0x[jx] + B * dsx[jx]) * dsy[jy]) * qz;
        float       ww = 0;
        for (int jz = 0; jz < N - 1; jz++) {
          ww += dsz[jz] * w;
          add(gz + jz + 1, gy + jy, gx + jx, 3, ww);
        }
      }
  } else {
    // 2D (esirkepov.hpp:76-143): Jz from the velocity (current.hpp:285)
    const float vz = (p.rz[i] - ro2[i]) / dc.delt;
    const float qx = qs * dc.dxdt, qy = qs * dc.dydt;
    for (int jy = 0; jy < N; jy++) {
      const float w  = -(s0y[jy] + A * dsy[jy]) * qx;
      float       ww = 0;
      for (int jx = 0; jx < N - 1; jx++) {
        ww += dsx[jx] * w;
        add(gz, gy + jy, gx + jx + 1, 1, ww);
      }
    }
    for (int jx = 0; jx < N; jx++) {
      const float w  = -(s0x[jx] + A * dsx[jx]) * qy;
      float       ww = 0;
      for (int jy = 0; jy < N - 1; jy++) {
        ww += dsy[jy] * w;
        add(gz, gy + jy + 1, gx + jx, 2, ww);
      }
    }
    for (int jy = 0; jy < N; jy++)
      for (int jx = 0; jx < N; jx++) {
        const float w =
            ((s0x[jx] + A * dsx[jx]) * s0y[jy] + (A * s0x[jx] + B * dsx[jx]) * dsy[jy]) * (qs * vz);
        add(gz, gy + jy, gx + jx, 3, w);
      }
  }
}

// Compensated FP32 accumulation into LDS pairs (hi, lo).  The atomic add on hi
// returns the value it replaced; the lane forms old + x again with the same
// IEEE operation the LDS unit applied, and the rounding error of that addition
// (Knuth's TwoSum, exact in round-to-nearest) goes to lo.  Concurrent adds keep
// this exact: each one captures the error of the one addition it made.  The
// node's value is hi + lo (formed in FP64 at the flush).  A lane first issues
// all its hi adds and only then the lo adds, so the returns of the hi adds are
// awaited once, not one at a time.
__device__ __forceinline__ float two_sum_err(float old, float x)
{
  const float s  = old + x;
  const float bb = s - old;
  return (old - (s - bb)) + (x - bb);
}

constexpr int kTileStaged = 0;
constexpr int kTileF64    = 1;
constexpr int kTilePlain  = 2;
constexpr int kTileComp   = 3;

// The value of ring-tile node idx (tile_d: the FP64 form's nodes; hi, lo: the
// FP32 forms' pairs, lo unused by kTilePlain).
template <int Tile>
__device__ __forceinline__ double tile_value(const double* tile_d, const float* hi,
                                             const float* lo, int idx)
{
  if constexpr (Tile == kTileF64)
    return tile_d[idx];
  else if constexpr (Tile == kTileComp || Tile == kTileStaged)
    return (double)hi[idx] + (double)lo[idx];
  else
    return (double)hi[idx];
}

// Plane-lane march kernel (3D, even orders): the FP64 k_deposit_march in FP32
// registers, with the block's ring tile as compensated FP32 pairs in LDS.  All particles of a bin
// share the tile origin (even orders: the old cell), so no particle needs the
// own-origin path.
template <int Order>
struct March32 {
  static constexpr int N     = Order + 3;
  static constexpr int NPL   = 3 * N;
  static constexpr int PPS   = 64 / NPL >= 4 ? 4 : (64 / NPL >= 2 ? 2 : 1);
  static constexpr int BATCH = 16;
  static constexpr int NS    = N & ~1;
  static constexpr int ROWS  = N + NS / 2;
  static constexpr int PITCH = BATCH + 4;
  static constexpr int WPAIR = 3 * ROWS * PITCH;
  static constexpr int NWAVE = 4;
  static constexpr size_t WLDS = sizeof(float) * 2 * WPAIR * NWAVE;
  // waves per SIMD the register allocation targets (latency bound: at Order 2,
  // five waves measured 1.27 -> 1.14 ms at 64^3 against the unconstrained four;
  // six the same, three slower)
  static constexpr int WAVES = Order == 2 ? 5 : 2;
  static constexpr int waves(int) { return WAVES; }
};

// Tile selects how cell sums reach the ring tile.
//   kTileF64 (the default): one FP64 value per node (LDS double atomics).
//   kTileStaged: compensated FP32 pairs (hi, lo) updated without atomics.  A
//     claimed cell's wave stores its per-cell sums (one N x (N-1) block per
//     plane lane) into the cell's slot of a stage of kst cells; at the end of
//     each phase of kst cells one thread per tile node adds the slots that
//     reach it, in ascending cell order, with TwoSum into (hi, lo).  The
//     block's tile is therefore the same whichever wave claimed which cell.
//   kTileComp: the pairs take LDS float atomics; the error of each add comes
//     from its returned old value (TwoSum).
//   kTilePlain: one FP32 value per node, no compensation.
// All four flush to J with FP32 global atomics, whose rounding bounds J's
// error (a few 1e-7 of max |J|) for every form.  At 128^3 the FP64 tile is the
// fastest (LDS float atomics are several times slower than double ones on this
// GPU, and the staged form pays for its stage), so it is the default; the
// others stay selectable to measure against it.
template <int Order, bool Queue, int Tile>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(March32<Order>::waves(Tile), March32<Order>::waves(Tile))))
    k32_deposit_march(Geom g, const P32 p, const float* __restrict__ ro0,
                      const float* __restrict__ ro1, const float* __restrict__ ro2,
                      const int32_t* __restrict__ cellstart, const int* __restrict__ off,
                      float* __restrict__ J, Dep32 dc, int zrows,
                      int* __restrict__ wq, int gx, int nitems, int kst)
{
  using MP            = March32<Order>;
  constexpr int N     = MP::N;
  constexpr int NPL   = MP::NPL;
  constexpr int PPS   = MP::PPS;
  constexpr int BATCH = MP::BATCH;
  constexpr int NS    = MP::NS;
  constexpr int ROWS  = MP::ROWS;
  constexpr int PITCH = MP::PITCH;
  constexpr int NWAVE = MP::NWAVE;
  static_assert(Order % 2 == 0, "the FP32 march kernel takes even orders");
  static_assert(PITCH % 2 == 0 && (MP::WPAIR * sizeof(float2)) % 16 == 0,
                "phase 2 reads the weights of particle pairs as 16-byte values");

  __shared__ __align__(16) float2 wts[NWAVE][MP::WPAIR];
  __shared__ int    next_cell;
  // kTileComp: hi [3][N z-ring][N][TX] and lo (same) as floats; kTileF64: the
  // same nodes as doubles; kTilePlain: hi only.  Then rcs.
  extern __shared__ double tile_d[];
  float* tile = reinterpret_cast<float*>(tile_d);

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
  const int ylo = by + lb - (Order / 2) - 1; // tile origin (local node index)
  const int xlo = lb - (Order / 2) - 1;
  const int TX  = nbx + N - 1;
  const int psz = N * TX;
  const int tsz = N * psz;
  // kTileStaged: the stage holds the per-cell sums of the kst cells of one
  // phase of the row, [kst][NPL][N][N - 1] (zeros for an empty cell); the
  // ring (hi, lo) is updated only by the phase reductions.
  constexpr int CB    = NPL * N * (N - 1); // staged floats per cell
  float*        thi   = tile;
  float*        tlo   = tile + 3 * tsz;
  float*        stage = tile + 2 * 3 * tsz;
  int*          rcs   = reinterpret_cast<int*>(stage + (Tile == kTileStaged ? kst * CB : 0));
  for (int q = threadIdx.x; q < 3 * tsz; q += blockDim.x) {
    if constexpr (Tile == kTileF64)
      tile_d[q] = 0.0;
    else
      thi[q] = tlo[q] = 0.0f;
  }
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

  const int   lane = threadIdx.x & 63;
  const int   wv   = threadIdx.x >> 6;
  const long  v    = g.vol;
  const float A    = 0.5f;
  const float B    = 1.0f / 3.0f;
  const bool  act   = lane < NPL * PPS;
  const int   plane = act ? lane / PPS : 0;
  const int   k     = lane & (PPS - 1);
  const int   comp  = plane / N;
  const int   t     = plane - comp * N;
  const int   axF   = comp == 0 ? 1 : (comp == 1 ? 2 : 0);
  const int   axS   = comp == 0 ? 2 : (comp == 1 ? 0 : 1);
  const int   axC   = comp;
  const float ddt   = comp == 0 ? dc.dxdt : (comp == 1 ? dc.dydt : dc.dzdt);
  const int   a1    = lane / BATCH;
  const int   k1    = lane - a1 * BATCH;
  const float* R0   = a1 == 0 ? ro0 : (a1 == 1 ? ro1 : ro2);
  const float* R1   = a1 == 0 ? p.rx : (a1 == 1 ? p.ry : p.rz);
  float2*      W    = &wts[wv][0];
  // phase 2 reads the weights of two neighbouring particles at once, the second
  // possibly past the batch: the slots start as zeros, so such a read is finite
  for (int q = lane; q < MP::WPAIR; q += 64)
    W[q] = make_float2(0.0f, 0.0f);

  // the next cell of [next_cell, cend) that holds particles of any species
  // (uniform per wave); false when the range is done
  int  cbeg = 0, cend = nbx; // the cells of the current phase (kTileStaged) or row
  auto claim = [&](int& bx) {
    for (;;) {
      int q = 0;
      if (lane == 0)
        q = atomicAdd(&next_cell, 1);
      q = __builtin_amdgcn_readfirstlane(q);
      if (q >= cend)
        return false;
      bx = q;
      for (int sp = 0; sp < ns; sp++)
        if (rcs[sp * (nbx + 1) + bx + 1] > rcs[sp * (nbx + 1) + bx])
          return true;
      if constexpr (Tile == kTileStaged) // an empty cell stages zeros
        for (int e = lane; e < CB; e += 64)
          stage[(bx - cbeg) * CB + e] = 0.0f;
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
  auto fetch = [&](int seg, int pb, int p1, float& xo, float& xn) {
    if (a1 < 3 && pb + k1 < p1) {
      const long i = (long)seg * g.cap + pb + k1;
      xo           = R0[i];
      xn           = R1[i];
    }
  };

  const int nph = Tile == kTileStaged ? (nbx + kst - 1) / kst : 1; // phases per row
  for (int bz = bz0; bz < bz1; bz++) {
    const int zlo = bz + lb - (Order / 2) - 1;
    const int zs  = ((zlo % N) + N) % N;
    for (int ph = 0; ph < nph; ph++) {
    cbeg = Tile == kTileStaged ? ph * kst : 0;
    cend = Tile == kTileStaged ? min(nbx, cbeg + kst) : nbx;
    // one claimed cell at a time; all its species accumulate into the same
    // registers, flushed to the tile once per cell
    int       bx, sp = 0, p0 = 0, p1 = 0;
    bool      have = claim(bx);
    float     nxo = 0, nxn = 0;
    if (have) {
      species_from(bx, sp, p0, p1);
      fetch(c * ns + sp, p0, p1, nxo, nxn);
    }
    while (have) {
      int         bxn, spn = 0, p0n = 0, p1n = 0;
      const bool  more = claim(bxn);
      if (more)
        species_from(bxn, spn, p0n, p1n);
      float       acc[N][N - 1];
#pragma unroll
      for (int r = 0; r < N; r++)
#pragma unroll
        for (int j = 0; j < N - 1; j++)
          acc[r][j] = 0.0f;
      for (;;) { // species of cell bx
      const int   seg = c * ns + sp;
      const float qd  = dc.q[sp] * ddt;
      int         sp2 = sp + 1, q0 = 0, q1 = 0;
      const bool  next_sp = species_from(bx, sp2, q0, q1);
      for (int pb = p0; pb < p1; pb += BATCH) {
        const int   nbt = min(BATCH, p1 - pb);
        const float xo  = nxo, xn = nxn;
        if (pb + BATCH < p1)
          fetch(seg, pb + BATCH, p1, nxo, nxn);
        else if (next_sp)
          fetch(c * ns + sp2, q0, q1, nxo, nxn);
        else if (more)
          fetch(c * ns + spn, p0n, p1n, nxo, nxn);
        if (a1 < 3 && k1 < nbt) {
          float s0[N], ds[N];
          esk_axis<Order>(xo, xn, s0, ds);
          float2* pr = W + a1 * ROWS * PITCH + k1;
#pragma unroll
          for (int l = 0; l < N; l++)
            pr[l * PITCH] = make_float2(s0[l], ds[l]);
          float pre[NS];
          pre[0] = ds[0];
#pragma unroll
          for (int l = 1; l < NS; l++)
            pre[l] = l < N - 1 ? pre[l - 1] + ds[l] : 0.0f;
#pragma unroll
          for (int l = 0; l < NS / 2; l++)
            pr[(N + l) * PITCH] = make_float2(pre[2 * l], pre[2 * l + 1]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // each lane takes two neighbouring particles kk, kk + 1 per pass, whose
        // weights one 16-byte read returns (the second counts only inside the
        // batch: its weights are zeroed otherwise)
        for (int k0 = 0; k0 < nbt; k0 += 2 * PPS) {
          const int kk = k0 + 2 * k;
          if (act && kk < nbt) {
            const float4* pp = reinterpret_cast<const float4*>(W + kk);
            constexpr int P4 = PITCH / 2; // the pitch in float4
            const float4  pf = pp[(axF * ROWS + t) * P4];
            float4        psr[N]; // (s0S[r], dsS[r]) of both, every row read before use
#pragma unroll
            for (int r = 0; r < N; r++)
              psr[r] = pp[(axS * ROWS + r) * P4];
            const bool  two = kk + 1 < nbt;
            const float aF0 = fmaf(A, pf.y, pf.x);
            const float bF0 = fmaf(B, pf.y, A * pf.x);
            const float aF1 = two ? fmaf(A, pf.w, pf.z) : 0.0f;
            const float bF1 = two ? fmaf(B, pf.w, A * pf.z) : 0.0f;
            float       P0[NS], P1[NS];
#pragma unroll
            for (int j = 0; j < NS / 2; j++) {
              const float4 d4 = pp[(axC * ROWS + N + j) * P4];
              P0[2 * j]       = d4.x;
              P0[2 * j + 1]   = d4.y;
              P1[2 * j]       = d4.z;
              P1[2 * j + 1]   = d4.w;
            }
#pragma unroll
            for (int r = 0; r < N; r++) {
              const float4 ps = psr[r];
              const float  w0 = -fmaf(aF0, ps.x, bF0 * ps.y) * qd;
              const float  w1 = -fmaf(aF1, ps.z, bF1 * ps.w) * qd;
#pragma unroll
              for (int j = 0; j < N - 1; j++)
                acc[r][j] = fmaf(P1[j], w1, fmaf(P0[j], w0, acc[r][j]));
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
#pragma unroll
      for (int r = 0; r < N; r++)
#pragma unroll
        for (int j = 0; j < N - 1; j++) {
          float s = acc[r][j];
          if (PPS >= 2)
            s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0xB1, 0xF, 0xF, true));
          if (PPS >= 4)
            s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x4E, 0xF, 0xF, true));
          acc[r][j] = s;
        }
      if (act && k == 0) {
        int ro[N], co[N - 1];
        if constexpr (Tile != kTileStaged)
          plane_offsets<N>(comp, t, zs, TX, psz, tsz, bx, ro, co);
        if constexpr (Tile == kTileStaged) {
          float* st = stage + (bx - cbeg) * CB + plane * (N * (N - 1));
#pragma unroll
          for (int r = 0; r < N; r++)
#pragma unroll
            for (int j = 0; j < N - 1; j++)
              st[r * (N - 1) + j] = acc[r][j];

        } else if constexpr (Tile == kTileComp) {
          // per row: the N - 1 hi adds, then their N - 1 error adds
#pragma unroll
          for (int r = 0; r < N; r++) {
            float old[N - 1];
#pragma unroll
            for (int j = 0; j < N - 1; j++)
              old[j] = atomicAdd(&thi[ro[r] + co[j]], acc[r][j]);
#pragma unroll
            for (int j = 0; j < N - 1; j++)
              atomicAdd(&tlo[ro[r] + co[j]], two_sum_err(old[j], acc[r][j]));
          }
        } else {
#pragma unroll
          for (int r = 0; r < N; r++)
#pragma unroll
            for (int j = 0; j < N - 1; j++) {
              if constexpr (Tile == kTileF64)
                atomicAdd(&tile_d[ro[r] + co[j]], (double)acc[r][j]);
              else
                atomicAdd(&thi[ro[r] + co[j]], acc[r][j]);
            }
        }
      }
      have = more;
      sp   = spn;
      bx   = bxn;
      p0   = p0n;
      p1   = p1n;
    }
    __syncthreads();
    if constexpr (Tile == kTileStaged) {
      // Every tile node that the phase's cells reach (columns cbeg .. cend+N-2)
      // takes their staged sums in ascending cell order, TwoSum into (hi, lo),
      // one thread per node: thread -> (column, row group), wp = the window
      // width rounded up to a power of two.
      const bool last = ph == nph - 1;
      const int  wx   = (cend - cbeg) + N - 1;
      int        lw   = 0;
      while ((1 << lw) < wx)
        lw++;
      const int col = (int)threadIdx.x & ((1 << lw) - 1);
      const int tx  = cbeg + col;
      if (col < wx) {
        constexpr int PB = N * (N - 1); // floats per plane block
        const int     cb0 = tx - N + 1;
        for (int rr = (int)threadIdx.x >> lw; rr < 3 * N * N; rr += 256 >> lw) {
          const int ty = rr % N;
          const int sl = (rr / N) % N;
          const int kc = rr / (N * N);
          const int q  = kc * tsz + sl * psz + ty * TX + tx;
          const int oz = sl >= zs ? sl - zs : sl - zs + N; // z offset from zlo
          // the cells that reach this node are cb0 + i (ascending), their staged
          // entries base + i * step:
          //   Jx: z = r, y = t, x = j + 1 -> i < N - 1, j = N - 2 - i
          //   Jy: z = t, y = j + 1, x = r -> i < N, r = N - 1 - i
          //   Jz: z = j + 1, y = r, x = t -> i < N, t = N - 1 - i
          if (kc == 1 ? ty < 1 : (kc == 2 && oz < 1))
            continue;
          const int cnt  = kc == 0 ? N - 1 : N;
          const int base = (cb0 - cbeg) * CB +
                           (kc == 0 ? ty * PB + oz * (N - 1) + N - 2
                            : kc == 1 ? (N + oz) * PB + (N - 1) * (N - 1) + ty - 1
                                      : (3 * N - 1) * PB + ty * (N - 1) + oz - 1);
          const int step = CB - (kc == 0 ? 1 : kc == 1 ? N - 1 : PB);
          // the phase's cells among cb0 .. cb0 + cnt - 1
          const int i0 = max(0, cbeg - cb0), i1 = min(cnt, cend - cb0);
          float     hi = thi[q], lo = tlo[q];
          for (int i = i0; i < i1; i++) {
            const float x = stage[base + i * step];
            lo += two_sum_err(hi, x);
            hi += x;
          }
          thi[q] = hi;
          tlo[q] = lo;
        }
      }
      if (last) {
        // plane zlo is complete: to J, and its ring slot cleared
        __syncthreads();
        for (int q = threadIdx.x; q < 3 * psz; q += blockDim.x) {
          const int    kc  = q / psz;
          const int    r2  = q - kc * psz;
          const int    idx = kc * tsz + zs * psz + r2;
          const double val = (double)thi[idx] + (double)tlo[idx];
          if (val != 0.0) {
            const int ty = r2 / TX;
            const int tx2 = r2 - ty * TX;
            const int gz = off[c * 3 + 0] + zlo, gy = off[c * 3 + 1] + ylo + ty,
                      gx = off[c * 3 + 2] + xlo + tx2;
            if ((unsigned)gz < (unsigned)g.P[0] && (unsigned)gy < (unsigned)g.P[1] &&
                (unsigned)gx < (unsigned)g.P[2])
              atomicAdd(&J[(kc + 1) * v + ((long)gz * g.P[1] + gy) * g.P[2] + gx], (float)val);
            thi[idx] = tlo[idx] = 0.0f;
          }
        }
      }
      if (!last) {
        if (threadIdx.x == 0)
          next_cell = cend;
        __syncthreads();
      }
    } else {
    for (int q = threadIdx.x; q < 3 * psz; q += blockDim.x) {
      const int    kc  = q / psz;
      const int    r2  = q - kc * psz;
      const int    idx = kc * tsz + zs * psz + r2;
      const double val = tile_value<Tile>(tile_d, thi, tlo, idx);
      if (val != 0.0) {
        const int ty = r2 / TX;
        const int tx = r2 - ty * TX;
        const int gz = off[c * 3 + 0] + zlo, gy = off[c * 3 + 1] + ylo + ty,
                  gx = off[c * 3 + 2] + xlo + tx;
        if ((unsigned)gz < (unsigned)g.P[0] && (unsigned)gy < (unsigned)g.P[1] &&
            (unsigned)gx < (unsigned)g.P[2])
          atomicAdd(&J[(kc + 1) * v + ((long)gz * g.P[1] + gy) * g.P[2] + gx], (float)val);
        if constexpr (Tile == kTileF64)
          tile_d[idx] = 0.0;
        else
          thi[idx] = tlo[idx] = 0.0f;
      }
    }
    }
    } // phase
    if (bz + 1 < bz1)
      load_row(bz + 1);
    if (threadIdx.x == 0)
      next_cell = 0;
    __syncthreads();
  }
  const int zl = (bz1 - 1) + lb - (Order / 2) - 1;
  const int zs = ((zl % N) + N) % N;
  for (int q = threadIdx.x; q < 3 * (N - 1) * psz; q += blockDim.x) {
    const int    kc  = q / ((N - 1) * psz);
    const int    r   = q - kc * (N - 1) * psz;
    const int    dz  = 1 + r / psz;
    const int    r2  = r - (dz - 1) * psz;
    const int    sl  = zs + dz >= N ? zs + dz - N : zs + dz;
    const double val = tile_value<Tile>(tile_d, thi, tlo, kc * tsz + sl * psz + r2);
    if (val == 0.0)
      continue;
    const int ty = r2 / TX;
    const int tx = r2 - ty * TX;
    const int gz = off[c * 3 + 0] + zl + dz, gy = off[c * 3 + 1] + ylo + ty,
              gx = off[c * 3 + 2] + xlo + tx;
    if ((unsigned)gz < (unsigned)g.P[0] && (unsigned)gy < (unsigned)g.P[1] &&
        (unsigned)gx < (unsigned)g.P[2])
      atomicAdd(&J[(kc + 1) * v + ((long)gz * g.P[1] + gy) * g.P[2] + gx], (float)val);
  }
  } while (false);
  if constexpr (!Queue)
    return;
  __syncthreads(); // the LDS of this item is read out before the next one
  }
  if constexpr (Queue)
    wq_finish(wq);
}

// rho from the current (sorted or not) positions, one thread per particle:
// qs times the (Order+1)^Dim cell-centred weights (esirkepov.hpp:145-155).
template <int Dim, int Order>
__global__ void __launch_bounds__(256)
    k32_rho_global(Geom g, P32 p, const int32_t* __restrict__ np, const int* __restrict__ off,
                   float* __restrict__ J, const int32_t* __restrict__ perm,
                   const int* __restrict__ nb)
{
  constexpr int S   = Order + 1;
  const int     seg = blockIdx.y;
  const long    ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const int  c   = seg / g.Ns;
  const int  sp  = seg - c * g.Ns;
  const long i   = (long)seg * g.cap + ip;
  int        bin;
  float      r[3];
  if (perm) {
    // a pending sort: sorted slot i holds record perm[i], whose bin is found
    // again and whose offset is folded into [0, 1), as k32_apply_perm does
    const long  j = perm[i];
    const float x = p.rx[j], y = p.ry[j], z = p.rz[j];
    int         ds, k;
    dest32(g, (int)(j / g.cap), p.cell[j], x, y, z, nb, ds, bin, k);
    r[0] = g.has[2] ? frac01(x) : x;
    r[1] = g.has[1] ? frac01(y) : y;
    r[2] = g.has[0] ? frac01(z) : z;
  } else {
    bin  = p.cell[i];
    r[0] = p.rx[i];
    r[1] = p.ry[i];
    r[2] = p.rz[i];
  }
  if (bin < 0)
    return;
  const Cell3 cl   = cell_of(g, bin);
  const int   ci[3] = {cl.x, cl.y, cl.z};
  float       w[3][S];
  int         o[3];
  for (int a = 0; a < 3; a++) {
    if (a < Dim) {
      // unwrapped r (after a position push) is allowed: split off its cell shift
      const float fl  = floorf(r[a]);
      const float rr  = r[a] - fl;
      int         fst = 0;
      shape_mc<Order, float>(int_delta<Order>(rr, fst), 0.0f, 1.0f, w[a]);
      o[a] = ci[a] + (int)fl + fst + g.Nb - (Order / 2);
    } else {
      w[a][0] = 1.0f;
      o[a]    = g.Nb;
    }
  }
  const float qs = (float)g.q[sp];
  for (int jz = 0; jz < (Dim == 3 ? S : 1); jz++)
    for (int jy = 0; jy < (Dim >= 2 ? S : 1); jy++)
      for (int jx = 0; jx < S; jx++) {
        const float val = Dim == 3   ? qs * w[0][jx] * w[1][jy] * w[2][jz]
                          : Dim == 2 ? qs * w[0][jx] * w[1][jy]
                                     : qs * w[0][jx];
        const int   gz  = g.has[0] ? off[c * 3 + 0] + o[2] + jz : 0;
        const int   gy  = g.has[1] ? off[c * 3 + 1] + o[1] + jy : 0;
        const int   gx  = off[c * 3 + 2] + o[0] + jx;
        if ((unsigned)gz >= (unsigned)g.P[0] || (unsigned)gy >= (unsigned)g.P[1] ||
            (unsigned)gx >= (unsigned)g.P[2])
          continue;
        atomicAdd(&J[((long)gz * g.P[1] + gy) * g.P[2] + gx], val);
      }
}

// rho after the sort (3D, even orders): the FP64 k_rho_march walk in FP32
// registers with a compensated FP32 LDS ring tile.  Sorted FP32 particles carry
// r in [0, 1) and sit in their cell's bin, so every particle of a bin has the
// bin's stencil origin.
template <int Order>
struct Rho32 {
  static constexpr int S     = Order + 1;
  static constexpr int PPS   = 64 / S >= 32 ? 32 : (64 / S >= 16 ? 16 : 8);
  static constexpr int BATCH = 64;
  static constexpr int NWAVE = 4;
  static constexpr size_t WLDS = sizeof(float) * NWAVE * BATCH * 3 * S;
};

template <int Order>
__global__ void __launch_bounds__(256)
    k32_rho_march(Geom g, P32 p, const int32_t* __restrict__ cellstart,
                  const int* __restrict__ off, float* __restrict__ J, int zrows,
                  const int32_t* __restrict__ perm)
{
  using RP            = Rho32<Order>;
  constexpr int S     = RP::S;
  constexpr int PPS   = RP::PPS;
  constexpr int BATCH = RP::BATCH;
  constexpr int NWAVE = RP::NWAVE;
  static_assert(Order % 2 == 0, "the FP32 rho march kernel takes even orders");

  __shared__ float wts[NWAVE][BATCH][3 * S];
  __shared__ int   next_cell;
  extern __shared__ float tile[]; // compensated pairs: hi [S z-ring][S][TX], lo (same), then rcs

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
  const int ylo = by + lb - (Order / 2);
  const int xlo = lb - (Order / 2);
  const int TX  = nbx + Order;
  const int psz = S * TX;
  const int tsz = S * psz;
  float*    thi = tile;
  float*    tlo = tile + tsz;
  int*      rcs = reinterpret_cast<int*>(tile + 2 * tsz);
  for (int q = threadIdx.x; q < tsz; q += blockDim.x)
    thi[q] = tlo[q] = 0.0f;
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

  const int  lane = threadIdx.x & 63;
  const int  wv   = threadIdx.x >> 6;
  const bool act  = lane < S * PPS;
  const int  t    = act ? lane / PPS : 0; // z plane
  const int  k    = lane & (PPS - 1);
  float*     W    = &wts[wv][0][0];

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
  // a pending sort: sorted slot i holds record perm[i], whose offset (still
  // relative to its old cell) is folded into [0, 1) of the slot's bin
  auto fetch = [&](int seg, int pb, int p1, float* x3) {
    if (pb + lane < p1) {
      const long i = (long)seg * g.cap + pb + lane;
      if (perm) {
        const long j = perm[i];
        x3[0]        = g.has[2] ? frac01(p.rx[j]) : p.rx[j];
        x3[1]        = g.has[1] ? frac01(p.ry[j]) : p.ry[j];
        x3[2]        = g.has[0] ? frac01(p.rz[j]) : p.rz[j];
      } else {
        x3[0] = p.rx[i];
        x3[1] = p.ry[i];
        x3[2] = p.rz[i];
      }
    }
  };

  for (int bz = bz0; bz < bz1; bz++) {
    const int zlo = bz + lb - (Order / 2);
    const int zs  = ((zlo % S) + S) % S;
    int       sp, bx, p0, p1;
    bool      have   = claim(sp, bx, p0, p1);
    float     nx3[3] = {0, 0, 0};
    if (have)
      fetch(c * ns + sp, p0, p1, nx3);
    while (have) {
      int         spn, bxn, p0n, p1n;
      const bool  more = claim(spn, bxn, p0n, p1n);
      const int   seg  = c * ns + sp;
      const float qs   = (float)g.q[sp];
      float       acc[S][S];
#pragma unroll
      for (int a = 0; a < S; a++)
#pragma unroll
        for (int e = 0; e < S; e++)
          acc[a][e] = 0.0f;
      for (int pb = p0; pb < p1; pb += BATCH) {
        const int   nbt   = min(BATCH, p1 - pb);
        const float r3[3] = {nx3[0], nx3[1], nx3[2]};
        if (pb + BATCH < p1)
          fetch(seg, pb + BATCH, p1, nx3);
        else if (more)
          fetch(c * ns + spn, p0n, p1n, nx3);
        if (lane < nbt) {
          float* rec = W + lane * 3 * S;
#pragma unroll
          for (int a = 0; a < 3; a++) {
            int fst;
            shape_mc<Order, float>(int_delta<Order>(r3[a], fst), 0.0f, 1.0f, &rec[a * S]);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int k0 = 0; k0 < nbt; k0 += PPS) {
          const int kk = k0 + k;
          if (act && kk < nbt) {
            const float* rec = W + kk * 3 * S;
            const float  sz  = rec[2 * S + t];
            float        sx[S];
#pragma unroll
            for (int jx = 0; jx < S; jx++)
              sx[jx] = rec[jx];
#pragma unroll
            for (int jy = 0; jy < S; jy++) {
              const float sy = rec[S + jy];
#pragma unroll
              for (int jx = 0; jx < S; jx++)
                acc[jy][jx] = fmaf(qs * sx[jx] * sy, sz, acc[jy][jx]);
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
          float x = acc[a][e];
          x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));
          x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, true));
#pragma unroll
          for (int m = 4; m < PPS; m <<= 1)
            x += __shfl_xor(x, m);
          acc[a][e] = x;
        }
      if (act && k == 0) {
        const int sl = zs + t >= S ? zs + t - S : zs + t;
        float     old[S][S];
#pragma unroll
        for (int jy = 0; jy < S; jy++)
#pragma unroll
          for (int jx = 0; jx < S; jx++)
            old[jy][jx] = atomicAdd(&thi[sl * psz + jy * TX + bx + jx], acc[jy][jx]);
#pragma unroll
        for (int jy = 0; jy < S; jy++)
#pragma unroll
          for (int jx = 0; jx < S; jx++)
            atomicAdd(&tlo[sl * psz + jy * TX + bx + jx], two_sum_err(old[jy][jx], acc[jy][jx]));
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
      const double val = (double)thi[idx] + (double)tlo[idx];
      if (val != 0.0) {
        const int ty = q / TX;
        const int tx = q - ty * TX;
        const int gz = off[c * 3 + 0] + zlo, gy = off[c * 3 + 1] + ylo + ty,
                  gx = off[c * 3 + 2] + xlo + tx;
        if ((unsigned)gz < (unsigned)g.P[0] && (unsigned)gy < (unsigned)g.P[1] &&
            (unsigned)gx < (unsigned)g.P[2])
          atomicAdd(&J[((long)gz * g.P[1] + gy) * g.P[2] + gx], (float)val);
        thi[idx] = tlo[idx] = 0.0f;
      }
    }
    if (bz + 1 < bz1)
      load_row(bz + 1);
    if (threadIdx.x == 0)
      next_cell = 0;
    __syncthreads();
  }
  const int zl = (bz1 - 1) + lb - (Order / 2);
  const int zs = ((zl % S) + S) % S;
  for (int q = threadIdx.x; q < (S - 1) * psz; q += blockDim.x) {
    const int    dz  = 1 + q / psz;
    const int    r2  = q - (dz - 1) * psz;
    const int    sl  = zs + dz >= S ? zs + dz - S : zs + dz;
    const double val = (double)thi[sl * psz + r2] + (double)tlo[sl * psz + r2];
    if (val == 0.0)
      continue;
    const int ty = r2 / TX;
    const int tx = r2 - ty * TX;
    const int gz = off[c * 3 + 0] + zl + dz, gy = off[c * 3 + 1] + ylo + ty,
              gx = off[c * 3 + 2] + xlo + tx;
    if ((unsigned)gz < (unsigned)g.P[0] && (unsigned)gy < (unsigned)g.P[1] &&
        (unsigned)gx < (unsigned)g.P[2])
      atomicAdd(&J[((long)gz * g.P[1] + gy) * g.P[2] + gx], (float)val);
  }
}

void march32_split(const Geom& g, int fixed, int& zrows, int& nzr)
{
  const int  nby   = g.bsz / g.bsy;
  const int  nbz   = g.nbin / g.bsz;
  const long cols  = (long)nby * g.nchunk;
  const int  split = (int)std::min<long>(nbz, std::max<long>(1, (4096 + cols - 1) / cols));
  zrows            = fixed > 0 ? std::min(fixed, nbz) : (nbz + split - 1) / split;
  nzr              = (nbz + zrows - 1) / zrows;
}

} // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch32_from_f64(Sim& s, int seg0, int nseg)
{
  const Geom& g = s.g;
  hipLaunchKernelGGL(k32_from_f64, dim3(nblk(g.cap, 256), nseg), dim3(256), 0, s.stream, g, s.pa,
                     s.qa, s.d_np, s.d_lim, seg0, nseg);
}

void launch32_to_f64(Sim& s, int seg0, int nseg, bool old)
{
  const Geom& g = s.g;
  hipLaunchKernelGGL(k32_to_f64, dim3(nblk(g.cap, 256), nseg), dim3(256), 0, s.stream, g, s.qa,
                     s.pa, s.f_ro[0], s.f_ro[1], s.f_ro[2], old ? s.d_xo[0] : nullptr,
                     old ? s.d_xo[1] : nullptr, old ? s.d_xo[2] : nullptr, s.d_np, s.d_lim, seg0);
}

void launch32_fields_to_f64(Sim& s, int which)
{
  const long v    = s.g.vol;
  auto       conv = [&](const float* a, double* b, long n) {
    hipLaunchKernelGGL((k32_convert<float, double>), dim3(4096), dim3(256), 0, s.stream, a, b, n);
  };
  if (which == PN_FIELD_UF) {
    conv(s.f_E, s.d_E, 3 * v);
    conv(s.f_B, s.d_B, 3 * v);
  } else {
    conv(s.f_J, s.d_J, 4 * v);
  }
}

void launch32_fields_from_f64(Sim& s, int which)
{
  const long v    = s.g.vol;
  auto       conv = [&](const double* a, float* b, long n) {
    hipLaunchKernelGGL((k32_convert<double, float>), dim3(4096), dim3(256), 0, s.stream, a, b, n);
  };
  if (which == PN_FIELD_UF) {
    conv(s.d_E, s.f_E, 3 * v);
    conv(s.d_B, s.f_B, 3 * v);
  } else {
    conv(s.d_J, s.f_J, 4 * v);
  }
}

static Pos32 pos32_args(Sim& s, double delt, int on)
{
  const Geom& g = s.g;
  Pos32       ps{};
  ps.ro0     = s.f_ro[0];
  ps.ro1     = s.f_ro[1];
  ps.ro2     = s.f_ro[2];
  ps.delt    = (float)delt;
  ps.rc      = (float)(1 / g.cc);
  ps.rdx     = (float)(1 / g.del[2]);
  ps.rdy     = (float)(1 / g.del[1]);
  ps.rdz     = (float)(1 / g.del[0]);
  ps.nb      = s.d_nb;
  ps.off     = s.d_off;
  ps.cellcnt = s.ss.cellcnt;
  ps.key     = s.ss.key;
  ps.slot    = s.ss.slot;
  ps.on      = on;
  return ps;
}

// with_position: when the row gather runs, it also runs each particle's
// position push (k32_push_position's work); returns true if it did.
// The FP32 row gather's launch shape: yr bin rows per block and the LDS they
// stage; false when the row form does not apply.
static bool vel32_rows_shape(const Sim& s, int& yr, size_t& lds)
{
  const Geom& g      = s.g;
  const int   T      = g.order + 2;
  auto        lds_of = [&](int r) {
    return sizeof(float) * 6 * (g.dim == 3 ? T : 1) * (T + r - 1) * (g.bsy + g.order + 1);
  };
  yr = s.vel_rows > 0 ? s.vel_rows : 4; // 0: four rows, or fewer where LDS is short
  while (yr > 1 && lds_of(yr) > 65536)
    yr = (yr + 1) / 2;
  lds = lds_of(yr);
  return g.dim >= 2 && s.sorted && !s.generic_kernels && !s.vel_generic && lds <= 65536;
}

bool gather32_takes_perm(const Sim& s)
{
  int    yr;
  size_t lds;
  return vel32_rows_shape(s, yr, lds);
}

// The record move of a deferred FP32 sort: sorted slot i of every segment
// takes record perm[i] (any segment), its bin found again (dest32) and its
// offset folded into [0, 1), as the sort's scatter writes it.
__global__ void __launch_bounds__(256)
    k32_apply_perm(Geom g, P32 src, P32 dst, const int32_t* __restrict__ np,
                   const int32_t* __restrict__ perm, const int* __restrict__ nb)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long  i = (long)seg * g.cap + ip;
  const long  j = perm[i];
  const float x = src.rx[j], y = src.ry[j], z = src.rz[j];
  int         ds, b, k;
  dest32(g, (int)(j / g.cap), src.cell[j], x, y, z, nb, ds, b, k);
  dst.rx[i]   = g.has[2] ? frac01(x) : x;
  dst.ry[i]   = g.has[1] ? frac01(y) : y;
  dst.rz[i]   = g.has[0] ? frac01(z) : z;
  dst.ux[i]   = src.ux[j];
  dst.uy[i]   = src.uy[j];
  dst.uz[i]   = src.uz[j];
  dst.id[i]   = src.id[j];
  dst.cell[i] = b;
}

bool launch32_push_velocity(Sim& s, double delt, bool with_position)
{
  PhaseTimer  pt(s, 1);
  const Geom& g = s.g;
  Push32      pc{};
  for (int sp = 0; sp < g.Ns; sp++)
    pc.qmdt[sp] = (float)(0.5 * g.q[sp] / g.m[sp] * delt);
  pc.pusher = s.cfg.pusher;
  pc.wt     = s.cfg.interp;
  for (int a = 0; a < 3; a++) {
    pc.cdt[a]  = (float)(g.cc * delt / g.del[2 - a]);
    pc.rcdt[a] = 1.0f / pc.cdt[a];
  }
  const int    nby  = g.dim == 3 ? g.bsz / g.bsy : g.nbin / g.bsy;
  int          yr;
  size_t       lds;
  const bool   rows = vel32_rows_shape(s, yr, lds);
  // a sort left pending by the previous step: the row gather applies it;
  // any other form needs the particles in place first
  if (s.perm_pending && !rows)
    materialize(s);
  const bool     use_perm = s.perm_pending;
  const int32_t* perm     = use_perm ? s.ss.perm : nullptr;
  const P32      dst      = use_perm ? s.qb : s.qa;
  const int    nblk_rows = (g.nbin / g.bsy / nby) * ((nby + yr - 1) / yr);
  const int    nsplit    = s.vel_split ? g.Ns : 1;
  const bool   fused     = rows && with_position;
  const Pos32  ps        = pos32_args(s, delt, fused ? 1 : 0);
  if (fused)
    launch_zero(s, s.ss.cellcnt, (long)g.nseg * g.nbin);
#define PN32_VEL_ROWS(D, O, Q, F)                                                             \
  hipLaunchKernelGGL((k32_push_velocity_rows<D, O, Q, F>),                                    \
                     Q ? dim3(queue_blocks(                                                   \
                             s, reinterpret_cast<const void*>(&k32_push_velocity_rows<D, O, Q, F>), \
                             lds, nitems))                                                    \
                       : dim3(nblk_rows, gy),                                                 \
                     dim3(256), lds, s.stream, g, s.qa, s.ss.cellstart, s.d_off, s.f_E, s.f_B, \
                     pc, ps, yr, nsplit, wq, nblk_rows, nitems, perm, dst)
#define PN32_VEL(D, O)                                                                        \
  do {                                                                                        \
    if (rows) {                                                                               \
      const int gy = g.nchunk * nsplit, nitems = nblk_rows * gy;                              \
      int*      wq = s.d_wq + 2 * kWqGather32;                                                \
      if (s.vel_queue && fused)                                                               \
        PN32_VEL_ROWS(D, O, true, true);                                                      \
      else if (s.vel_queue)                                                                   \
        PN32_VEL_ROWS(D, O, true, false);                                                     \
      else if (fused)                                                                         \
        PN32_VEL_ROWS(D, O, false, true);                                                     \
      else                                                                                    \
        PN32_VEL_ROWS(D, O, false, false);                                                    \
    } else {                                                                                  \
      hipLaunchKernelGGL((k32_push_velocity<D, O>), dim3(nblk(g.cap, 256), g.nseg), dim3(256), \
                         0, s.stream, g, s.qa, s.d_np, s.d_off, s.f_E, s.f_B, pc);            \
    }                                                                                         \
  } while (0)
  if (g.dim == 1) {
    const dim3 grid(nblk(g.cap, 256), g.nseg);
    switch (g.order) {
    case 1: hipLaunchKernelGGL((k32_push_velocity<1, 1>), grid, dim3(256), 0, s.stream, g, s.qa, s.d_np, s.d_off, s.f_E, s.f_B, pc); break;
    case 2: hipLaunchKernelGGL((k32_push_velocity<1, 2>), grid, dim3(256), 0, s.stream, g, s.qa, s.d_np, s.d_off, s.f_E, s.f_B, pc); break;
    case 3: hipLaunchKernelGGL((k32_push_velocity<1, 3>), grid, dim3(256), 0, s.stream, g, s.qa, s.d_np, s.d_off, s.f_E, s.f_B, pc); break;
    default: hipLaunchKernelGGL((k32_push_velocity<1, 4>), grid, dim3(256), 0, s.stream, g, s.qa, s.d_np, s.d_off, s.f_E, s.f_B, pc); break;
    }
  } else if (g.dim == 3) {
    switch (g.order) {
    case 1: PN32_VEL(3, 1); break;
    case 2: PN32_VEL(3, 2); break;
    case 3: PN32_VEL(3, 3); break;
    default: PN32_VEL(3, 4); break;
    }
  } else {
    switch (g.order) {
    case 1: PN32_VEL(2, 1); break;
    case 2: PN32_VEL(2, 2); break;
    case 3: PN32_VEL(2, 3); break;
    default: PN32_VEL(2, 4); break;
    }
  }
#undef PN32_VEL
#undef PN32_VEL_ROWS
  if (use_perm) {
    std::swap(s.qa, s.qb);
    s.perm_pending = false;
  }
  if (fused) {
    s.counted = true;
    s.sorted  = false;
  }
  return fused;
}

void materialize32(Sim& s)
{
  if (!s.perm_pending)
    return;
  const Geom& g = s.g;
  hipLaunchKernelGGL(k32_apply_perm, dim3(s.ss.ntile, g.nseg), dim3(256), 0, s.stream, g, s.qa,
                     s.qb, s.d_np, s.ss.perm, s.d_nb);
  std::swap(s.qa, s.qb);
  s.perm_pending = false;
}

void launch32_push_position(Sim& s, double delt)
{
  materialize(s);
  PhaseTimer  pt(s, 2);
  const Geom& g = s.g;
  launch_zero(s, s.ss.cellcnt, (long)g.nseg * g.nbin);
  hipLaunchKernelGGL(k32_push_position, dim3(s.ss.ntile, g.nseg), dim3(256), 0, s.stream, g, s.qa,
                     pos32_args(s, delt, 0), s.d_np);
  s.counted = true;
  s.sorted  = false;
}

void launch32_wall_apply(Sim& s)
{
  const int nk = (int)s.shk.useg.size();
  if (nk == 0)
    return;
  hipLaunchKernelGGL(k32_wall_apply, dim3(nk), dim3(256), 0, s.stream, s.g, s.qa, s.shk.d_useg,
                     s.shk.d_list, s.shk.d_count, s.shk.d_voff, s.shk.d_vals, s.d_lim, s.d_nb,
                     s.ss.cellcnt, s.ss.key, s.ss.slot);
}

void launch32_inject_append(Sim& s)
{
  const int nk = (int)s.shk.useg.size();
  if (nk == 0)
    return;
  hipLaunchKernelGGL(k32_inject_append, dim3(nk), dim3(256), 0, s.stream, s.g, s.qa, s.f_ro[0],
                     s.f_ro[1], s.f_ro[2], s.shk.d_useg, s.shk.d_count, s.shk.d_voff,
                     s.shk.d_vals, s.d_lim, s.d_nb, s.d_np, s.ss.cellcnt, s.ss.key, s.ss.slot,
                     s.ss.errflag);
}

// The FP32 kernels deposit J only; rho follows as the mode asks.
RhoDone launch32_deposit(Sim& s, double delt, RhoMode mode)
{
  materialize(s);
  const Geom& g = s.g;
  Dep32       dc{};
  for (int sp = 0; sp < g.Ns; sp++)
    dc.q[sp] = (float)g.q[sp];
  dc.delt = (float)delt;
  dc.dxdt = (float)(g.del[2] / delt);
  dc.dydt = (float)(g.del[1] / delt);
  dc.dzdt = (float)(g.del[0] / delt);
  {
    PhaseTimer pz(s, 4);
    if (mode == RhoMode::Keep)
      launch_zero(s, s.f_J + g.vol, 3 * g.vol);
    else
      launch_zero(s, s.f_J, 4 * g.vol);
  }
  bool march = false;
  {
    PhaseTimer pt(s, 3);
    const int    N    = g.order + 3;
    const size_t mbase = sizeof(double) * 3 * (size_t)N * N * (g.bsy + N - 1) +
                         sizeof(int) * (size_t)g.Ns * (g.bsy + 1);
    // kTileStaged adds the stage of one phase of kst cells (per cell: 3N planes
    // x N x (N-1) floats)
    const int    kst    = std::max(1, std::min(s.stage_cells > 0 ? s.stage_cells : 4, g.bsy));
    const size_t mstage = mbase + sizeof(float) * (size_t)kst * 3 * N * N * (N - 1);
    const size_t wlds  = g.order == 2 ? March32<2>::WLDS : March32<4>::WLDS;
    const int    tile  = s.tile32 != kTileStaged ? s.tile32
                         : (mstage + wlds + 64 <= 65536 ? kTileStaged : kTileF64);
    const size_t mlds  = tile == kTileStaged ? mstage : mbase;
    int zrows, nzr;
    march32_split(g, s.dep_rows ? s.dep_rows : s.march_rows, zrows, nzr);
    const int  gx     = nzr * (g.bsz / g.bsy);
    const int  nitems = gx * g.nchunk;
    int*       wq     = s.d_wq + 2 * kWqDeposit32;
#define PN32_MARCH_T(O, T)                                                                     \
  do {                                                                                         \
    if (s.dep_queue) {                                                                         \
      const int nb = queue_blocks(                                                             \
          s, reinterpret_cast<const void*>(&k32_deposit_march<O, true, T>), mlds, nitems);     \
      hipLaunchKernelGGL((k32_deposit_march<O, true, T>), dim3(nb), dim3(256), mlds, s.stream, \
                         g, s.qa, s.f_ro[0], s.f_ro[1], s.f_ro[2], s.ss.cellstart, s.d_off,     \
                         s.f_J, dc, zrows, wq, gx, nitems, kst);                               \
    } else {                                                                                   \
      hipLaunchKernelGGL((k32_deposit_march<O, false, T>), dim3(gx, g.nchunk), dim3(256),      \
                         mlds, s.stream, g, s.qa, s.f_ro[0], s.f_ro[1], s.f_ro[2],              \
                         s.ss.cellstart, s.d_off, s.f_J, dc, zrows, wq, gx, nitems, kst);      \
    }                                                                                          \
  } while (0)
#define PN32_MARCH(O)                                                                          \
  do {                                                                                         \
    if (tile == kTileF64)                                                                      \
      PN32_MARCH_T(O, kTileF64);                                                               \
    else if (tile == kTilePlain)                                                               \
      PN32_MARCH_T(O, kTilePlain);                                                             \
    else if (tile == kTileComp)                                                                \
      PN32_MARCH_T(O, kTileComp);                                                              \
    else                                                                                       \
      PN32_MARCH_T(O, kTileStaged);                                                            \
  } while (0)
    if (g.dim == 3 && s.sorted && !s.generic_kernels && g.order == 2 &&
        mlds + March32<2>::WLDS + 64 <= 65536) {
      PN32_MARCH(2);
      march = true;
    } else if (g.dim == 3 && s.sorted && !s.generic_kernels && g.order == 4 &&
               mlds + March32<4>::WLDS + 64 <= 65536) {
      PN32_MARCH(4);
      march = true;
#undef PN32_MARCH
#undef PN32_MARCH_T
    } else {
      const dim3 grid(nblk(g.cap, 256), g.nseg);
#define PN32_DEP(D, O)                                                                         \
  hipLaunchKernelGGL((k32_deposit_global<D, O>), grid, dim3(256), 0, s.stream, g, s.qa,        \
                     s.f_ro[0], s.f_ro[1], s.f_ro[2], s.d_np, s.d_off, s.f_J, dc)
      if (g.dim == 3) {
        switch (g.order) {
        case 1: PN32_DEP(3, 1); break;
        case 2: PN32_DEP(3, 2); break;
        case 3: PN32_DEP(3, 3); break;
        default: PN32_DEP(3, 4); break;
        }
      } else if (g.dim == 1) {
        switch (g.order) {
        case 1: PN32_DEP(1, 1); break;
        case 2: PN32_DEP(1, 2); break;
        case 3: PN32_DEP(1, 3); break;
        default: PN32_DEP(1, 4); break;
        }
      } else {
        switch (g.order) {
        case 1: PN32_DEP(2, 1); break;
        case 2: PN32_DEP(2, 2); break;
        case 3: PN32_DEP(2, 3); break;
        default: PN32_DEP(2, 4); break;
        }
      }
#undef PN32_DEP
    }
  }
  (void)march;
  if (mode == RhoMode::Keep)
    return RhoDone::Kept;
  if (mode == RhoMode::Now) {
    launch32_rho(s);
    return RhoDone::Deposited;
  }
  return RhoDone::Deferred;
}

void launch32_rho(Sim& s)
{
  PhaseTimer     pt(s, 7);
  PhaseTimer     pr(s, kPhaseRhoDeposit);
  const Geom&    g    = s.g;
  const int32_t* perm = s.perm_pending ? s.ss.perm : nullptr; // read through by the kernels
  if (g.dim == 3 && s.sorted && !s.generic_kernels && (g.order == 2 || g.order == 4)) {
    const int    S    = g.order + 1;
    const size_t mlds = sizeof(double) * (size_t)S * S * (g.bsy + g.order) +
                        sizeof(int) * (size_t)g.Ns * (g.bsy + 1);
    const size_t wlds = g.order == 2 ? Rho32<2>::WLDS : Rho32<4>::WLDS;
    if (mlds + wlds + 64 <= 65536) {
      int zrows, nzr;
      march32_split(g, s.rho_rows ? s.rho_rows : s.march_rows, zrows, nzr);
      const dim3 mgrid(nzr * (g.bsz / g.bsy), g.nchunk);
      if (g.order == 2)
        hipLaunchKernelGGL(k32_rho_march<2>, mgrid, dim3(256), mlds, s.stream, g, s.qa,
                           s.ss.cellstart, s.d_off, s.f_J, zrows, perm);
      else
        hipLaunchKernelGGL(k32_rho_march<4>, mgrid, dim3(256), mlds, s.stream, g, s.qa,
                           s.ss.cellstart, s.d_off, s.f_J, zrows, perm);
      return;
    }
  }
  const dim3 grid(nblk(g.cap, 256), g.nseg);
#define PN32_RHO(D, O)                                                                        \
  hipLaunchKernelGGL((k32_rho_global<D, O>), grid, dim3(256), 0, s.stream, g, s.qa, s.d_np,    \
                     s.d_off, s.f_J, perm, s.d_nb)
  if (g.dim == 3) {
    switch (g.order) {
    case 1: PN32_RHO(3, 1); break;
    case 2: PN32_RHO(3, 2); break;
    case 3: PN32_RHO(3, 3); break;
    default: PN32_RHO(3, 4); break;
    }
  } else if (g.dim == 1) {
    switch (g.order) {
    case 1: PN32_RHO(1, 1); break;
    case 2: PN32_RHO(1, 2); break;
    case 3: PN32_RHO(1, 3); break;
    default: PN32_RHO(1, 4); break;
    }
  } else {
    switch (g.order) {
    case 1: PN32_RHO(2, 1); break;
    case 2: PN32_RHO(2, 2); break;
    case 3: PN32_RHO(2, 3); break;
    default: PN32_RHO(2, 4); break;
    }
  }
#undef PN32_RHO
}

} // namespace picnix
