// pn_sort.hip -- particle migration between chunks and the per-cell counting sort.
//
// The reference sorts each chunk's particles by the key (cell, ip mod 8, ip)
// (XtensorParticle::sort, nix/xtensor_particle.hpp:260-321) after appending the
// particles received from its 26 neighbours in direction order and, within a
// direction, in the sender's order (XtensorHaloParticle3D, nix/xtensor_halo3d.hpp:
// 214-499; slot order nix/chunk.hpp:426-461).  The GPU version computes, for every
// particle, its destination segment, its index `idx` in the destination's pre-sort
// array (own particles keep ip, arrivals get Np + receive offset + rank) and its
// bin.  Because (cell, idx mod 8, idx) is a total order, the final position is
//   start(cell) + #{members q of the cell : (q mod 8, q) < (idx mod 8, idx)}
// which is computed exactly from an unordered member list, so atomics only hand
// out scratch slots and never decide the result.
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Exclusive prefix over tiles of the per-direction leaver counts of each segment.
__global__ void k_tile_scan(Geom g, const int32_t* __restrict__ np, int32_t* __restrict__ tilecnt,
                            int32_t* __restrict__ nout, int ntile)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nseg * 27)
    return;
  int seg = t / 27, d = t % 27;
  int nt  = (np[seg] + 255) / 256;
  int run = 0;
  for (int k = 0; k < nt; k++) {
    long a      = ((long)seg * ntile + k) * 27 + d;
    int  v      = tilecnt[a];
    tilecnt[a]  = run;
    run        += v;
  }
  nout[seg * 27 + d] = run;
}

// Receive offsets: slot order (iz, iy, ix) lexicographic, skipping the centre;
// a slot's message comes from the neighbour in that direction, which packed it
// for the opposite direction 26 - slot (nix/chunk.hpp:465-524).
__global__ void k_recv_offsets(Geom g, const int32_t* __restrict__ np, const int* __restrict__ nb,
                               const int32_t* __restrict__ nout, int32_t* __restrict__ recvoff)
{
  int segR = blockIdx.x * blockDim.x + threadIdx.x;
  if (segR >= g.nseg)
    return;
  int R = segR / g.Ns, sp = segR - R * g.Ns;
  int run = np[segR];
  for (int slot = 0; slot < 27; slot++) {
    recvoff[segR * 27 + slot] = run;
    if (slot == 13)
      continue;
    int S = nb[R * 27 + slot];
    run += nout[(S * g.Ns + sp) * 27 + (26 - slot)];
  }
}

__device__ __forceinline__ int bin_in(const Geom& g, const double* __restrict__ lim,
                                      const double* __restrict__ rng, int c, double x, double y,
                                      double z)
{
  const int    is_odd  = (g.order % 2 == 1) ? 1 : 0;
  const double xoffset = lim[c * 6 + 4] - 0.5 * g.del[2] * is_odd;
  const double yoffset = lim[c * 6 + 2] - 0.5 * g.del[1] * is_odd;
  const double zoffset = lim[c * 6 + 0] - 0.5 * g.del[0] * is_odd;
  const double rdx = 1 / g.del[2], rdy = 1 / g.del[1], rdz = 1 / g.del[0];
  if (g.has[2] && (x < rng[c * 6 + 4] || x >= rng[c * 6 + 5]))
    return -1;
  if (g.has[1] && (y < rng[c * 6 + 2] || y >= rng[c * 6 + 3]))
    return -1;
  if (g.has[0] && (z < rng[c * 6 + 0] || z >= rng[c * 6 + 1]))
    return -1;
  int ix = g.has[2] ? digitize(x, xoffset, rdx) : 0;
  int iy = g.has[1] ? digitize(y, yoffset, rdy) : 0;
  int iz = g.has[0] ? digitize(z, zoffset, rdz) : 0;
  return iz * g.bsz + iy * g.bsy + ix;
}

// Destination of every particle.  Leavers: rank among the tile's earlier leavers
// with the same direction, receive offset in the neighbour, periodic wrap of the
// position (XtensorParticle::set_boundary_periodic, nix/xtensor_particle.hpp:
// 359-376, applied to arrivals only), and the bin in the receiving chunk.
__global__ void __launch_bounds__(256)
    k_migrate(Geom g, PArr p, const int32_t* __restrict__ np, const double* __restrict__ lim,
              const double* __restrict__ rng, const int* __restrict__ nb,
              const int32_t* __restrict__ tilecnt, const int32_t* __restrict__ recvoff,
              int32_t* __restrict__ cellcnt, int32_t* __restrict__ key,
              int32_t* __restrict__ slot, int32_t* __restrict__ dseg,
              int32_t* __restrict__ didx, int ntile)
{
  __shared__ int dirs[256];
  const int  seg    = blockIdx.y;
  const long ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = ip < np[seg];
  const long i      = (long)seg * g.cap + ip;
  int        k      = active ? key[i] : 0;
  int        d      = (active && k < 0) ? (-1 - k) : -1;
  dirs[threadIdx.x] = d;
  __syncthreads();
  if (!active)
    return;
  if (d < 0) {
    dseg[i] = seg;
    didx[i] = (int32_t)ip;
    return;
  }
  int rank = tilecnt[((long)seg * ntile + blockIdx.x) * 27 + d];
  for (int t = 0; t < (int)threadIdx.x; t++)
    rank += (dirs[t] == d);
  const int c  = seg / g.Ns;
  const int sp = seg - c * g.Ns;
  const int R  = nb[c * 27 + d];
  const int sR = R * g.Ns + sp;
  const int idx = recvoff[sR * 27 + (26 - d)] + rank;

  const double X1 = g.glim[2][0], X2 = g.glim[2][1];
  const double Y1 = g.glim[1][0], Y2 = g.glim[1][1];
  const double Z1 = g.glim[0][0], Z2 = g.glim[0][1];
  const double X  = g.has[2] * (X2 - X1);
  const double Y  = g.has[1] * (Y2 - Y1);
  const double Z  = g.has[0] * (Z2 - Z1);
  double       x = p.x[i], y = p.y[i], z = p.z[i];
  x += (x < X1) * X - (x >= X2) * X;
  y += (y < Y1) * Y - (y >= Y2) * Y;
  z += (z < Z1) * Z - (z >= Z2) * Z;
  p.x[i] = x;
  p.y[i] = y;
  p.z[i] = z;
  int b  = bin_in(g, lim, rng, R, x, y, z);
  if (b < 0) {
    dseg[i] = -1; // outside the receiver as well: discarded by its sort
    return;
  }
  slot[i] = atomicAdd(&cellcnt[(long)sR * g.nbin + b], 1);
  key[i]  = b;
  dseg[i] = sR;
  didx[i] = idx;
}

// Exclusive scan of the bin counts of one segment (one block per segment).
__global__ void __launch_bounds__(256)
    k_cell_scan(Geom g, const int32_t* __restrict__ cellcnt, int32_t* __restrict__ cellstart,
                int32_t* __restrict__ npnew, int32_t* __restrict__ errflag)
{
  __shared__ int part[256];
  const int      seg = blockIdx.x;
  const int      nb  = g.nbin;
  const int      per = (nb + 255) / 256;
  const int      b0  = threadIdx.x * per;
  const int32_t* cnt = cellcnt + (long)seg * nb;
  int            sum = 0;
  for (int b = b0; b < b0 + per && b < nb; b++)
    sum += cnt[b];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int      run   = part[threadIdx.x] - sum;
  int32_t* start = cellstart + (long)seg * (nb + 1);
  for (int b = b0; b < b0 + per && b < nb; b++) {
    start[b] = run;
    run += cnt[b];
  }
  if (threadIdx.x == 255) {
    start[nb]  = part[255];
    npnew[seg] = part[255];
    if (part[255] > g.cap)
      atomicOr(&errflag[0], 1);
  }
}

// Member lists: every bin receives the pre-sort indices of its particles.
__global__ void __launch_bounds__(256)
    k_members(Geom g, const int32_t* __restrict__ np, const int32_t* __restrict__ cellstart,
              const int32_t* __restrict__ key, const int32_t* __restrict__ slot,
              const int32_t* __restrict__ dseg, const int32_t* __restrict__ didx,
              int32_t* __restrict__ mkey)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long i  = (long)seg * g.cap + ip;
  const int  ds = dseg[i];
  if (ds < 0)
    return;
  long pos = (long)cellstart[(long)ds * (g.nbin + 1) + key[i]] + slot[i];
  if (pos < g.cap)
    mkey[(long)ds * g.cap + pos] = didx[i];
}

// Rank inside the bin by (idx mod 8, idx) and move the record.
__global__ void __launch_bounds__(256)
    k_rank_move(Geom g, PArr src, PArr dst, const int32_t* __restrict__ np,
                const int32_t* __restrict__ cellstart, const int32_t* __restrict__ key,
                const int32_t* __restrict__ dseg, const int32_t* __restrict__ didx,
                const int32_t* __restrict__ mkey)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long i  = (long)seg * g.cap + ip;
  const int  ds = dseg[i];
  if (ds < 0)
    return;
  const int32_t* cs   = cellstart + (long)ds * (g.nbin + 1);
  const int      b    = key[i];
  const int      base = cs[b];
  if (base >= g.cap)
    return; // capacity exceeded (flagged by k_cell_scan)
  const int      n    = min(cs[b + 1], (int)g.cap) - base;
  const int      me   = didx[i];
  const int      ml   = me & (kSimdWidth - 1);
  const int32_t* mk   = mkey + (long)ds * g.cap + base;
  int            rank = 0;
  for (int q = 0; q < n; q++) {
    int o  = mk[q];
    int ol = o & (kSimdWidth - 1);
    rank += (ol < ml) || (ol == ml && o < me);
  }
  long dst_i = (long)base + rank;
  if (dst_i >= g.cap)
    return;
  dst_i += (long)ds * g.cap;
  dst.x[dst_i]  = src.x[i];
  dst.y[dst_i]  = src.y[i];
  dst.z[dst_i]  = src.z[i];
  dst.ux[dst_i] = src.ux[i];
  dst.uy[dst_i] = src.uy[i];
  dst.uz[dst_i] = src.uz[i];
  dst.id[dst_i] = src.id[i];
}

void finish_sort(Sim& s)
{
  const Geom& g = s.g;
  hipLaunchKernelGGL(k_cell_scan, dim3(g.nseg), dim3(256), 0, s.stream, g, s.ss.cellcnt,
                     s.ss.cellstart, s.ss.npnew, s.ss.errflag);
  dim3 grid(s.ss.ntile, g.nseg);
  hipLaunchKernelGGL(k_members, grid, dim3(256), 0, s.stream, g, s.d_np, s.ss.cellstart,
                     s.ss.key, s.ss.slot, s.ss.dseg, s.ss.didx, s.ss.mkey);
  hipLaunchKernelGGL(k_rank_move, grid, dim3(256), 0, s.stream, g, s.pa, s.pb, s.d_np,
                     s.ss.cellstart, s.ss.key, s.ss.dseg, s.ss.didx, s.ss.mkey);
  hipMemcpyAsync(s.d_np, s.ss.npnew, sizeof(int32_t) * g.nseg, hipMemcpyDeviceToDevice,
                 s.stream);
  std::swap(s.pa, s.pb);
  s.np_host_valid = false;
  s.counted       = false;
  s.sorted        = true;
}

} // namespace

void launch_migrate(Sim& s)
{
  const Geom& g = s.g;
  hipLaunchKernelGGL(k_tile_scan, dim3(nblk((long)g.nseg * 27, 256)), dim3(256), 0, s.stream, g,
                     s.d_np, s.ss.tilecnt, s.ss.nout, s.ss.ntile);
  hipLaunchKernelGGL(k_recv_offsets, dim3(nblk(g.nseg, 256)), dim3(256), 0, s.stream, g, s.d_np,
                     s.d_nb, s.ss.nout, s.ss.recvoff);
  dim3 grid(s.ss.ntile, g.nseg);
  hipLaunchKernelGGL(k_migrate, grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_lim, s.d_rng,
                     s.d_nb, s.ss.tilecnt, s.ss.recvoff, s.ss.cellcnt, s.ss.key, s.ss.slot,
                     s.ss.dseg, s.ss.didx, s.ss.ntile);
  finish_sort(s);
}

void launch_sort(Sim& s)
{
  launch_count(s);
  finish_sort(s);
}

} // namespace picnix
