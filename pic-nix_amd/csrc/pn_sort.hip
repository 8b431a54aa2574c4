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

// Exclusive prefix over tiles of the per-direction leaver counts of each
// segment: one block per segment.  The counts ([tile][27]) are read and written in
// order, 9 tiles x 27 directions per pass (one element per thread, so every access
// is coalesced); each pass scans its 9 tiles per direction in LDS and carries the
// running sums to the next.
__global__ void __launch_bounds__(256)
    k_tile_scan(Geom g, const int32_t* __restrict__ np, int32_t* __restrict__ tilecnt,
                int32_t* __restrict__ nout, int ntile)
{
  constexpr int TPP = 9; // tiles per pass: 9 x 27 = 243 threads
  __shared__ int v[TPP][27];
  __shared__ int carry[27];
  const int  seg = blockIdx.x;
  const int  nt  = (np[seg] + 255) / 256;
  int32_t*   tc  = tilecnt + (long)seg * ntile * 27;
  const bool on  = threadIdx.x < TPP * 27;
  const int  s   = (int)threadIdx.x / 27;
  const int  d   = (int)threadIdx.x - s * 27;
  if (threadIdx.x < 27)
    carry[threadIdx.x] = 0;
  __syncthreads();
  for (int base = 0; base < nt; base += TPP) {
    const int  t   = base + s;
    const bool in  = on && t < nt;
    const int  val = in ? tc[(long)t * 27 + d] : 0;
    if (on)
      v[s][d] = val;
    __syncthreads();
    if (in) {
      int pre = carry[d];
      for (int q = 0; q < s; q++)
        pre += v[q][d];
      tc[(long)t * 27 + d] = pre;
    }
    __syncthreads();
    if (threadIdx.x < 27) {
      int sum = 0;
      for (int q = 0; q < TPP; q++)
        sum += v[q][threadIdx.x];
      carry[threadIdx.x] += sum;
    }
    __syncthreads(); // carry is read, and v rewritten, by the next pass
  }
  if (threadIdx.x < 27) // carry[d] is only ever written by thread d
    nout[seg * 27 + threadIdx.x] = carry[threadIdx.x];
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
  __shared__ int wcnt[4][27]; // leavers per direction in each wavefront of the tile
  const int  seg    = blockIdx.y;
  const long ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = ip < np[seg];
  const long i      = (long)seg * g.cap + ip;
  int        k      = active ? key[i] : 0;
  int        d      = (active && k < 0) ? (-1 - k) : -1;
  const int  lane   = (int)(threadIdx.x & 63);
  const int  wv     = (int)(threadIdx.x >> 6);
  for (int q = threadIdx.x; q < 4 * 27; q += blockDim.x)
    (&wcnt[0][0])[q] = 0;
  __syncthreads();
  // a leaver's rank in the tile counts the earlier threads with its direction:
  // the earlier lanes of its wavefront (one ballot per direction present in the
  // wavefront), then the earlier wavefronts' totals
  int                rank_w  = 0;
  unsigned long long pending = __ballot(d >= 0);
  while (pending) {
    const int                ld = __builtin_ctzll(pending);
    const int                dd = __shfl(d, ld);
    const unsigned long long m  = __ballot(d == dd);
    if (d == dd)
      rank_w = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == ld)
      wcnt[wv][dd] = __popcll(m);
    pending &= ~m;
  }
  __syncthreads();
  // arrivals: destination segment, pre-sort index, wrapped position and bin
  int  sR = -1, b = -1, idx = 0;
  if (d >= 0) {
    int rank = tilecnt[((long)seg * ntile + blockIdx.x) * 27 + d] + rank_w;
    for (int w = 0; w < wv; w++)
      rank += wcnt[w][d];
    const int c  = seg / g.Ns;
    const int sp = seg - c * g.Ns;
    const int R  = nb[c * 27 + d];
    sR           = R * g.Ns + sp;
    idx          = recvoff[sR * 27 + (26 - d)] + rank;

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
    b      = bin_in(g, lim, rng, R, x, y, z);
  }
  const bool arrive = d >= 0 && b >= 0;
  int        sl     = wave_slot(cellcnt, arrive ? sR * g.nbin + b : 0, arrive);
  if (!active)
    return;
  if (d < 0) {
    dseg[i] = seg;
    didx[i] = (int32_t)ip;
    return;
  }
  if (!arrive) {
    dseg[i] = -1; // outside the receiver as well: discarded by its sort
    return;
  }
  slot[i] = sl;
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
  // On overflow every start and the new count are clamped to the capacity, so
  // that the kernels bounded by them (the record moves here, the next step's
  // gathers, deposits and pushes) stay inside the segment; the flag reports it.
  const int cap = (int)g.cap;
  int       run = part[threadIdx.x] - sum;
  int32_t* start = cellstart + (long)seg * (nb + 1);
  for (int b = b0; b < b0 + per && b < nb; b++) {
    start[b] = min(run, cap);
    run += cnt[b];
  }
  if (threadIdx.x == 255) {
    start[nb]  = min(part[255], cap);
    npnew[seg] = min(part[255], cap);
    if (part[255] > cap)
      atomicOr(&errflag[0], 1);
  }
}

// Total order of the members of one bin: (idx mod 8, idx).
__device__ __forceinline__ unsigned int sort_key(int idx)
{
  return ((unsigned int)(idx & (kSimdWidth - 1)) << 28) | ((unsigned int)idx >> 3);
}

// Member lists: every bin receives the sort key of each of its particles (from
// its pre-sort index) and where the particle's record currently lives.
__global__ void __launch_bounds__(256)
    k_members(Geom g, const int32_t* __restrict__ np, const int32_t* __restrict__ cellstart,
              const int32_t* __restrict__ key, const int32_t* __restrict__ slot,
              const int32_t* __restrict__ dseg, const int32_t* __restrict__ didx,
              int32_t* __restrict__ mkey, int32_t* __restrict__ msrc, int ntile)
{
  // consecutive tiles on one XCD: their members land in neighbouring bins, so
  // the partially written lines of the member lists meet in one L2
  const long nblocks = (long)g.nseg * ntile;
  const long blk     = xcd_block(blockIdx.x, nblocks);
  if (blk >= nblocks)
    return;
  const int  seg = (int)(blk / ntile);
  const long ip  = (blk - (long)seg * ntile) * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long i  = (long)seg * g.cap + ip;
  const int  ds = dseg[i];
  if (ds < 0)
    return;
  long pos = (long)cellstart[(long)ds * (g.nbin + 1) + key[i]] + slot[i];
  if (pos < g.cap) {
    mkey[(long)ds * g.cap + pos] = (int32_t)sort_key(didx[i]);
    msrc[(long)ds * g.cap + pos] = (int32_t)i;
  }
}


// The member list seen through the constant address space: its entries are
// read-only while the rank kernels run, so wave-uniform reads of it become
// scalar loads.
typedef const __attribute__((address_space(4))) int32_t* member_ptr;

// Adds to r0 (and r1 when Two) the number of members in [lo, hi) of a member
// list whose keys are below m0 (m1).  The keys are wave-uniform: they are read
// 16 at a time with scalar loads.  A lane whose member is not in this range
// passes key 0, which no key is below.  The last group reads up to 15 entries
// past hi and ignores them (the member list is allocated with kMemberPad spare
// entries).
template <bool Two>
__device__ __forceinline__ void range_ranks(member_ptr ck, int lo, int hi, unsigned int m0,
                                            unsigned int m1, int& r0, int& r1)
{
  int c1 = lo;
  for (; c1 + 16 <= hi; c1 += 16) {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const unsigned int o = (unsigned int)ck[c1 + j];
      r0 += o < m0;
      if constexpr (Two)
        r1 += o < m1;
    }
  }
  if (c1 < hi) {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const unsigned int o = c1 + j < hi ? (unsigned int)ck[c1 + j] : 0xFFFFFFFFu;
      r0 += o < m0;
      if constexpr (Two)
        r1 += o < m1;
    }
  }
}

struct Rec64 {
  double  x, y, z, ux, uy, uz;
  int64_t id;
};

__device__ __forceinline__ Rec64 load_rec(const PArr& p, long i)
{
  return Rec64{p.x[i], p.y[i], p.z[i], p.ux[i], p.uy[i], p.uz[i], p.id[i]};
}

__device__ __forceinline__ void store_rec(const PArr& p, long i, const Rec64& r)
{
  p.x[i]  = r.x;
  p.y[i]  = r.y;
  p.z[i]  = r.z;
  p.ux[i] = r.ux;
  p.uy[i] = r.uy;
  p.uz[i] = r.uz;
  p.id[i] = r.id;
}

// Rank the members of bins b and b + 1 of segment seg (b even; b + 1 may be
// past the last bin) by (idx mod 8, idx) -- the order of XtensorParticle::sort
// -- each within its own bin, and move each record to start(bin) + rank
// (Defer: record only where it comes from, perm[start(bin) + rank] = source).
// The two bins' members are contiguous in the member list, so one wavefront
// covers both with one member per lane (two per lane per pass of 128
// members); all lanes call.
template <bool Defer>
__device__ __forceinline__ void rank_bin_pair(const Geom& g, const PArr& src, const PArr& dst,
                                              const int32_t* __restrict__ cellstart,
                                              const int32_t* __restrict__ mkey,
                                              const int32_t* __restrict__ msrc,
                                              int32_t* __restrict__ perm, int seg, int b,
                                              int lane)
{
  const int32_t* cs   = cellstart + (long)seg * (g.nbin + 1);
  const int      cap  = (int)g.cap;
  const int      base = min(cs[b], cap);
  const int      mid  = min(cs[b + 1], cap);
  const int      end  = b + 1 < g.nbin ? min(cs[b + 2], cap) : mid;
  if (base >= end)
    return;
  const int        n  = end - base;
  const int        nA = mid - base;
  const int32_t*   mk = mkey + (long)seg * g.cap + base;
  const int32_t*   ms = msrc + (long)seg * g.cap + base;
  const member_ptr ck = (member_ptr)mk;
  const long       d0 = (long)seg * g.cap;
  for (int c0 = 0; c0 < n; c0 += 128) {
    const int          q0 = c0 + lane, q1 = q0 + 64;
    const unsigned int m0 = q0 < n ? (unsigned int)mk[q0] : 0xFFFFFFFFu;
    const unsigned int m1 = q1 < n ? (unsigned int)mk[q1] : 0xFFFFFFFFu;
    // the records are loaded before the ranking, which overlaps their latency
    Rec64   a{}, c{};
    int32_t s0 = 0, s1 = 0;
    if (q0 < n) {
      s0 = ms[q0];
      if constexpr (!Defer)
        a = load_rec(src, s0);
    }
    if (q1 < n) {
      s1 = ms[q1];
      if constexpr (!Defer)
        c = load_rec(src, s1);
    }
    // each member is ranked among its own bin's members only
    const bool inA0 = q0 < nA, inA1 = q1 < nA;
    int        r0 = 0, r1 = 0;
    if (c0 + 64 < n) {
      // two members per lane: each range is compared only with the members that
      // can lie in it (wave-uniform tests), e.g. for two bins of 64 the first
      // members are all in bin b and the second all in bin b + 1
      const bool a0 = c0 < nA, a1 = c0 + 64 < nA;        // some q0 / q1 in bin b
      const bool b0 = c0 + 63 >= nA, b1 = c0 + 127 >= nA; // some q0 / q1 in bin b + 1
      int        none = 0;
      if (a0 && a1)
        range_ranks<true>(ck, 0, nA, inA0 ? m0 : 0u, inA1 ? m1 : 0u, r0, r1);
      else if (a0)
        range_ranks<false>(ck, 0, nA, inA0 ? m0 : 0u, 0u, r0, none);
      else if (a1)
        range_ranks<false>(ck, 0, nA, inA1 ? m1 : 0u, 0u, r1, none);
      if (b0 && b1)
        range_ranks<true>(ck, nA, n, inA0 ? 0u : m0, inA1 ? 0u : m1, r0, r1);
      else if (b0)
        range_ranks<false>(ck, nA, n, inA0 ? 0u : m0, 0u, r0, none);
      else if (b1)
        range_ranks<false>(ck, nA, n, inA1 ? 0u : m1, 0u, r1, none);
    } else {
      range_ranks<false>(ck, 0, nA, inA0 ? m0 : 0u, 0u, r0, r1);
      range_ranks<false>(ck, nA, n, inA0 ? 0u : m0, 0u, r0, r1);
    }
    if constexpr (Defer) {
      if (q0 < n)
        perm[d0 + (inA0 ? base : mid) + r0] = s0;
      if (q1 < n)
        perm[d0 + (inA1 ? base : mid) + r1] = s1;
    } else {
      if (q0 < n)
        store_rec(dst, d0 + (inA0 ? base : mid) + r0, a);
      if (q1 < n)
        store_rec(dst, d0 + (inA1 ? base : mid) + r1, c);
    }
  }
}

// A block owns kRankPairs consecutive bin pairs of one segment, one per
// wavefront; consecutive blocks run on one XCD (xcd_block), so records moving
// between neighbouring bins are read through the same L2.
constexpr int kRankPairs = 4;

template <bool Defer>
__global__ void __launch_bounds__(256)
    k_rank_bin(Geom g, PArr src, PArr dst, const int32_t* __restrict__ cellstart,
               const int32_t* __restrict__ mkey, const int32_t* __restrict__ msrc,
               int32_t* __restrict__ perm, int npairs, long nblocks)
{
  const long blk = xcd_block(blockIdx.x, nblocks);
  if (blk >= nblocks)
    return;
  const int lane = threadIdx.x & 63;
  // the wavefront index, marked wave-uniform so that the member lists are
  // read with scalar loads
  const int  wv   = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long pair = blk * kRankPairs + wv; // segment-major: npairs per segment
  if (pair >= (long)g.nseg * npairs)
    return;
  const int seg = (int)(pair / npairs);
  const int b   = 2 * (int)(pair - (long)seg * npairs);
  rank_bin_pair<Defer>(g, src, dst, cellstart, mkey, msrc, perm, seg, b, lane);
}

// The record move of a deferred sort: sorted slot i of every segment takes
// record perm[i] (any segment) of src.
__global__ void __launch_bounds__(256)
    k_apply_perm(Geom g, PArr src, PArr dst, const int32_t* __restrict__ np,
                 const int32_t* __restrict__ perm)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long i = (long)seg * g.cap + ip;
  store_rec(dst, i, load_rec(src, perm[i]));
}

// defer: the record move is left to the next row gather (perm_pending)
void finish_sort(Sim& s, bool defer = false)
{
  const Geom& g = s.g;
  hipLaunchKernelGGL(k_cell_scan, dim3(g.nseg), dim3(256), 0, s.stream, g, s.ss.cellcnt,
                     s.ss.cellstart, s.ss.npnew, s.ss.errflag);
  hipLaunchKernelGGL(k_members, dim3(xcd_grid((long)g.nseg * s.ss.ntile)), dim3(256), 0, s.stream,
                     g, s.d_np, s.ss.cellstart, s.ss.key, s.ss.slot, s.ss.dseg, s.ss.didx,
                     s.ss.mkey, s.ss.msrc, s.ss.ntile);
  const int  npairs = (g.nbin + 1) / 2;
  const long nrb    = nblk((long)g.nseg * npairs, kRankPairs);
  if (defer)
    hipLaunchKernelGGL(k_rank_bin<true>, dim3(xcd_grid(nrb)), dim3(256), 0, s.stream, g, s.pa,
                       s.pb, s.ss.cellstart, s.ss.mkey, s.ss.msrc, s.ss.perm, npairs, nrb);
  else
    hipLaunchKernelGGL(k_rank_bin<false>, dim3(xcd_grid(nrb)), dim3(256), 0, s.stream, g, s.pa,
                       s.pb, s.ss.cellstart, s.ss.mkey, s.ss.msrc, s.ss.perm, npairs, nrb);
  launch_copy(s, s.d_np, s.ss.npnew, g.nseg);
  if (defer)
    s.perm_pending = true;
  else
    std::swap(s.pa, s.pb);
  s.np_host_valid = false;
  s.counted       = false;
  s.sorted        = true;
}

// ---- FP32 form (cell-relative particles, struct P32) ----

__device__ __forceinline__ void cell3_of(const Geom& g, int bin, int& x, int& y, int& z)
{
  z           = bin / g.bsz;
  const int r = bin - z * g.bsz;
  y           = r / g.bsy;
  x           = r - y * g.bsy;
}

// count(reset) for the FP32 form: the bin of cell + floor(r); outside the chunk
// (or never placed in it) -> dropped by the sort.
__global__ void __launch_bounds__(256)
    k32_count(Geom g, P32 p, const int32_t* __restrict__ np, int32_t* __restrict__ cellcnt,
              int32_t* __restrict__ key, int32_t* __restrict__ slot)
{
  const int  seg    = blockIdx.y;
  const long ip     = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = ip < np[seg];
  const long i      = (long)seg * g.cap + ip;
  int        b      = -1;
  if (active && p.cell[i] >= 0) {
    int cx, cy, cz;
    cell3_of(g, p.cell[i], cx, cy, cz);
    cx += g.has[2] ? (int)floorf(p.rx[i]) : 0;
    cy += g.has[1] ? (int)floorf(p.ry[i]) : 0;
    cz += g.has[0] ? (int)floorf(p.rz[i]) : 0;
    if (cx >= 0 && cx < g.dims[2] && cy >= 0 && cy < g.dims[1] && cz >= 0 && cz < g.dims[0])
      b = cz * g.bsz + cy * g.bsy + cx;
  }
  const bool stay = active && b >= 0;
  int        sl   = wave_slot(cellcnt + (long)seg * g.nbin, b, stay);
  if (!active)
    return;
  if (stay) {
    slot[i] = sl;
    key[i]  = b;
  } else {
    key[i] = -1; // dropped
  }
}

struct Rec32 {
  float   rx, ry, rz, ux, uy, uz;
  int64_t id;
};

__device__ __forceinline__ Rec32 load_rec32(const P32& p, long i)
{
  return Rec32{p.rx[i], p.ry[i], p.rz[i], p.ux[i], p.uy[i], p.uz[i], p.id[i]};
}

__device__ __forceinline__ void store_rec32(const Geom& g, const P32& p, long i, const Rec32& r,
                                            int b)
{
  p.rx[i]   = g.has[2] ? frac01(r.rx) : r.rx;
  p.ry[i]   = g.has[1] ? frac01(r.ry) : r.ry;
  p.rz[i]   = g.has[0] ? frac01(r.rz) : r.rz;
  p.ux[i]   = r.ux;
  p.uy[i]   = r.uy;
  p.uz[i]   = r.uz;
  p.id[i]   = r.id;
  p.cell[i] = b;
}

// The FP32 form's sort: every record goes straight to start(bin) + the slot
// its count handed out, with its offset folded back into [0, 1) and its cell
// set to the bin.  One pass, no member lists and no ranking: inside a bin the
// order is the order of the slot atomics, not the reference's (idx mod 8, idx)
// (the FP32 form is compared with the reference particle by particle, by id).
// Defer: the record stays where it is; perm[destination] = its index (the
// next row gather, or materialize, moves it).
template <bool Defer>
__global__ void __launch_bounds__(256)
    k32_scatter(Geom g, P32 src, P32 dst, const int32_t* __restrict__ np,
                const int32_t* __restrict__ cellstart, const int32_t* __restrict__ key,
                const int32_t* __restrict__ slot, const int* __restrict__ nb, int ntile,
                int32_t* __restrict__ perm)
{
  // consecutive tiles on one XCD (as in k_members): records moving to
  // neighbouring bins meet in one L2
  const long nblocks = (long)g.nseg * ntile;
  const long blk     = xcd_block(blockIdx.x, nblocks);
  if (blk >= nblocks)
    return;
  const int  seg = (int)(blk / ntile);
  const long ip  = (blk - (long)seg * ntile) * blockDim.x + threadIdx.x;
  if (ip >= np[seg])
    return;
  const long i = (long)seg * g.cap + ip;
  const int  k = key[i];
  if (k == -1)
    return; // dropped
  // key >= 0: the bin in this segment; else -2 - (d * nbin + bin) in the
  // neighbour chunk in direction d
  int ds = seg, b = k;
  if (k < -1) {
    const int v  = -2 - k;
    const int d  = v / g.nbin;
    const int c  = seg / g.Ns;
    b            = v - d * g.nbin;
    ds           = nb[c * 27 + d] * g.Ns + (seg - c * g.Ns);
  }
  const long pos = (long)cellstart[(long)ds * (g.nbin + 1) + b] + slot[i];
  if (pos >= g.cap)
    return; // overflow, flagged by k_cell_scan
  if constexpr (Defer)
    perm[(long)ds * g.cap + pos] = (int32_t)i;
  else
    store_rec32(g, dst, (long)ds * g.cap + pos, load_rec32(src, i), b);
}

void finish_sort32(Sim& s, bool defer = false)
{
  const Geom& g = s.g;
  hipLaunchKernelGGL(k_cell_scan, dim3(g.nseg), dim3(256), 0, s.stream, g, s.ss.cellcnt,
                     s.ss.cellstart, s.ss.npnew, s.ss.errflag);
  const dim3 grid(xcd_grid((long)g.nseg * s.ss.ntile));
  if (defer)
    hipLaunchKernelGGL(k32_scatter<true>, grid, dim3(256), 0, s.stream, g, s.qa, s.qb, s.d_np,
                       s.ss.cellstart, s.ss.key, s.ss.slot, s.d_nb, s.ss.ntile, s.ss.perm);
  else
    hipLaunchKernelGGL(k32_scatter<false>, grid, dim3(256), 0, s.stream, g, s.qa, s.qb, s.d_np,
                       s.ss.cellstart, s.ss.key, s.ss.slot, s.d_nb, s.ss.ntile, s.ss.perm);
  launch_copy(s, s.d_np, s.ss.npnew, g.nseg);
  if (defer)
    s.perm_pending = true;
  else
    std::swap(s.qa, s.qb);
  s.np_host_valid = false;
  s.counted       = false;
  s.sorted        = true;
}

} // namespace

void materialize(Sim& s)
{
  if (!s.perm_pending)
    return;
  if (s.fp32) {
    materialize32(s);
    return;
  }
  const Geom& g = s.g;
  hipLaunchKernelGGL(k_apply_perm, dim3(s.ss.ntile, g.nseg), dim3(256), 0, s.stream, g, s.pa,
                     s.pb, s.d_np, s.ss.perm);
  std::swap(s.pa, s.pb);
  s.perm_pending = false;
}

void launch_migrate(Sim& s, bool defer)
{
  materialize(s);
  PhaseTimer  pt(s, 6);
  const Geom& g = s.g;
  hipLaunchKernelGGL(k_tile_scan, dim3(g.nseg), dim3(256), 0, s.stream, g, s.d_np,
                     s.ss.tilecnt, s.ss.nout, s.ss.ntile);
  hipLaunchKernelGGL(k_recv_offsets, dim3(nblk(g.nseg, 256)), dim3(256), 0, s.stream, g, s.d_np,
                     s.d_nb, s.ss.nout, s.ss.recvoff);
  dim3 grid(s.ss.ntile, g.nseg);
  hipLaunchKernelGGL(k_migrate, grid, dim3(256), 0, s.stream, g, s.pa, s.d_np, s.d_lim, s.d_rng,
                     s.d_nb, s.ss.tilecnt, s.ss.recvoff, s.ss.cellcnt, s.ss.key, s.ss.slot,
                     s.ss.dseg, s.ss.didx, s.ss.ntile);
  finish_sort(s, defer);
}

void launch_sort(Sim& s)
{
  materialize(s);
  PhaseTimer pt(s, 6);
  launch_count(s);
  finish_sort(s);
}

// The FP32 position push already gave every particle its destination segment,
// bin and slot, leavers included, so the migration + sort is the bin scan and
// the record scatter.
void launch32_migrate(Sim& s, bool defer)
{
  materialize(s);
  PhaseTimer pt(s, 6);
  finish_sort32(s, defer);
}

void launch32_sort(Sim& s)
{
  materialize(s);
  PhaseTimer  pt(s, 6);
  const Geom& g = s.g;
  launch_zero(s, s.ss.cellcnt, (long)g.nseg * g.nbin);
  hipLaunchKernelGGL(k32_count, dim3(s.ss.ntile, g.nseg), dim3(256), 0, s.stream, g, s.qa,
                     s.d_np, s.ss.cellcnt, s.ss.key, s.ss.slot);
  finish_sort32(s);
}

} // namespace picnix
