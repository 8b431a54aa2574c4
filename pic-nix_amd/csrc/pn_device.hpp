// pn_device.hpp -- device-side numerics shared by the gfx950 kernels.
//
// Every routine evaluates the same expression, in the same order, as the
// reference routine cited next to it.  The FP64 kernels are compiled with
// -ffp-contract=off, so each operation is one IEEE rounding and the results are
// bit-identical to the reference algorithm evaluated without contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace picnix {

constexpr int kSimdWidth = 8; // lane sub-bins of the counting sort (nix/nix.hpp:102-105)
constexpr int kMaxSpecies = 8;

// nix/primitives.hpp:46-58
template <typename T>
__device__ __forceinline__ int digitize(T x, T xmin, T rdx)
{
  return static_cast<int>(floor((x - xmin) * rdx));
}

// nix/primitives.hpp:157-161
template <typename T>
__device__ __forceinline__ T lorentz_factor(T ux, T uy, T uz, T rc)
{
  return sqrt(T(1) + (ux * ux + uy * uy + uz * uz) * rc * rc);
}

// nix/primitives.hpp:163-189
template <typename T>
__device__ __forceinline__ void push_boris(T& ux, T& uy, T& uz, T ex, T ey, T ez, T bx, T by,
                                           T bz, T cc)
{
  ux += ex;
  uy += ey;
  uz += ez;
  T gm = T(1) / sqrt(cc * cc + ux * ux + uy * uy + uz * uz);
  bx *= gm;
  by *= gm;
  bz *= gm;
  T bb = T(2.0) / (T(1.0) + bx * bx + by * by + bz * bz);
  T vx = ux + (uy * bz - uz * by);
  T vy = uy + (uz * bx - ux * bz);
  T vz = uz + (ux * by - uy * bx);
  ux += (vy * bz - vz * by) * bb + ex;
  uy += (vz * bx - vx * bz) * bb + ey;
  uz += (vx * by - vy * bx) * bb + ez;
}

// Vay (2008) pusher, nix/primitives.hpp:191-220 (gamma of Vay's Eq. 11)
template <typename T>
__device__ __forceinline__ void push_vay(T& ux, T& uy, T& uz, T ex, T ey, T ez, T bx, T by, T bz,
                                         T cc)
{
  T gm = T(1) / sqrt(cc * cc + ux * ux + uy * uy + uz * uz);
  T vx = ux + 2 * ex + gm * (uy * bz - uz * by);
  T vy = uy + 2 * ey + gm * (uz * bx - ux * bz);
  T vz = uz + 2 * ez + gm * (ux * by - uy * bx);
  gm   = (cc * cc + vx * vx + vy * vy + vz * vz);
  T bb = bx * bx + by * by + bz * bz;
  T bu = bx * vx + by * vy + bz * vz;
  T xx = gm - bb;
  T yy = bb + bu * bu;
  gm   = T(1) / sqrt(T(0.5) * (xx + sqrt(xx * xx + 4 * yy)));
  bx *= gm;
  by *= gm;
  bz *= gm;
  bu = bx * vx + by * vy + bz * vz;
  bb = T(1.0) / (T(1.0) + bx * bx + by * by + bz * bz);
  ux = (vx + bu * bx + (vy * bz - vz * by)) * bb;
  uy = (vy + bu * by + (vz * bx - vx * bz)) * bb;
  uz = (vz + bu * bz + (vx * by - vy * bx)) * bb;
}

// Higuera-Cary (2017) pusher, nix/primitives.hpp:222-253 (gamma of their Eq. 20)
template <typename T>
__device__ __forceinline__ void push_higuera_cary(T& ux, T& uy, T& uz, T ex, T ey, T ez, T bx,
                                                  T by, T bz, T cc)
{
  ux += ex;
  uy += ey;
  uz += ez;
  T gm = cc * cc + ux * ux + uy * uy + uz * uz;
  T bb = bx * bx + by * by + bz * bz;
  T bu = bx * ux + by * uy + bz * uz;
  T xx = gm - bb;
  T yy = bb + bu * bu;
  gm   = T(1) / sqrt(T(0.5) * (xx + sqrt(xx * xx + 4 * yy)));
  bx *= gm;
  by *= gm;
  bz *= gm;
  bb   = T(2.0) / (T(1.0) + bx * bx + by * by + bz * bz);
  T vx = ux + (uy * bz - uz * by);
  T vy = uy + (uz * bx - ux * bz);
  T vz = uz + (ux * by - uy * bx);
  ux += (vy * bz - vz * by) * bb + ex;
  uy += (vz * bx - vx * bz) * bb + ey;
  uz += (vx * by - vy * bx) * bb + ez;
}

// BaseVelocity::push_impl (pic/engine/velocity.hpp:81-99): 0 Boris, 1 Vay,
// 2 Higuera-Cary; `pusher` is uniform across a launch.
template <typename T>
__device__ __forceinline__ void push_momentum(int pusher, T& ux, T& uy, T& uz, const T eb[6], T cc)
{
  if (pusher == 0)
    push_boris(ux, uy, uz, eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
  else if (pusher == 1)
    push_vay(ux, uy, uz, eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
  else
    push_higuera_cary(ux, uy, uz, eb[0], eb[1], eb[2], eb[3], eb[4], eb[5], cc);
}

// nix/primitives.hpp:256-328 (shape_mc1..4)
template <int Order, typename T>
__device__ __forceinline__ void shape_mc(T x, T X, T rdx, T* s)
{
  T delta = (x - X) * rdx;
  if constexpr (Order == 1) {
    s[0] = T(1) - delta;
    s[1] = delta;
  } else if constexpr (Order == 2) {
    T w0 = delta;
    T w1 = T(0.5) - w0;
    T w2 = T(0.5) + w0;
    s[0] = T(0.50) * w1 * w1;
    s[1] = T(0.75) - w0 * w0;
    s[2] = T(0.50) * w2 * w2;
  } else if constexpr (Order == 3) {
    const T a  = T(1) / T(6.0);
    T       w1 = delta;
    T       w2 = T(1) - delta;
    T       p1 = w1 * w1;
    T       p2 = w2 * w2;
    T       q1 = p1 * w1;
    T       q2 = p2 * w2;
    s[0]       = a * q2;
    s[1]       = a * (T(4) - T(6) * p1 + T(3) * q1);
    s[2]       = a * (T(4) - T(6) * p2 + T(3) * q2);
    s[3]       = a * q1;
  } else {
    const T a  = T(1) / T(384.0);
    const T b  = T(1) / T(96.0);
    const T c  = T(115) / T(192.0);
    const T d  = T(1) / T(8.0);
    T       w1 = T(1) + delta;
    T       w2 = T(1) - delta;
    T       w3 = T(1) + delta * T(2);
    T       w4 = T(1) - delta * T(2);
    T       p0 = delta * delta;
    T       p1 = w1 * w1;
    T       p2 = w2 * w2;
    T       q1 = p1 * w1;
    T       q2 = p2 * w2;
    T       r1 = q1 * w1;
    T       r2 = q2 * w2;
    T       r3 = w3 * w3 * w3 * w3;
    T       r4 = w4 * w4 * w4 * w4;
    s[0]       = a * r4;
    s[1]       = b * (T(55) + T(20) * w1 - T(120) * p1 + T(80) * q1 - T(16) * r1);
    s[2]       = c + d * p0 * (T(2) * p0 - T(5));
    s[3]       = b * (T(55) + T(20) * w2 - T(120) * p2 + T(80) * q2 - T(16) * r2);
    s[4]       = a * r3;
  }
}

template <typename T>
__device__ __forceinline__ T wt_abs(T v)
{
  return fabs(v);
}

// WT shape functions (nix/primitives.hpp:331-566; Lu et al., J. Comput. Phys. 413,
// 109388 (2020)): the B-spline of degree Order-1 smeared by a top hat of half
// width dt = c dt / dx; rdt = 1 / dt.  The branch factors t1..t4 combine the
// left/middle/right pieces exactly as the reference does.
template <int Order, typename T>
__device__ __forceinline__ void shape_wt(T x, T X, T rdx, T dt, T rdt, T* s)
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

// Slot allocation with one atomic per distinct counter in the wave: lanes that
// want a slot from the same counter share a single atomicAdd and take
// consecutive slots in lane order.  The groups are formed first (ballots and
// shuffles only); then every group's lowest lane issues its atomic in the same
// instruction, so the wave waits for one memory round trip, not one per group.
// Every active lane of the wave must call this; lanes with want = false get -1.
__device__ __forceinline__ int wave_slot(int32_t* counters, int key, bool want)
{
  const int          lane    = (int)__lane_id();
  unsigned long long pending = __ballot(want);
  int                leader  = lane; // lowest lane of this lane's group
  int                rank    = 0;    // place in the group
  int                count   = 0;    // group size (on the leader)
  while (pending) {
    const int          ld = __builtin_ctzll(pending);
    const int          lk = __shfl(key, ld);
    unsigned long long m  = __ballot(want && key == lk);
    if ((m >> lane) & 1ull) {
      leader = ld;
      rank   = __popcll(m & ((1ull << lane) - 1ull));
    }
    if (lane == ld)
      count = __popcll(m);
    pending &= ~m;
  }
  int base = 0;
  if (want && leader == lane)
    base = atomicAdd(&counters[key], count);
  base = __shfl(base, leader);
  return want ? base + rank : -1;
}

// Workgroups of a launch are dispatched to the 8 XCDs of an MI355X round-robin
// (id % 8), and each XCD has its own L2.  xcd_block maps the hardware block id
// of a grid of xcd_grid(nblocks) blocks so that every XCD runs one contiguous
// range of logical ids: blocks that touch neighbouring data share an L2.  The
// result is >= nblocks for the few padding blocks, which must exit.
constexpr int kNumXcd = 8;
inline long xcd_grid(long nblocks) { return (nblocks + kNumXcd - 1) / kNumXcd * kNumXcd; }
__device__ __forceinline__ long xcd_block(long hw, long nblocks)
{
  const long per = (nblocks + kNumXcd - 1) / kNumXcd;
  return (hw % kNumXcd) * per + hw / kNumXcd;
}

// Device view of the domain: global padded periodic grid + chunk tables.
struct Geom {
  int    dim;      // 1, 2 or 3
  int    order;    // shape order
  int    Nb;       // ghost margin (order + 3) / 2
  int    Ns;       // species
  int    N[3];     // global interior cells (z, y, x)
  int    C[3];     // chunk counts (z, y, x)
  int    dims[3];  // chunk interior cells (z, y, x)
  int    pad[3];   // padding of the global arrays (Nb, or 0 for an absent z)
  int    P[3];     // padded extents (z, y, x)
  long   plane;    // P[1] * P[2]
  long   vol;      // P[0] * P[1] * P[2]
  double del[3];   // cell sizes (z, y, x)
  double glim[3][2];
  double cc;
  int    nbin;     // sort bins per chunk ((nz+1)(ny+1)(nx+1), or 2D analogue)
  int    bsy, bsz; // flat-bin strides (nix/xtensor_particle.hpp:231-240)
  int    nchunk;
  int    nseg;     // nchunk * Ns
  long   cap;      // particle capacity of one (chunk, species) segment
  double q[kMaxSpecies];
  double m[kMaxSpecies];
  bool   has[3];
  bool   xwall;    // x is not periodic: a wall at x = 0 and an open end at x = Lx
};

// key of a particle that left through the open end x = Lx and waits for its
// replacement (the shock problem's upstream boundary); below every leaver code
// of both forms (FP64: -1 - direction, FP32: -2 - (direction * nbin + bin))
constexpr int kKeyReplace = -2147483647 - 1;

// Bin index of a position inside chunk c, or -1 when the particle lies outside
// the chunk.  XtensorParticle::count (nix/xtensor_particle.hpp:324-357).
__device__ __forceinline__ int bin_of(const Geom& g, const double* __restrict__ lim,
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

// Work queue of the persistent row kernels: wq[0] is the next item, wq[1]
// counts the blocks that have drained the queue.  Thread 0 takes an item with a
// global atomic and hands it to the block through LDS; the last block to drain
// the queue resets both counters, so they are zero again for the next launch
// (replays of a captured graph included) without a separate reset.
__device__ __forceinline__ int wq_next(int* wq, int* item)
{
  if (threadIdx.x == 0)
    *item = atomicAdd(&wq[0], 1);
  __syncthreads();
  return *item;
}

__device__ __forceinline__ void wq_finish(int* wq)
{
  if (threadIdx.x == 0 && atomicAdd(&wq[1], 1) == (int)(gridDim.x * gridDim.y) - 1) {
    atomicExch(&wq[0], 0);
    atomicExch(&wq[1], 0);
  }
}

// Exchange direction slot (0..26, 13 = stays) of XtensorHaloParticle3D::pre_pack
// (nix/xtensor_halo3d.hpp:230-244).
__device__ __forceinline__ int dir_of(const double* __restrict__ rng, int c, double x, double y,
                                      double z)
{
  int iz = (z >= rng[c * 6 + 1]) - (z < rng[c * 6 + 0]) + 1;
  int iy = (y >= rng[c * 6 + 3]) - (y < rng[c * 6 + 2]) + 1;
  int ix = (x >= rng[c * 6 + 5]) - (x < rng[c * 6 + 4]) + 1;
  return 9 * iz + 3 * iy + ix;
}



// Node offsets of a plane lane of the march deposits.  The block's ring tile
// is [3 components][N z-planes][N][TX] (z ring slot, y, x), its lowest z-plane
// in ring slot zs; the lane holds component comp at index t along that
// component's axis, and its accumulator entry (r, j) is the node
//   Jx: z = r,     y = t,     x = j + 1
//   Jy: z = t,     y = j + 1, x = r
//   Jz: z = j + 1, y = r,     x = t
// relative to the cell bx's tile origin.  Its offset is ro[r] + co[j]; the
// wrap of the z ring is folded into the two tables, so the flush adds them
// with no per-node select.
template <int N>
__device__ __forceinline__ void plane_offsets(int comp, int t, int zs, int TX, int psz, int tsz,
                                              int bx, int (&ro)[N], int (&co)[N - 1])
{
  const int lim  = N - zs; // z offsets >= lim wrap to the start of the ring
  const int wrap = N * psz;
  const int zt   = zs + t >= N ? zs + t - N : zs + t;
  const int base = comp == 0 ? zs * psz + t * TX
                             : (comp == 1 ? tsz + zt * psz : 2 * tsz + zs * psz + t);
  const int sr   = comp == 0 ? psz : (comp == 1 ? 1 : TX);
  const int sj   = comp == 0 ? 1 : (comp == 1 ? TX : psz);
#pragma unroll
  for (int r = 0; r < N; r++)
    ro[r] = base + bx + r * sr - ((comp == 0 && r >= lim) ? wrap : 0);
#pragma unroll
  for (int j = 0; j < N - 1; j++)
    co[j] = (j + 1) * sj - ((comp == 2 && j + 1 >= lim) ? wrap : 0);
}

// An FP32 offset folded back into [0, 1) (the largest float below 1 when the
// subtraction rounds up to 1).
__device__ __forceinline__ float frac01(float r)
{
  const float f = r - floorf(r);
  return f < 1.0f ? f : 0.99999994f;
}

// FP32 particle storage (precision = "fp32"): position as the cell index (the
// chunk-local bin, packed like the sort bins) plus the offset inside the cell
// in cell units, r in [0, 1) after every sort; between the position push and
// the sort, r holds the new position relative to the old cell (in [-1, 2)).
struct P32 {
  float*   rx;
  float*   ry;
  float*   rz;
  float*   ux;
  float*   uy;
  float*   uz;
  int64_t* id;
  int32_t* cell;
};

// Structure-of-arrays particle storage; segment s occupies [s * cap, (s + 1) * cap).
struct PArr {
  double*  x;
  double*  y;
  double*  z;
  double*  ux;
  double*  uy;
  double*  uz;
  int64_t* id;
};

} // namespace picnix
