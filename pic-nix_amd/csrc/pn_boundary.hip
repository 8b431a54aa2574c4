// pn_boundary.hip -- physical x boundaries of the shock problem (example/shock) on
// the global padded grid: the wall and upstream field values, the particles that
// cross x = Lx, the appended injection particles and the moment-based kinetic
// energy.
//
// With x not periodic, the x padding of the global arrays plays the part of the
// wall chunks' x ghosts: it is never filled from the periodic image (k_fill_ghost
// leaves it alone) and receives the values MainChunk::set_boundary_emf writes
// (example/shock/main.cpp:232-340).
#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

__device__ __forceinline__ long gidx(const Geom& g, int z, int y, int x)
{
  return ((long)z * g.P[1] + y) * g.P[2] + x;
}

// MainChunk::set_boundary_emf on every (z, y) row of the padded arrays, y and z
// padding included.  Lower wall (x columns [0, 2 Nb), the lower ghosts and the
// first Nb cells): E = 0 and B copied from column 2 Nb.  Upper end (columns
// [N, N + 2 Nb), the last Nb cells and the upper ghosts): Ey, Ez, Bx fixed from
// column N, Ex, By, Bz from column N + 1.  The reference leaves some ghost
// rows of its per-chunk arrays out of these ranges (Ex at the top ghost row, Bx
// at the bottom one); in the global arrays those rows are another chunk's cells,
// which its own call sets.  One thread per (row, column) of the two wall blocks.
template <typename T>
__global__ void __launch_bounds__(256) k_wall_emf(Geom g, T* __restrict__ E, T* __restrict__ B,
                                                  T ex0, T ey0, T ez0, T bx0, T by0, T bz0)
{
  const int  w    = 2 * g.Nb;
  const long rows = (long)g.P[0] * g.P[1];
  const long t    = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * 2 * w)
    return;
  const long row  = t / (2 * w);
  const int  k    = (int)(t % (2 * w));
  const int  z    = (int)(row / g.P[1]);
  const int  y    = (int)(row % g.P[1]);
  const long v    = g.vol;
  if (k < w) {
    const int  x   = k;
    const long i   = gidx(g, z, y, x);
    const long src = gidx(g, z, y, w);
    E[i]         = 0;
    E[v + i]     = 0;
    E[2 * v + i] = 0;
    B[i]         = B[src];
    B[v + i]     = B[v + src];
    B[2 * v + i] = B[2 * v + src];
  } else {
    const int  x = g.N[2] + (k - w);
    const long i = gidx(g, z, y, x);
    E[v + i]     = ey0;
    E[2 * v + i] = ez0;
    B[i]         = bx0;
    if (x >= g.N[2] + 1) {
      E[i]         = ex0;
      B[v + i]     = by0;
      B[2 * v + i] = bz0;
    }
  }
}

// Indices of the particles that crossed x = Lx (key == kKeyReplace), in particle
// order, for every segment of the upper chunks: one block per segment, one wave
// ballot and one LDS prefix per 1024-particle stretch.
__global__ void __launch_bounds__(1024) k_wall_list(Geom g, const int32_t* __restrict__ useg,
                                                    const int32_t* __restrict__ np,
                                                    const int32_t* __restrict__ key,
                                                    int32_t* __restrict__ list,
                                                    int32_t* __restrict__ count)
{
  __shared__ int wsum[16];
  __shared__ int base;
  const int k    = blockIdx.x;
  const int seg  = useg[k];
  const int n    = np[seg];
  const int lane = threadIdx.x & 63;
  const int wv   = threadIdx.x >> 6;
  if (threadIdx.x == 0)
    base = 0;
  __syncthreads();
  for (int s0 = 0; s0 < n; s0 += 1024) {
    const int  ip   = s0 + (int)threadIdx.x;
    const bool hit  = ip < n && key[(long)seg * g.cap + ip] == kKeyReplace;
    const unsigned long long m = __ballot(hit);
    const int  pre  = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0)
      wsum[wv] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wv; w++)
      off += wsum[w];
    if (hit)
      list[(long)k * g.cap + off + pre] = ip;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < 16; w++)
        tot += wsum[w];
      base += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    count[k] = base;
}

// Bin and slot of a particle that is back in its chunk, or its exchange direction
// (counted in its 256-particle tile) when y or z took it into a neighbour
// (XtensorParticle::count then the Particle exchange).
__device__ __forceinline__ void place(const Geom& g, const double* __restrict__ lim,
                                      const double* __restrict__ rng, int seg, long ip, double x,
                                      double y, double z, int32_t* __restrict__ cellcnt,
                                      int32_t* __restrict__ key, int32_t* __restrict__ slot,
                                      int32_t* __restrict__ tilecnt, int ntile)
{
  const int  c = seg / g.Ns;
  const long i = (long)seg * g.cap + ip;
  const int  b = bin_of(g, lim, rng, c, x, y, z);
  if (b >= 0) {
    key[i]  = b;
    slot[i] = atomicAdd(&cellcnt[(long)seg * g.nbin + b], 1);
  } else {
    const int d = dir_of(rng, c, x, y, z);
    key[i]      = -1 - d;
    atomicAdd(&tilecnt[((long)seg * ntile + ip / 256) * 27 + d], 1);
  }
}

// New values for the particles listed by k_wall_list (x, ux, uy, uz from the
// host's MainChunk::set_boundary_particle draws; y and z stay), then their count.
__global__ void __launch_bounds__(256) k_wall_apply(Geom g, PArr p, const int32_t* __restrict__ useg,
                                                    const int32_t* __restrict__ list,
                                                    const int32_t* __restrict__ count,
                                                    const int64_t* __restrict__ voff,
                                                    const double* __restrict__ vals,
                                                    const double* __restrict__ lim,
                                                    const double* __restrict__ rng,
                                                    int32_t* __restrict__ cellcnt,
                                                    int32_t* __restrict__ key,
                                                    int32_t* __restrict__ slot,
                                                    int32_t* __restrict__ tilecnt, int ntile)
{
  const int k   = blockIdx.x;
  const int seg = useg[k];
  const int n   = count[k];
  for (int t = (int)threadIdx.x; t < n; t += blockDim.x) {
    const long    ip = list[(long)k * g.cap + t];
    const long    i  = (long)seg * g.cap + ip;
    const double* v  = vals + voff[k] + 4 * (long)t;
    p.x[i]           = v[0];
    p.ux[i]          = v[1];
    p.uy[i]          = v[2];
    p.uz[i]          = v[3];
    place(g, lim, rng, seg, ip, v[0], p.y[i], p.z[i], cellcnt, key, slot, tilecnt, ntile);
  }
}

// Append the injected particles of every upper segment (MainChunk::inject_particle
// + count(np_prev, np_next - 1, false), example/shock/main.cpp:445-535) after its
// particles; the old position is set to the new one.
__global__ void __launch_bounds__(256) k_inject_append(Geom g, PArr p, double* __restrict__ xo,
                                                       double* __restrict__ yo,
                                                       double* __restrict__ zo,
                                                       const int32_t* __restrict__ useg,
                                                       const int32_t* __restrict__ ninj,
                                                       const int64_t* __restrict__ voff,
                                                       const double* __restrict__ vals,
                                                       const double* __restrict__ lim,
                                                       const double* __restrict__ rng,
                                                       int32_t* __restrict__ np,
                                                       int32_t* __restrict__ cellcnt,
                                                       int32_t* __restrict__ key,
                                                       int32_t* __restrict__ slot,
                                                       int32_t* __restrict__ tilecnt, int ntile,
                                                       int32_t* __restrict__ errflag)
{
  __shared__ int n0s;
  const int k   = blockIdx.x;
  const int seg = useg[k];
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
    p.x[i]          = v[0];
    p.y[i]          = v[1];
    p.z[i]          = v[2];
    p.ux[i]         = v[3];
    p.uy[i]         = v[4];
    p.uz[i]         = v[5];
    p.id[i]         = __double_as_longlong(v[6]);
    xo[i]           = v[0];
    yo[i]           = v[1];
    zo[i]           = v[2];
    place(g, lim, rng, seg, n0 + t, v[0], v[1], v[2], cellcnt, key, slot, tilecnt, ntile);
  }
  __syncthreads();
  if (threadIdx.x == 0)
    np[seg] = n0 + n;
}

// Kinetic energy per species as the reference forms it, sum over interior cells of
// um[4] c - um[0] c^2 (pic/pic_chunk.cpp:433-438), from the folded moments
// [Ns][14][vol].  Block partials per species, summed on the host.
__global__ void __launch_bounds__(256) k_moment_energy(Geom g, const double* __restrict__ M,
                                                       double* __restrict__ partial)
{
  __shared__ double red[256];
  const long n   = (long)g.N[0] * g.N[1] * g.N[2];
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int  sp  = blockIdx.y;
  double     acc = 0.0;
  if (tid < n) {
    int  x = (int)(tid % g.N[2]) + g.pad[2];
    long r = tid / g.N[2];
    int  y = (int)(r % g.N[1]) + g.pad[1];
    int  z = (int)(r / g.N[1]) + g.pad[0];
    long i = gidx(g, z, y, x);
    const double* m = M + (long)sp * 14 * g.vol;
    acc = m[4 * g.vol + i] * g.cc - m[i] * g.cc * g.cc;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partial[(long)sp * gridDim.x + blockIdx.x] = red[0];
}

} // namespace

void launch_wall_emf(Sim& s)
{
  const Geom& g = s.g;
  if (!g.xwall)
    return;
  const long   n  = (long)g.P[0] * g.P[1] * 4 * g.Nb;
  const double* e = s.shk.E0;
  const double* b = s.shk.B0;
  if (s.fp32)
    hipLaunchKernelGGL(k_wall_emf<float>, dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, s.f_E,
                       s.f_B, (float)e[0], (float)e[1], (float)e[2], (float)b[0], (float)b[1],
                       (float)b[2]);
  else
    hipLaunchKernelGGL(k_wall_emf<double>, dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, s.d_E,
                       s.d_B, e[0], e[1], e[2], b[0], b[1], b[2]);
}

void launch_wall_list(Sim& s)
{
  const int nk = (int)s.shk.useg.size();
  if (nk == 0)
    return;
  hipLaunchKernelGGL(k_wall_list, dim3(nk), dim3(1024), 0, s.stream, s.g, s.shk.d_useg, s.d_np,
                     s.ss.key, s.shk.d_list, s.shk.d_count);
}

void launch_wall_apply(Sim& s)
{
  const int nk = (int)s.shk.useg.size();
  if (nk == 0)
    return;
  hipLaunchKernelGGL(k_wall_apply, dim3(nk), dim3(256), 0, s.stream, s.g, s.pa, s.shk.d_useg,
                     s.shk.d_list, s.shk.d_count, s.shk.d_voff, s.shk.d_vals, s.d_lim, s.d_rng,
                     s.ss.cellcnt, s.ss.key, s.ss.slot, s.ss.tilecnt, s.ss.ntile);
}

void launch_inject_append(Sim& s)
{
  const int nk = (int)s.shk.useg.size();
  if (nk == 0)
    return;
  hipLaunchKernelGGL(k_inject_append, dim3(nk), dim3(256), 0, s.stream, s.g, s.pa, s.d_xo[0],
                     s.d_xo[1], s.d_xo[2], s.shk.d_useg, s.shk.d_count, s.shk.d_voff,
                     s.shk.d_vals, s.d_lim, s.d_rng, s.d_np, s.ss.cellcnt, s.ss.key, s.ss.slot,
                     s.ss.tilecnt, s.ss.ntile, s.ss.errflag);
}

void launch_moment_exchange(Sim& s, bool deposit)
{
  if (!s.d_M)
    HIPCHK(hipMalloc((void**)&s.d_M, sizeof(double) * 14 * (size_t)s.g.Ns * s.g.vol));
  if (deposit)
    launch_deposit_moment(s);
  // the Mom exchange; the shock's set_boundary_mom writes only the z = 0 ghost
  // plane of its per-chunk array (four indices into a 5-D xtensor, aligned to the
  // trailing axes), so no interior moment changes and nothing is done for it here
  launch_fold(s, s.d_M, 0, 14 * s.g.Ns, PN_PHASE_RHO);
}

void launch_moment_energy(Sim& s, double* out)
{
  const Geom& g  = s.g;
  const long  n  = (long)g.N[0] * g.N[1] * g.N[2];
  const int   nb = nblk(n, 256);
  double*     d  = nullptr;
  (void)hipMallocAsync((void**)&d, sizeof(double) * nb * g.Ns, s.stream);
  hipLaunchKernelGGL(k_moment_energy, dim3(nb, g.Ns), dim3(256), 0, s.stream, g, s.d_M, d);
  std::vector<double> h((size_t)nb * g.Ns);
  (void)hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s.stream);
  (void)hipFreeAsync(d, s.stream);
  (void)hipStreamSynchronize(s.stream);
  for (int sp = 0; sp < g.Ns; sp++) {
    double acc = 0;
    for (int b = 0; b < nb; b++)
      acc += h[(size_t)sp * nb + b];
    out[sp] = acc;
  }
}

} // namespace picnix
