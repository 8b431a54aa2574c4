// pn_diag.hip -- energy / divergence reductions (FP64) and chunk views of the
// global fields for the C-ABI boundary.
#include <mutex>
#include <vector>

#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace picnix {

namespace {

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

__device__ __forceinline__ long gidx(const Geom& g, int z, int y, int x)
{
  return ((long)z * g.P[1] + y) * g.P[2] + x;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out)
{
  __shared__ double red[NV][256];
  for (int k = 0; k < NV; k++)
    red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < NV; k++)
        red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++)
      out[k] = red[k][0];
}

// Per interior cell: E^2/2, B^2/2 (PicChunk::get_energy, pic/pic_chunk.cpp:407-431),
// div(E) - rho and div(B) (BaseMaxwell::get_diverror_{2,3}d,
// pic/engine/maxwell.hpp:203-229,350-381) of interior cell `tid` (x fastest).
template <int Dim>
__device__ __forceinline__ void field_terms(const Geom& g, const double* __restrict__ E,
                                            const double* __restrict__ B,
                                            const double* __restrict__ J, long tid, double acc[4])
{
  int  x = (int)(tid % g.N[2]) + g.pad[2];
  long r = tid / g.N[2];
  int  y = (int)(r % g.N[1]) + g.pad[1];
  int  z = (int)(r / g.N[1]) + g.pad[0];
  long i = gidx(g, z, y, x);
  long v = g.vol, sy = g.P[2], sz = g.plane;
  const double rdx = 1 / g.del[2], rdy = 1 / g.del[1], rdz = 1 / g.del[0];
  double ex = E[i], ey = E[v + i], ez = E[2 * v + i];
  double bx = B[i], by = B[v + i], bz = B[2 * v + i];
  acc[0] = 0.5 * (ex * ex + ey * ey + ez * ez);
  acc[1] = 0.5 * (bx * bx + by * by + bz * bz);
  if constexpr (Dim == 1) {
    // get_diverror_1d (pic/engine/maxwell.hpp:63-83)
    acc[2] = (E[i + 1] - ex) * rdx - J[i];
    acc[3] = (bx - B[i - 1]) * rdx;
    (void)rdy;
    (void)rdz;
    (void)sz;
  } else if constexpr (Dim == 3) {
    acc[2] = (E[i + 1] - ex) * rdx + (E[v + i + sy] - ey) * rdy + (E[2 * v + i + sz] - ez) * rdz -
             J[i];
    acc[3] = (bx - B[i - 1]) * rdx + (by - B[v + i - sy]) * rdy + (bz - B[2 * v + i - sz]) * rdz;
  } else {
    acc[2] = (E[i + 1] - ex) * rdx + (E[v + i + sy] - ey) * rdy - J[i];
    acc[3] = (bx - B[i - 1]) * rdx + (by - B[v + i - sy]) * rdy;
  }
  // cells next to a physical boundary are left out of the divergence errors
  // (pic/pic_engine.hpp:42-71)
  if (g.xwall && (x < g.pad[2] + g.Nb || x >= g.pad[2] + g.N[2] - g.Nb)) {
    acc[2] = 0;
    acc[3] = 0;
  }
}

// One thread per interior cell; block partials, summed on the host.
template <int Dim>
__global__ void __launch_bounds__(256) k_field_energy(Geom g, const double* __restrict__ E,
                                                      const double* __restrict__ B,
                                                      const double* __restrict__ J,
                                                      double* __restrict__ partial)
{
  long   n   = (long)g.N[0] * g.N[1] * g.N[2];
  long   tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[4] = {0, 0, 0, 0};
  if (tid < n)
    field_terms<Dim>(g, E, B, J, tid, acc);
  block_sum<4>(acc, partial + 4 * blockIdx.x);
}

// Kinetic energy m c^2 (gamma - 1) of particle i of species sp; the reference
// forms the same sum through the moment deposit, sum(um[...,4] c - um[...,0] c^2)
// (pic/pic_chunk.cpp:433-438).
__device__ __forceinline__ double kinetic(const Geom& g, const PArr& p, long i, int sp)
{
  const double rc = 1 / g.cc;
  double       ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
  double       u2 = (ux * ux + uy * uy + uz * uz) * rc * rc;
  double       gm = sqrt(1 + u2);
  return g.m[sp] * g.cc * g.cc * (u2 / (gm + 1));
}

__global__ void __launch_bounds__(256) k_particle_energy(Geom g, PArr p,
                                                         const int32_t* __restrict__ np,
                                                         double* __restrict__ partial, int ntile)
{
  const int  seg = blockIdx.y;
  const long ip  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double     acc[1] = {0};
  if (ip < np[seg])
    acc[0] = kinetic(g, p, (long)seg * g.cap + ip, seg % g.Ns);
  block_sum<1>(acc, partial + ((long)seg * ntile + blockIdx.x));
}

// ---- the same reductions into a fixed number of partial rows ----
// A grid of at most kEnergyRows blocks walks the cells (or the particle tiles)
// by stride; every block writes its row of partial sums, and one block sums the
// rows in row order.  All of it lives in the file-scope scratch array below, so
// these kernels need no allocation and no clearing (every row is written before
// it is read).
constexpr int kEnergyRows = 2048;
constexpr int kEnergyCols = 4 + kMaxSpecies;

template <int Dim>
__global__ void __launch_bounds__(256) k_field_energy_rows(Geom g, const double* __restrict__ E,
                                                           const double* __restrict__ B,
                                                           const double* __restrict__ J,
                                                           double* __restrict__ rows)
{
  const long n      = (long)g.N[0] * g.N[1] * g.N[2];
  const long stride = (long)gridDim.x * blockDim.x;
  double     acc[4] = {0, 0, 0, 0};
  for (long tid = (long)blockIdx.x * blockDim.x + threadIdx.x; tid < n; tid += stride) {
    double t[4];
    field_terms<Dim>(g, E, B, J, tid, t);
    for (int k = 0; k < 4; k++)
      acc[k] += t[k];
  }
  block_sum<4>(acc, rows + 4 * blockIdx.x);
}

__global__ void __launch_bounds__(256) k_particle_energy_rows(Geom g, PArr p,
                                                             const int32_t* __restrict__ np,
                                                             double* __restrict__ rows, int ntile)
{
  const long nitem = (long)g.nseg * ntile;
  double     acc[kMaxSpecies];
#pragma unroll
  for (int k = 0; k < kMaxSpecies; k++)
    acc[k] = 0;
  for (long it = blockIdx.x; it < nitem; it += gridDim.x) {
    const int  seg = (int)(it / ntile);
    const long ip  = (it - (long)seg * ntile) * blockDim.x + threadIdx.x;
    if (ip < np[seg]) {
      const int    sp = seg % g.Ns;
      const double e  = kinetic(g, p, (long)seg * g.cap + ip, sp);
#pragma unroll
      for (int k = 0; k < kMaxSpecies; k++)
        acc[k] += k == sp ? e : 0.0;
    }
  }
  block_sum<kMaxSpecies>(acc, rows + kMaxSpecies * blockIdx.x);
}

// out[0..3] = column sums of the field rows, out[4 + s] = of the particle rows
__global__ void __launch_bounds__(64) k_energy_rows_sum(const double* __restrict__ frows, int nf,
                                                        const double* __restrict__ prows, int npr,
                                                        double* __restrict__ out)
{
  const int k = threadIdx.x;
  if (k >= kEnergyCols)
    return;
  double s = 0;
  if (k < 4)
    for (int r = 0; r < nf; r++)
      s += frows[4 * r + k];
  else
    for (int r = 0; r < npr; r++)
      s += prows[kMaxSpecies * r + (k - 4)];
  out[k] = s;
}

// Field rows, particle rows, then the kEnergyCols sums.  Each GPU holds its own
// copy until the process ends; every call on one GPU uses it, so calls that
// overlap on one GPU race.
static __device__ __align__(256) double
    scratch[kEnergyRows * 4 + kEnergyRows * kMaxSpecies + kEnergyCols];

__device__ __forceinline__ int wrapg(int l, int off, int pad, int n)
{
  // chunk-local ghost-inclusive index -> global padded index of the periodic image
  int j = off + l - pad;
  j     = j < 0 ? j + n : (j >= n ? j - n : j);
  return j + pad;
}

// Ghost-inclusive per-chunk view (PicChunk::uf / uj layout): writes components
// [coff, coff + ncomp) of records of `stride` doubles.
__global__ void __launch_bounds__(256) k_chunk_get(Geom g, const double* __restrict__ F,
                                                   int ncomp, int coff, int stride, int oz,
                                                   int oy, int ox, double* __restrict__ out)
{
  const int nz = g.dims[0] + 2 * g.Nb, ny = g.dims[1] + 2 * g.Nb, nx = g.dims[2] + 2 * g.Nb;
  long      t  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nz * ny * nx)
    return;
  int  lx = (int)(t % nx);
  long r  = t / nx;
  int  ly = (int)(r % ny);
  int  lz = (int)(r / ny);
  // an absent dimension (z in 2D, y and z in 1D) has one global plane: the
  // chunk's ghost planes along it read as zero
  bool ok = (g.has[0] || lz == g.Nb) && (g.has[1] || ly == g.Nb);
  int  gz = g.has[0] ? wrapg(lz, oz, g.pad[0], g.N[0]) : 0;
  int  gy = g.has[1] ? wrapg(ly, oy, g.pad[1], g.N[1]) : 0;
  int  gx = g.xwall ? ox + lx : wrapg(lx, ox, g.pad[2], g.N[2]);
  long i  = gidx(g, gz, gy, gx);
  for (int k = 0; k < ncomp; k++)
    out[t * stride + coff + k] = ok ? F[k * g.vol + i] : 0.0;
}

// Interior cells of a per-chunk array -> global array (ghost cells are ignored).
__global__ void __launch_bounds__(256) k_chunk_set(Geom g, double* __restrict__ F, int ncomp,
                                                   int coff, int stride, int oz, int oy, int ox,
                                                   const double* __restrict__ in)
{
  const int nz = g.dims[0] + 2 * g.Nb, ny = g.dims[1] + 2 * g.Nb, nx = g.dims[2] + 2 * g.Nb;
  long      t  = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nz * ny * nx)
    return;
  int  lx = (int)(t % nx);
  long r  = t / nx;
  int  ly = (int)(r % ny);
  int  lz = (int)(r / ny);
  bool in_x = lx >= g.Nb && lx < g.Nb + g.dims[2];
  bool in_y = g.has[1] ? (ly >= g.Nb && ly < g.Nb + g.dims[1]) : ly == g.Nb;
  bool in_z = g.has[0] ? (lz >= g.Nb && lz < g.Nb + g.dims[0]) : lz == g.Nb;
  if (!(in_x && in_y && in_z))
    return;
  int  gz = g.has[0] ? oz + lz : 0;
  int  gy = g.has[1] ? oy + ly : 0;
  long i  = gidx(g, gz, gy, ox + lx);
  for (int k = 0; k < ncomp; k++)
    F[k * g.vol + i] = in[t * stride + coff + k];
}

} // namespace

void launch_energy(Sim& s, double* out)
{
  const Geom& g   = s.g;
  long        n   = (long)g.N[0] * g.N[1] * g.N[2];
  int         nb1 = nblk(n, 256);
  int         nt  = s.ss.ntile;
  size_t      np1 = 4 * (size_t)nb1, np2 = (size_t)g.nseg * nt;
  double*     d_part = nullptr;
  if (s.fp32) {
    // FP64 reductions over the FP64 mirrors of the FP32 state
    launch32_fields_to_f64(s, PN_FIELD_UF);
    launch32_fields_to_f64(s, PN_FIELD_UJ);
    launch32_to_f64(s, 0, g.nseg, false);
  }
  hipMallocAsync((void**)&d_part, sizeof(double) * (np1 + np2), s.stream);
  hipMemsetAsync(d_part, 0, sizeof(double) * (np1 + np2), s.stream);
  if (g.dim == 3)
    hipLaunchKernelGGL(k_field_energy<3>, dim3(nb1), dim3(256), 0, s.stream, g, s.d_E, s.d_B,
                       s.d_J, d_part);
  else if (g.dim == 1)
    hipLaunchKernelGGL(k_field_energy<1>, dim3(nb1), dim3(256), 0, s.stream, g, s.d_E, s.d_B,
                       s.d_J, d_part);
  else
    hipLaunchKernelGGL(k_field_energy<2>, dim3(nb1), dim3(256), 0, s.stream, g, s.d_E, s.d_B,
                       s.d_J, d_part);
  if (!g.xwall)
    hipLaunchKernelGGL(k_particle_energy, dim3(nt, g.nseg), dim3(256), 0, s.stream, g, s.pa,
                       s.d_np, d_part + np1, nt);
  std::vector<double> h(np1 + np2);
  hipMemcpyAsync(h.data(), d_part, sizeof(double) * (np1 + np2), hipMemcpyDeviceToHost, s.stream);
  hipFreeAsync(d_part, s.stream);
  hipStreamSynchronize(s.stream);
  for (int k = 0; k < 4 + g.Ns; k++)
    out[k] = 0.0;
  for (int b = 0; b < nb1; b++)
    for (int k = 0; k < 4; k++)
      out[k] += h[4 * (size_t)b + k];
  for (int seg = 0; seg < g.nseg; seg++)
    for (int b = 0; b < nt; b++)
      out[4 + seg % g.Ns] += h[np1 + (size_t)seg * nt + b];
  if (g.xwall) {
    // with physical boundaries the moment sum differs from the particle sum (the
    // weight beyond a wall is not folded back): form it as the reference does
    launch_moment_exchange(s, true);
    launch_moment_energy(s, out + 4);
  }
}

namespace {

// Device address of `scratch` on every GPU the library has run on (each GPU has
// its own copy of the array at its own address).  On the host the array's name
// is only a stand-in: kernels and copies take these addresses.
constexpr int kMaxDevices = 64;
std::mutex    scratch_mu;
double*       scratch_ptr[kMaxDevices] = {};

} // namespace

void energy_static_init(Sim& s)
{
  if (s.device < 0 || s.device >= kMaxDevices)
    throw HipError{"device index out of the range of the energy scratch table"};
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[s.device]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    scratch_ptr[s.device] = static_cast<double*>(ptr);
  }
  s.energy_scratch = scratch_ptr[s.device];
}

// launch_energy without any allocation: the partial rows and the sums live in
// the file-scope scratch array of this GPU.
int launch_energy_static(Sim& s, double* out)
{
  const Geom& g = s.g;
  if (!s.energy_scratch)
    return PN_ERR_STATE;
  if (g.xwall || g.Ns > kMaxSpecies)
    return PN_ERR_UNSUPPORTED;
  double*    frows = s.energy_scratch;
  double*    prows = frows + 4 * kEnergyRows;
  double*    sums  = prows + kMaxSpecies * kEnergyRows;
  const long n     = (long)g.N[0] * g.N[1] * g.N[2];
  const int  nf    = (int)std::min<long>(kEnergyRows, nblk(n, 256));
  const int  npr   = (int)std::min<long>(kEnergyRows, (long)g.nseg * s.ss.ntile);
  if (s.fp32) {
    // FP64 reductions over the FP64 mirrors of the FP32 state
    launch32_fields_to_f64(s, PN_FIELD_UF);
    launch32_fields_to_f64(s, PN_FIELD_UJ);
    launch32_to_f64(s, 0, g.nseg, false);
  }
  if (g.dim == 3)
    hipLaunchKernelGGL(k_field_energy_rows<3>, dim3(nf), dim3(256), 0, s.stream, g, s.d_E, s.d_B,
                       s.d_J, frows);
  else if (g.dim == 1)
    hipLaunchKernelGGL(k_field_energy_rows<1>, dim3(nf), dim3(256), 0, s.stream, g, s.d_E, s.d_B,
                       s.d_J, frows);
  else
    hipLaunchKernelGGL(k_field_energy_rows<2>, dim3(nf), dim3(256), 0, s.stream, g, s.d_E, s.d_B,
                       s.d_J, frows);
  hipLaunchKernelGGL(k_particle_energy_rows, dim3(npr), dim3(256), 0, s.stream, g, s.pa, s.d_np,
                     prows, s.ss.ntile);
  hipLaunchKernelGGL(k_energy_rows_sum, dim3(1), dim3(64), 0, s.stream, frows, nf, prows, npr,
                     sums);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, sums, sizeof(double) * (4 + g.Ns), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  return PN_OK;
}

void launch_chunk_fields_get(Sim& s, int chunk, int which, double* d_out)
{
  const Geom& g  = s.g;
  long        nn = (long)(g.dims[0] + 2 * g.Nb) * (g.dims[1] + 2 * g.Nb) * (g.dims[2] + 2 * g.Nb);
  const int*  o  = &s.offset[chunk * 3];
  dim3        grid(nblk(nn, 256));
  if (which == 0) {
    hipLaunchKernelGGL(k_chunk_get, grid, dim3(256), 0, s.stream, g, s.d_E, 3, 0, 6, o[0], o[1],
                       o[2], d_out);
    hipLaunchKernelGGL(k_chunk_get, grid, dim3(256), 0, s.stream, g, s.d_B, 3, 3, 6, o[0], o[1],
                       o[2], d_out);
  } else if (which == PN_FIELD_UJ) {
    hipLaunchKernelGGL(k_chunk_get, grid, dim3(256), 0, s.stream, g, s.d_J, 4, 0, 4, o[0], o[1],
                       o[2], d_out);
  } else {
    // [z][y][x][Ns][14], PicChunk::um
    hipLaunchKernelGGL(k_chunk_get, grid, dim3(256), 0, s.stream, g, s.d_M, 14 * g.Ns, 0,
                       14 * g.Ns, o[0], o[1], o[2], d_out);
  }
}

void launch_chunk_fields_set(Sim& s, int chunk, int which, const double* d_in)
{
  const Geom& g  = s.g;
  long        nn = (long)(g.dims[0] + 2 * g.Nb) * (g.dims[1] + 2 * g.Nb) * (g.dims[2] + 2 * g.Nb);
  const int*  o  = &s.offset[chunk * 3];
  dim3        grid(nblk(nn, 256));
  if (which == 0) {
    hipLaunchKernelGGL(k_chunk_set, grid, dim3(256), 0, s.stream, g, s.d_E, 3, 0, 6, o[0], o[1],
                       o[2], d_in);
    hipLaunchKernelGGL(k_chunk_set, grid, dim3(256), 0, s.stream, g, s.d_B, 3, 3, 6, o[0], o[1],
                       o[2], d_in);
  } else {
    hipLaunchKernelGGL(k_chunk_set, grid, dim3(256), 0, s.stream, g, s.d_J, 4, 0, 4, o[0], o[1],
                       o[2], d_in);
  }
}

} // namespace picnix
