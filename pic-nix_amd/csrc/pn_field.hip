// pn_field.hip -- Yee FDTD updates, periodic ghost refresh and current fold on
// the global padded grid.
//
// The reference updates every chunk's ghost-inclusive block and then copies
// neighbour edges into ghosts (pic/engine/maxwell.hpp:383-496, EMF halo
// nix/xtensor_halo3d.hpp:18-71).  Here the whole domain is one periodic array
// with an Nb-wide padding that mirrors the opposite edge, so every interior
// cell sees the same neighbour values as in the reference and gets the same
// floating-point result; the chunk edges need no copies at all.
#include <algorithm>

#include "pn_sim.hpp"

namespace picnix {

namespace {

__device__ __forceinline__ long gidx(const Geom& g, int z, int y, int x)
{
  return ((long)z * g.P[1] + y) * g.P[2] + x;
}

// B -= c dt curl F: pic/engine/maxwell.hpp:442-496 (3D) and :292-348 (2D), where
// F = ff(...,0,k) is the Friedman-filtered E (k_friedman_combine), or E itself
// for theta = 0 (then ff(...,0,k) = 1 E + (-0) ff1 + 0 ff2 == E exactly).  One
// thread per interior cell.
template <int Dim, typename T>
__global__ void __launch_bounds__(256) k_push_bfd(Geom g, const T* __restrict__ E,
                                                  T* __restrict__ B, T cflx, T cfly, T cflz)
{
  long n   = (long)g.N[0] * g.N[1] * g.N[2];
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n)
    return;
  int  x  = (int)(tid % g.N[2]) + g.pad[2];
  long r  = tid / g.N[2];
  int  y  = (int)(r % g.N[1]) + g.pad[1];
  int  z  = (int)(r / g.N[1]) + g.pad[0];
  long i  = gidx(g, z, y, x);
  long sx = 1, sy = g.P[2], sz = g.plane, v = g.vol;
  const T* Ex = E;
  const T* Ey = E + v;
  const T* Ez = E + 2 * v;
  if constexpr (Dim == 3) {
    B[i] += (-cfly) * (Ez[i] - Ez[i - sy]) + (+cflz) * (Ey[i] - Ey[i - sz]);
    B[v + i] += (-cflz) * (Ex[i] - Ex[i - sz]) + (+cflx) * (Ez[i] - Ez[i - sx]);
    B[2 * v + i] += (-cflx) * (Ey[i] - Ey[i - sx]) + (+cfly) * (Ex[i] - Ex[i - sy]);
  } else if constexpr (Dim == 1) {
    // push_bfd_1d (pic/engine/maxwell.hpp:148-201): Bx unchanged
    B[v + i] += (+cflx) * (Ez[i] - Ez[i - sx]);
    B[2 * v + i] += (-cflx) * (Ey[i] - Ey[i - sx]);
    (void)Ex;
    (void)sy;
    (void)sz;
  } else {
    B[i] += (-cfly) * (Ez[i] - Ez[i - sy]);
    B[v + i] += (+cflx) * (Ez[i] - Ez[i - sx]);
    B[2 * v + i] += (-cflx) * (Ey[i] - Ey[i - sx]) + (+cfly) * (Ex[i] - Ex[i - sy]);
  }
}

// Friedman filter (pic/engine/maxwell.hpp:442-457 / :292-307):
//   ff0 = A E + B ff1 + C ff2,  A = 1 + theta/2, B = -theta (1 - theta/2),
//   C = theta (1 - theta)^2 / 2,
// over every padded cell (the reference covers the ghosts too; the padding holds
// periodic images, so each padding cell gets its image's value).  ff is
// [level][component][vol].
template <typename T>
__global__ void __launch_bounds__(256) k_friedman_combine(Geom g, const T* __restrict__ E,
                                                          T* __restrict__ ff, T A, T B, T C)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.vol)
    return;
  const long v = g.vol;
#pragma unroll
  for (int k = 0; k < 3; k++)
    ff[k * v + i] = A * E[k * v + i] + B * ff[(3 + k) * v + i] + C * ff[(6 + k) * v + i];
}

// History shift before the E update (maxwell.hpp:389-404 / :237-252):
//   ff2 = ff1 + theta ff2;  ff1 = E.
template <typename T>
__global__ void __launch_bounds__(256) k_friedman_shift(Geom g, const T* __restrict__ E,
                                                        T* __restrict__ ff, T theta)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.vol)
    return;
  const long v = g.vol;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ff[(6 + k) * v + i] = ff[(3 + k) * v + i] + theta * ff[(6 + k) * v + i];
    ff[(3 + k) * v + i] = E[k * v + i];
  }
}

// init_friedman (maxwell.hpp:44-60): every level = E.
template <typename T>
__global__ void __launch_bounds__(256) k_friedman_init(Geom g, const T* __restrict__ E,
                                                       T* __restrict__ ff)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.vol)
    return;
  const long v = g.vol;
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int k = 0; k < 3; k++)
      ff[(3 * l + k) * v + i] = E[k * v + i];
}

// E += c dt curl B - dt J: pic/engine/maxwell.hpp:383-440 (3D) and :231-290 (2D).
template <int Dim, typename T>
__global__ void __launch_bounds__(256) k_push_efd(Geom g, T* __restrict__ E,
                                                  const T* __restrict__ B,
                                                  const T* __restrict__ J, T cflx, T cfly, T cflz,
                                                  T delt)
{
  long n   = (long)g.N[0] * g.N[1] * g.N[2];
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n)
    return;
  int  x  = (int)(tid % g.N[2]) + g.pad[2];
  long r  = tid / g.N[2];
  int  y  = (int)(r % g.N[1]) + g.pad[1];
  int  z  = (int)(r / g.N[1]) + g.pad[0];
  long i  = gidx(g, z, y, x);
  long sx = 1, sy = g.P[2], sz = g.plane, v = g.vol;
  const T* Bx = B;
  const T* By = B + v;
  const T* Bz = B + 2 * v;
  if constexpr (Dim == 1) {
    // push_efd_1d (pic/engine/maxwell.hpp:85-146)
    E[i] += -delt * J[v + i];
    E[v + i] += (-cflx) * (Bz[i + sx] - Bz[i]) - delt * J[2 * v + i];
    E[2 * v + i] += (+cflx) * (By[i + sx] - By[i]) - delt * J[3 * v + i];
    (void)Bx;
    (void)sy;
    (void)sz;
  } else if constexpr (Dim == 3) {
    E[i] += (+cfly) * (Bz[i + sy] - Bz[i]) + (-cflz) * (By[i + sz] - By[i]) - delt * J[v + i];
    E[v + i] +=
        (+cflz) * (Bx[i + sz] - Bx[i]) + (-cflx) * (Bz[i + sx] - Bz[i]) - delt * J[2 * v + i];
    E[2 * v + i] +=
        (+cflx) * (By[i + sx] - By[i]) + (-cfly) * (Bx[i + sy] - Bx[i]) - delt * J[3 * v + i];
  } else {
    E[i] += (+cfly) * (Bz[i + sy] - Bz[i]) - delt * J[v + i];
    E[v + i] += (-cflx) * (Bz[i + sx] - Bz[i]) - delt * J[2 * v + i];
    E[2 * v + i] +=
        (+cflx) * (By[i + sx] - By[i]) + (-cfly) * (Bx[i + sy] - Bx[i]) - delt * J[3 * v + i];
  }
}

__device__ __forceinline__ int wrap(int i, int pad, int n)
{
  int j = i - pad;
  j     = j < 0 ? j + n : (j >= n ? j - n : j);
  return j + pad;
}

// Copy the periodic image into every padding cell (the EMF exchange,
// nix/xtensor_halo3d.hpp:18-71, on a single periodic array).
template <typename T>
__global__ void __launch_bounds__(256) k_fill_ghost(Geom g, T* __restrict__ f, int ncomp)
{
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.vol)
    return;
  int  x = (int)(tid % g.P[2]);
  long r = tid / g.P[2];
  int  y = (int)(r % g.P[1]);
  int  z = (int)(r / g.P[1]);
  bool in = x >= g.pad[2] && x < g.pad[2] + g.N[2] && y >= g.pad[1] && y < g.pad[1] + g.N[1] &&
            z >= g.pad[0] && z < g.pad[0] + g.N[0];
  if (in)
    return;
  // a non-periodic x keeps its padding (the wall values, k_wall_emf); the y and
  // z padding of those columns still takes the periodic image
  const int xs = g.xwall ? x : wrap(x, g.pad[2], g.N[2]);
  long src = gidx(g, wrap(z, g.pad[0], g.N[0]), wrap(y, g.pad[1], g.N[1]), xs);
  if (src == tid)
    return;
  for (int k = 0; k < ncomp; k++)
    f[k * g.vol + tid] = f[k * g.vol + src];
}

// Fold the current deposited into the padding back onto the periodic image
// (the Cur exchange adds a neighbour's ghost into the interior edge,
// nix/xtensor_halo3d.hpp:76-129).  One pass per dimension; each thread owns one
// line, so the additions happen in a fixed order (x, then y, then z).
template <typename T>
__global__ void __launch_bounds__(256) k_fold_line(Geom g, T* __restrict__ J, int dir, int k0,
                                                   int k1)
{
  // dir: 2 = x lines over (z, y); 1 = y lines over (z, x); 0 = z lines over (y, x)
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int  n0, n1;
  if (dir == 2) {
    n0 = g.P[0];
    n1 = g.P[1];
  } else if (dir == 1) {
    n0 = g.P[0];
    n1 = g.P[2];
  } else {
    n0 = g.P[1];
    n1 = g.P[2];
  }
  if (tid >= (long)n0 * n1)
    return;
  int  i0 = (int)(tid / n1), i1 = (int)(tid % n1);
  int  pad = g.pad[dir], n = g.N[dir];
  long stride = dir == 2 ? 1 : (dir == 1 ? g.P[2] : g.plane);
  long base   = dir == 2 ? gidx(g, i0, i1, 0) : (dir == 1 ? gidx(g, i0, 0, i1) : gidx(g, 0, i0, i1));
  for (int k = k0; k < k1; k++) {
    T* f = J + k * g.vol + base;
    for (int p = 0; p < pad; p++) {
      // low padding p -> interior n + p ; high padding pad + n + p -> interior pad + p
      f[(long)(n + p) * stride] += f[(long)p * stride];
      f[(long)(pad + p) * stride] += f[(long)(pad + n + p) * stride];
    }
  }
}

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Plain fills and copies as kernels: a step that may be captured into a
// hipGraph records no memset or memcpy nodes.
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ p, T v, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void __launch_bounds__(256) k_copy_i32(int32_t* __restrict__ d, const int32_t* __restrict__ s,
                                                  long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    d[i] = s[i];
}

inline int fill_grid(long n) { return (int)std::min<long>(std::max<long>(nblk(n, 256), 1), 4096); }

} // namespace

void launch_zero(Sim& s, double* p, long n)
{
  hipLaunchKernelGGL(k_fill<double>, dim3(fill_grid(n)), dim3(256), 0, s.stream, p, 0.0, n);
}

void launch_zero(Sim& s, float* p, long n)
{
  hipLaunchKernelGGL(k_fill<float>, dim3(fill_grid(n)), dim3(256), 0, s.stream, p, 0.0f, n);
}

void launch_zero(Sim& s, int32_t* p, long n)
{
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(fill_grid(n)), dim3(256), 0, s.stream, p, 0, n);
}

void launch_copy(Sim& s, int32_t* dst, const int32_t* src, long n)
{
  hipLaunchKernelGGL(k_copy_i32, dim3(fill_grid(n)), dim3(256), 0, s.stream, dst, src, n);
}

namespace {

// The interior cells the field pushes update: the whole domain (chunk < 0), or
// one chunk's interior (PicChunk::push_bfd / push_efd on one chunk).  The
// kernels walk g.N cells from g.pad, so a chunk is a Geom whose N is the chunk's
// extent and whose pad is shifted by the chunk's cell offset.
Geom update_box(const Sim& s, int chunk)
{
  Geom b = s.g;
  if (chunk < 0)
    return b;
  for (int d = 0; d < 3; d++)
    if (b.has[d]) {
      b.N[d] = b.dims[d];
      b.pad[d] += s.offset[chunk * 3 + d];
    }
  return b;
}

template <typename T>
void push_bfd_impl(Sim& s, const T* E, T* B, T* ff, double delt, int chunk)
{
  const Geom  g    = update_box(s, chunk);
  const T     cflx = (T)(g.cc * delt / g.del[2]);
  const T     cfly = (T)(g.cc * delt / g.del[1]);
  const T     cflz = (T)(g.cc * delt / g.del[0]);
  long        n    = (long)g.N[0] * g.N[1] * g.N[2];
  const T*    F    = E;
  if (ff) {
    const double th = s.cfg.friedman;
    const double A  = 1 + 0.5 * th;
    const double B  = -th * (1 - 0.5 * th);
    const double C  = 0.5 * th * (1 - th) * (1 - th);
    hipLaunchKernelGGL(k_friedman_combine<T>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream,
                       s.g, E, ff, (T)A, (T)B, (T)C);
    F = ff;
  }
  if (g.dim == 3)
    hipLaunchKernelGGL((k_push_bfd<3, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, F, B,
                       cflx, cfly, cflz);
  else if (g.dim == 1)
    hipLaunchKernelGGL((k_push_bfd<1, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, F, B,
                       cflx, cfly, cflz);
  else
    hipLaunchKernelGGL((k_push_bfd<2, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, F, B,
                       cflx, cfly, cflz);
  hipLaunchKernelGGL(k_fill_ghost<T>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream, s.g, B, 3);
}

template <typename T>
void push_efd_impl(Sim& s, T* E, const T* B, const T* J, T* ff, double delt, int chunk)
{
  const Geom  g    = update_box(s, chunk);
  const T     cflx = (T)(g.cc * delt / g.del[2]);
  const T     cfly = (T)(g.cc * delt / g.del[1]);
  const T     cflz = (T)(g.cc * delt / g.del[0]);
  long        n    = (long)g.N[0] * g.N[1] * g.N[2];
  if (ff)
    hipLaunchKernelGGL(k_friedman_shift<T>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream,
                       s.g, E, ff, (T)s.cfg.friedman);
  if (g.dim == 3)
    hipLaunchKernelGGL((k_push_efd<3, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, E, B, J,
                       cflx, cfly, cflz, (T)delt);
  else if (g.dim == 1)
    hipLaunchKernelGGL((k_push_efd<1, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, E, B, J,
                       cflx, cfly, cflz, (T)delt);
  else
    hipLaunchKernelGGL((k_push_efd<2, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, E, B, J,
                       cflx, cfly, cflz, (T)delt);
  hipLaunchKernelGGL(k_fill_ghost<T>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream, s.g, E, 3);
}

// rho^{n+1} from the continuity equation that the Esirkepov current satisfies
// node by node (nix/esirkepov.hpp:76-337), written with the discrete
// divergence of BaseMaxwell::get_diverror_{2,3}d (pic/engine/maxwell.hpp:
// 203-229,350-381):
//   rho -= dt [(Jx(x+1) - Jx(x))/dx + (Jy(y+1) - Jy(y))/dy + (Jz(z+1) - Jz(z))/dz]
// on the interior of the periodic domain (upper neighbours wrap), after the
// current fold.  div(E) - rho is then carried exactly as E is updated.
template <int Dim, typename T>
__global__ void __launch_bounds__(256) k_rho_continuity(Geom g, T* __restrict__ J, T rdx, T rdy,
                                                        T rdz, T delt)
{
  long n   = (long)g.N[0] * g.N[1] * g.N[2];
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n)
    return;
  int       x  = (int)(tid % g.N[2]) + g.pad[2];
  long      r  = tid / g.N[2];
  int       y  = (int)(r % g.N[1]) + g.pad[1];
  int       z  = (int)(r / g.N[1]) + g.pad[0];
  const int xp = wrap(x + 1, g.pad[2], g.N[2]);
  const int yp = wrap(y + 1, g.pad[1], g.N[1]);
  const long v = g.vol;
  const long i = gidx(g, z, y, x);
  T div = (J[v + gidx(g, z, y, xp)] - J[v + i]) * rdx + (J[2 * v + gidx(g, z, yp, x)] - J[2 * v + i]) * rdy;
  if constexpr (Dim == 3) {
    const int zp = wrap(z + 1, g.pad[0], g.N[0]);
    div += (J[3 * v + gidx(g, zp, y, x)] - J[3 * v + i]) * rdz;
  }
  J[i] -= delt * div;
}

template <typename T>
void rho_continuity_impl(Sim& s, T* J, double delt)
{
  const Geom& g = s.g;
  const long  n = (long)g.N[0] * g.N[1] * g.N[2];
  const T     rdx = (T)(1 / g.del[2]), rdy = (T)(1 / g.del[1]), rdz = (T)(1 / g.del[0]);
  if (g.dim == 3)
    hipLaunchKernelGGL((k_rho_continuity<3, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, J,
                       rdx, rdy, rdz, (T)delt);
  else
    hipLaunchKernelGGL((k_rho_continuity<2, T>), dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, J,
                       rdx, rdy, rdz, (T)delt);
}

template <typename T>
void fold_impl(Sim& s, T* F, int k0, int k1)
{
  const Geom& g = s.g;
  for (int dir = 2; dir >= 0; dir--) {
    if (!g.has[dir] || g.pad[dir] == 0)
      continue;
    if (dir == 2 && g.xwall)
      continue; // no neighbour beyond a physical boundary: nothing is folded
    long n = dir == 2 ? (long)g.P[0] * g.P[1] : (dir == 1 ? (long)g.P[0] * g.P[2] : g.plane);
    hipLaunchKernelGGL(k_fold_line<T>, dim3(nblk(n, 256)), dim3(256), 0, s.stream, g, F, dir, k0,
                       k1);
  }
}

} // namespace

void launch_push_bfd(Sim& s, double delt, int chunk)
{
  PhaseTimer pt(s, 0);
  if (s.fp32)
    push_bfd_impl<float>(s, s.f_E, s.f_B, s.f_ff, delt, chunk);
  else
    push_bfd_impl<double>(s, s.d_E, s.d_B, s.d_ff, delt, chunk);
}

void launch_push_efd(Sim& s, double delt, int chunk)
{
  PhaseTimer pt(s, 5);
  if (s.fp32)
    push_efd_impl<float>(s, s.f_E, s.f_B, s.f_J, s.f_ff, delt, chunk);
  else
    push_efd_impl<double>(s, s.d_E, s.d_B, s.d_J, s.d_ff, delt, chunk);
}

void launch_init_friedman(Sim& s)
{
  if (s.d_ff)
    hipLaunchKernelGGL(k_friedman_init<double>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream,
                       s.g, s.d_E, s.d_ff);
  if (s.f_ff)
    hipLaunchKernelGGL(k_friedman_init<float>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream,
                       s.g, s.f_E, s.f_ff);
}

void launch_fill_ghost(Sim& s, double* f, int ncomp)
{
  hipLaunchKernelGGL(k_fill_ghost<double>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream, s.g,
                     f, ncomp);
}

void launch_fill_ghost(Sim& s, float* f, int ncomp)
{
  hipLaunchKernelGGL(k_fill_ghost<float>, dim3(nblk(s.g.vol, 256)), dim3(256), 0, s.stream, s.g,
                     f, ncomp);
}

void launch_fold(Sim& s, double* F, int k0, int k1, int phase)
{
  PhaseTimer pt(s, phase);
  fold_impl<double>(s, F, k0, k1);
}

void launch_fold(Sim& s, float* F, int k0, int k1, int phase)
{
  PhaseTimer pt(s, phase);
  fold_impl<float>(s, F, k0, k1);
}

void launch_rho_continuity(Sim& s, double delt)
{
  PhaseTimer pt(s, 7);
  if (s.fp32)
    rho_continuity_impl<float>(s, s.f_J, delt);
  else
    rho_continuity_impl<double>(s, s.d_J, delt);
}

void launch_fold_current(Sim& s, int k0, int k1)
{
  const int phase = k0 == 0 && k1 == 1 ? 7 : 4;
  if (s.fp32)
    launch_fold(s, s.f_J, k0, k1, phase);
  else
    launch_fold(s, s.d_J, k0, k1, phase);
}

} // namespace picnix
