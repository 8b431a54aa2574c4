// pn_capi.cpp -- extern "C" entry points (include/picnix.h) and simulation set-up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace picnix {

int queue_blocks(Sim& s, const void* kernel, size_t lds, long nitems)
{
  int per = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, lds));
  const long n = (long)std::max(per, 1) * std::max(s.ncu, 1);
  return (int)std::max<long>(1, std::min(n, nitems));
}
std::vector<int> make_chunk_map(int Cz, int Cy, int Cx);
void             shock_constants(Sim& s);
}

using namespace picnix;

struct pn_sim {
  Sim s;
};

Sim* picnix::sim_state(pn_sim* h) { return h ? &h->s : nullptr; }

namespace {


// Run `f` and convert failures into an error code + message on the handle.
// keep_perm: the call may leave a deferred sort pending (pn_step); every other
// entry point sees the particles in sorted order
template <typename F>
int guarded(pn_sim* h, F f, bool keep_perm = false)
{
  if (!h)
    return PN_ERR_ARG;
  try {
    if (!keep_perm)
      materialize(h->s);
    int rc = f(h->s);
    if (rc == PN_OK) {
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        h->s.err = std::string("HIP launch error: ") + hipGetErrorString(e);
        return PN_ERR_HIP;
      }
    }
    return rc;
  } catch (const HipError& e) {
    h->s.err = e.msg;
    return PN_ERR_HIP;
  } catch (const std::exception& e) {
    h->s.err = e.what();
    return PN_ERR_CONFIG;
  }
}

int fail(Sim& s, int code, const std::string& msg)
{
  s.err = msg;
  return code;
}

template <typename T>
T* dalloc(size_t n)
{
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

// ---------------------------------------------------------------------------
// geometry + chunk tables (nix/application.cpp:245-309, nix/chunk.cpp:118-237)
// ---------------------------------------------------------------------------
void setup_geometry(Sim& s)
{
  const Config& cfg = s.cfg;
  Geom&         g   = s.g;
  for (int d = 0; d < 3; d++)
    g.has[d] = !(cfg.N[d] == 1 && cfg.C[d] == 1);
  g.dim   = g.has[0] ? 3 : (g.has[1] ? 2 : 1);
  g.order = cfg.order;
  g.Nb    = (cfg.order + 3) / 2;
  g.Ns    = cfg.Ns;
  g.cc    = cfg.cc;
  for (int d = 0; d < 3; d++) {
    g.N[d]       = cfg.N[d];
    g.C[d]       = cfg.C[d];
    g.dims[d]    = cfg.N[d] / cfg.C[d];
    g.pad[d]     = g.has[d] ? g.Nb : 0;
    g.P[d]       = g.N[d] + 2 * g.pad[d];
    g.del[d]     = cfg.delh;
    g.glim[d][0] = 0.0;
    g.glim[d][1] = g.N[d] * g.del[d];
  }
  g.plane  = (long)g.P[1] * g.P[2];
  g.vol    = (long)g.P[0] * g.plane;
  // flat bin strides, XtensorParticle::flatindex (nix/xtensor_particle.hpp:231-240);
  // an absent dimension has Ub = Lb, so its stride factor is 2
  int ex = g.has[2] ? g.dims[2] + 1 : 2;
  int ey = g.has[1] ? g.dims[1] + 1 : 2;
  int ez = g.has[0] ? g.dims[0] + 1 : 1;
  g.bsy  = ex;
  g.bsz  = ex * ey;
  g.nbin = ex * ey * ez;
  g.nchunk = cfg.C[0] * cfg.C[1] * cfg.C[2];
  g.nseg   = g.nchunk * g.Ns;

  s.id_of_coord = make_chunk_map(cfg.C[0], cfg.C[1], cfg.C[2]);
  const int nc  = g.nchunk;
  s.coord.assign(nc * 3, 0);
  s.offset.assign(nc * 3, 0);
  s.lim.assign(nc * 6, 0.0);
  s.rng.assign(nc * 6, 0.0);
  s.nb.assign(nc * 27, 0);
  for (int cz = 0; cz < cfg.C[0]; cz++)
    for (int cy = 0; cy < cfg.C[1]; cy++)
      for (int cx = 0; cx < cfg.C[2]; cx++) {
        int id             = s.id_of_coord[(cz * cfg.C[1] + cy) * cfg.C[2] + cx];
        s.coord[id * 3 + 0] = cz;
        s.coord[id * 3 + 1] = cy;
        s.coord[id * 3 + 2] = cx;
      }
  for (int id = 0; id < nc; id++) {
    for (int d = 0; d < 3; d++) {
      int off              = s.coord[id * 3 + d] * cfg.N[d] / cfg.C[d];
      s.offset[id * 3 + d] = off;
      // nix/chunk.cpp:210-237 (set_coordinate)
      double lo              = off * g.del[d];
      double hi              = off * g.del[d] + g.dims[d] * g.del[d];
      s.lim[id * 6 + 2 * d]  = lo;
      s.lim[id * 6 + 2 * d + 1] = hi;
      // get_{x,y,z}range: +-max for an absent dimension (nix/chunk.hpp:295-335)
      s.rng[id * 6 + 2 * d]     = g.has[d] ? lo : -std::numeric_limits<double>::max();
      s.rng[id * 6 + 2 * d + 1] = g.has[d] ? hi : +std::numeric_limits<double>::max();
    }
    // periodic neighbours (nix/chunkvector.hpp:56-81, nix/chunkmap.cpp:100-112)
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int c3[3] = {s.coord[id * 3] + dz, s.coord[id * 3 + 1] + dy, s.coord[id * 3 + 2] + dx};
          for (int d = 0; d < 3; d++) {
            c3[d] = c3[d] >= 0 ? c3[d] : cfg.C[d] - 1;
            c3[d] = c3[d] < cfg.C[d] ? c3[d] : 0;
          }
          s.nb[id * 27 + 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)] =
              s.id_of_coord[(c3[0] * cfg.C[1] + c3[1]) * cfg.C[2] + c3[2]];
        }
  }
  // species: example/thermal/main.cpp:86-88
  int mp_max = 0;
  g.xwall    = cfg.problem == 1; // set_periodicity(1, 1, 0), example/shock/main.cpp:550-558
  if (cfg.problem == 1) {
    shock_constants(s);
    mp_max = s.shk.npc_max * g.dims[0] * g.dims[1] * g.dims[2];
  }
  for (int sp = 0; sp < g.Ns && cfg.problem == 0; sp++) {
    const SpeciesConfig& sc = cfg.species[sp];
    g.m[sp]                 = sc.ro / sc.np;
    g.q[sp]                 = sc.qm * g.m[sp];
    mp_max                  = std::max(mp_max, sc.np * g.dims[0] * g.dims[1] * g.dims[2]);
  }
  // capacity: round_up_alloc(mp * (1 + buffer_ratio)) (nix/particle.hpp:146-153)
  int req = static_cast<int>(mp_max * (1 + cfg.buffer_ratio));
  g.cap   = ((long)(req + 128) / 128) * 128;
}

void allocate(Sim& s)
{
  Geom&  g   = s.g;
  size_t vol = (size_t)g.vol;
  HIPCHK(hipStreamCreate(&s.stream));
  s.d_E = dalloc<double>(3 * vol);
  s.d_B = dalloc<double>(3 * vol);
  s.d_J = dalloc<double>(4 * vol);
  if (s.cfg.friedman != 0.0)
    s.d_ff = dalloc<double>(9 * vol);
  HIPCHK(hipMemset(s.d_E, 0, 3 * vol * sizeof(double)));
  HIPCHK(hipMemset(s.d_B, 0, 3 * vol * sizeof(double)));
  HIPCHK(hipMemset(s.d_J, 0, 4 * vol * sizeof(double)));

  s.d_wq = dalloc<int>(2 * kWqSlots);
  HIPCHK(hipMemset(s.d_wq, 0, 2 * kWqSlots * sizeof(int)));
  {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&s.ncu, hipDeviceAttributeMultiprocessorCount, dev));
    s.device = dev;
  }
  s.d_off = dalloc<int>(s.offset.size());
  s.d_lim = dalloc<double>(s.lim.size());
  s.d_rng = dalloc<double>(s.rng.size());
  s.d_nb  = dalloc<int>(s.nb.size());
  HIPCHK(hipMemcpy(s.d_off, s.offset.data(), s.offset.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s.d_lim, s.lim.data(), s.lim.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s.d_rng, s.rng.data(), s.rng.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s.d_nb, s.nb.data(), s.nb.size() * sizeof(int), hipMemcpyHostToDevice));

  size_t np = (size_t)g.nseg * g.cap;
  for (PArr* p : {&s.pa, &s.pb}) {
    if (s.fp32 && p == &s.pb)
      continue; // the FP32 form sorts P32 records
    p->x  = dalloc<double>(np);
    p->y  = dalloc<double>(np);
    p->z  = dalloc<double>(np);
    p->ux = dalloc<double>(np);
    p->uy = dalloc<double>(np);
    p->uz = dalloc<double>(np);
    p->id = dalloc<int64_t>(np);
  }
  for (int d = 0; d < 3; d++)
    s.d_xo[d] = dalloc<double>(np);
  if (s.fp32) {
    // FP32 state; pa, d_xo and the FP64 fields above stay as boundary mirrors
    s.f_E = dalloc<float>(3 * vol);
    s.f_B = dalloc<float>(3 * vol);
    s.f_J = dalloc<float>(4 * vol);
    HIPCHK(hipMemset(s.f_E, 0, 3 * vol * sizeof(float)));
    HIPCHK(hipMemset(s.f_B, 0, 3 * vol * sizeof(float)));
    HIPCHK(hipMemset(s.f_J, 0, 4 * vol * sizeof(float)));
    if (s.cfg.friedman != 0.0)
      s.f_ff = dalloc<float>(9 * vol);
    for (P32* q : {&s.qa, &s.qb}) {
      q->rx   = dalloc<float>(np);
      q->ry   = dalloc<float>(np);
      q->rz   = dalloc<float>(np);
      q->ux   = dalloc<float>(np);
      q->uy   = dalloc<float>(np);
      q->uz   = dalloc<float>(np);
      q->id   = dalloc<int64_t>(np);
      q->cell = dalloc<int32_t>(np);
    }
    for (int d = 0; d < 3; d++)
      s.f_ro[d] = dalloc<float>(np);
  }
  s.d_np = dalloc<int32_t>(g.nseg);
  HIPCHK(hipMemset(s.d_np, 0, g.nseg * sizeof(int32_t)));
  s.h_np.assign(g.nseg, 0);
  s.np_host_valid = true;

  SortScratch& ss = s.ss;
  ss.ntile        = (int)((g.cap + 255) / 256);
  ss.cellcnt      = dalloc<int32_t>((size_t)g.nseg * g.nbin);
  ss.cellstart    = dalloc<int32_t>((size_t)g.nseg * (g.nbin + 1));
  ss.key          = dalloc<int32_t>(np);
  ss.slot         = dalloc<int32_t>(np);
  ss.dseg         = dalloc<int32_t>(np);
  ss.didx         = dalloc<int32_t>(np);
  ss.mkey         = dalloc<int32_t>(np + kMemberPad);
  ss.msrc         = dalloc<int32_t>(np);
  ss.perm         = dalloc<int32_t>(np);
  ss.tilecnt      = dalloc<int32_t>((size_t)g.nseg * ss.ntile * 27);
  ss.nout         = dalloc<int32_t>((size_t)g.nseg * 27);
  ss.recvoff      = dalloc<int32_t>((size_t)g.nseg * 27);
  ss.npnew        = dalloc<int32_t>(g.nseg);
  ss.errflag      = dalloc<int32_t>(4);
  HIPCHK(hipMemset(ss.errflag, 0, 4 * sizeof(int32_t)));
  HIPCHK(hipMemset(ss.cellstart, 0, (size_t)g.nseg * (g.nbin + 1) * sizeof(int32_t)));
  HIPCHK(hipHostMalloc((void**)&s.h_flag, 4 * sizeof(int32_t), hipHostMallocDefault));
  std::memset(s.h_flag, 0, 4 * sizeof(int32_t));
  HIPCHK(hipEventCreateWithFlags(&s.flag_ev, hipEventDisableTiming));
}

void release(Sim& s)
{
  for (int k = 0; k < 2; k++)
    if (s.gexec[k])
      (void)hipGraphExecDestroy(s.gexec[k]);
  auto f = [](void* p) {
    if (p)
      (void)hipFree(p);
  };
  f(s.d_E);
  f(s.d_B);
  f(s.d_J);
  f(s.d_M);
  f(s.d_ff);
  f(s.d_off);
  f(s.d_lim);
  f(s.d_wq);
  f(s.d_rng);
  f(s.d_nb);
  for (PArr* p : {&s.pa, &s.pb}) {
    f(p->x);
    f(p->y);
    f(p->z);
    f(p->ux);
    f(p->uy);
    f(p->uz);
    f(p->id);
  }
  for (int d = 0; d < 3; d++)
    f(s.d_xo[d]);
  f(s.f_E);
  f(s.f_B);
  f(s.f_J);
  f(s.f_ff);
  for (P32* q : {&s.qa, &s.qb}) {
    f(q->rx);
    f(q->ry);
    f(q->rz);
    f(q->ux);
    f(q->uy);
    f(q->uz);
    f(q->id);
    f(q->cell);
  }
  for (int d = 0; d < 3; d++)
    f(s.f_ro[d]);
  f(s.d_np);
  SortScratch& ss = s.ss;
  for (void* p : {(void*)ss.cellcnt, (void*)ss.cellstart, (void*)ss.key, (void*)ss.slot,
                  (void*)ss.dseg, (void*)ss.didx, (void*)ss.mkey, (void*)ss.msrc, (void*)ss.perm, (void*)ss.tilecnt,
                  (void*)ss.nout, (void*)ss.recvoff, (void*)ss.npnew, (void*)ss.errflag})
    f(p);
  f(s.d_stage);
  f(s.d_cs_saved);
  if (s.h_flag)
    (void)hipHostFree(s.h_flag);
  if (s.flag_ev)
    (void)hipEventDestroy(s.flag_ev);
  shock_release(s);
  if (s.stream)
    (void)hipStreamDestroy(s.stream);
}

void ensure_stage(Sim& s, size_t elems)
{
  if (elems <= s.stage_elems)
    return;
  if (s.d_stage)
    HIPCHK(hipFree(s.d_stage));
  s.d_stage     = dalloc<double>(elems);
  s.stage_elems = elems;
}

void refresh_np(Sim& s)
{
  if (s.np_host_valid)
    return;
  HIPCHK(hipMemcpyAsync(s.h_np.data(), s.d_np, sizeof(int32_t) * s.g.nseg, hipMemcpyDeviceToHost,
                        s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  s.np_host_valid = true;
}

int check_device_flags(Sim& s)
{
  int32_t flags[4];
  HIPCHK(hipMemcpyAsync(flags, s.ss.errflag, sizeof(flags), hipMemcpyDeviceToHost, s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  if (flags[0])
    return fail(s, PN_ERR_CAPACITY,
                "particle segment capacity exceeded (raise [application.option] buffer_ratio)");
  return PN_OK;
}

// ---------------------------------------------------------------------------
// loader: example/beam/main.cpp:67-116 (example/thermal is the zero-drift case),
// with the reference's own generators: std::mt19937_64 seeded with the chunk id
// (seed_type = 'fixed', pic/pic_chunk.cpp:242-255), uniform_real_distribution
// and normal_distribution from libstdc++, mtp re-seeded per species.
// ---------------------------------------------------------------------------
void load_particles(Sim& s)
{
  const Geom& g   = s.g;
  const int   nc  = g.nchunk;
  size_t      tot = (size_t)g.nseg * g.cap;
  std::vector<double>  hx(tot), hy(tot), hz(tot), hux(tot), huy(tot), huz(tot);
  std::vector<int64_t> hid(tot);
  std::vector<int32_t> hnp(g.nseg, 0);
  std::vector<uint32_t> seeds(nc);
  {
    std::random_device rd;
    for (int c = 0; c < nc; c++)
      seeds[c] = s.cfg.seed_type == "fixed" ? (uint32_t)(c + s.cfg.seed_offset) : rd();
  }
  auto work = [&](int c) {
    int                                    seed = (int)seeds[c];
    std::mt19937_64                        mtp(seed);
    std::mt19937_64                        mtv(seed);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    std::normal_distribution<double>       normal(0.0, 1.0);
    const double* lim = &s.lim[c * 6];
    for (int sp = 0; sp < g.Ns; sp++) {
      const SpeciesConfig& sc  = s.cfg.species[sp];
      int                  mp  = sc.np * g.dims[0] * g.dims[1] * g.dims[2];
      int64_t              id  = (int64_t)mp * c;
      size_t               seg = (size_t)c * g.Ns + sp;
      size_t               b   = seg * g.cap;
      mtp.seed(seed);
      for (int ip = 0; ip < mp; ip++) {
        hx[b + ip]  = uniform(mtp) * (lim[5] - lim[4]) + lim[4];
        hy[b + ip]  = uniform(mtp) * (lim[3] - lim[2]) + lim[2];
        hz[b + ip]  = uniform(mtp) * (lim[1] - lim[0]) + lim[0];
        hux[b + ip] = normal(mtv) * sc.vt + sc.vx;
        huy[b + ip] = normal(mtv) * sc.vt + sc.vy;
        huz[b + ip] = normal(mtv) * sc.vt + sc.vz;
        hid[b + ip] = id + ip;
      }
      hnp[seg] = mp;
    }
  };
  unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthr; t++)
    pool.emplace_back([&, t] {
      for (int c = (int)t; c < nc; c += (int)nthr)
        work(c);
    });
  for (auto& th : pool)
    th.join();
  auto up = [&](void* d, const void* h, size_t bytes) {
    HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  };
  up(s.pa.x, hx.data(), tot * 8);
  up(s.pa.y, hy.data(), tot * 8);
  up(s.pa.z, hz.data(), tot * 8);
  up(s.pa.ux, hux.data(), tot * 8);
  up(s.pa.uy, huy.data(), tot * 8);
  up(s.pa.uz, huz.data(), tot * 8);
  up(s.pa.id, hid.data(), tot * 8);
  up(s.d_np, hnp.data(), g.nseg * sizeof(int32_t));
  s.h_np          = hnp;
  s.np_host_valid = true;
  if (s.fp32)
    launch32_from_f64(s, 0, g.nseg);
}

void init_fields(Sim& s)
{
  const Geom&         g = s.g;
  std::vector<double> h(3 * (size_t)g.vol, 0.0);
  for (int which = 0; which < 2; which++) {
    const double* v = which == 0 ? s.cfg.E0 : s.cfg.B0;
    std::fill(h.begin(), h.end(), 0.0);
    for (int z = g.pad[0]; z < g.pad[0] + g.N[0]; z++)
      for (int y = g.pad[1]; y < g.pad[1] + g.N[1]; y++)
        for (int x = g.pad[2]; x < g.pad[2] + g.N[2]; x++)
          for (int k = 0; k < 3; k++)
            h[k * g.vol + ((long)z * g.P[1] + y) * g.P[2] + x] = v[k];
    HIPCHK(hipMemcpy(which == 0 ? s.d_E : s.d_B, h.data(), h.size() * sizeof(double),
                     hipMemcpyHostToDevice));
  }
  launch_fill_ghost(s, s.d_E, 3);
  launch_fill_ghost(s, s.d_B, 3);
  if (s.fp32)
    launch32_fields_from_f64(s, PN_FIELD_UF);
}

// chunk = -1: every chunk (one omp-for phase of push_openmp); chunk >= 0: that
// chunk alone, as PicChunk's operators act (pic/pic_chunk.hpp:90-142).
int check_chunk_arg(Sim& s, int chunk)
{
  if (chunk == -1)
    return PN_OK;
  if (chunk < 0 || chunk >= s.g.nchunk)
    return fail(s, PN_ERR_ARG, "chunk index out of range (-1 selects every chunk)");
  if (s.fp32)
    return fail(s, PN_ERR_UNSUPPORTED,
                "per-chunk operator calls are built for precision = fp64; pass chunk = -1");
  if (s.cfg.problem != 0)
    return fail(s, PN_ERR_UNSUPPORTED,
                "per-chunk operator calls are built for periodic problems; pass chunk = -1");
  return PN_OK;
}

// The particle operators on one chunk run the one-thread-per-particle kernels,
// which stop at each segment's count: the other chunks' counts are set to zero
// for the call and restored afterwards (the chunk's own count stays as the
// operator leaves it).
struct ChunkWindow {
  Sim&                 s;
  int                  c;
  bool                 generic;
  std::vector<int32_t> saved;
  ChunkWindow(Sim& s_, int c_) : s(s_), c(c_), generic(s_.generic_kernels)
  {
    const Geom& g = s.g;
    refresh_np(s);
    saved = s.h_np;
    std::vector<int32_t> m(g.nseg, 0);
    for (int sp = 0; sp < g.Ns; sp++)
      m[c * g.Ns + sp] = saved[c * g.Ns + sp];
    HIPCHK(hipMemcpyAsync(s.d_np, m.data(), sizeof(int32_t) * g.nseg, hipMemcpyHostToDevice,
                          s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    s.generic_kernels = true;
  }
  void restore()
  {
    const Geom&          g = s.g;
    std::vector<int32_t> now(g.nseg);
    HIPCHK(hipMemcpyAsync(now.data(), s.d_np, sizeof(int32_t) * g.nseg, hipMemcpyDeviceToHost,
                          s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    for (int seg = 0; seg < g.nseg; seg++)
      if (seg / g.Ns != c)
        now[seg] = saved[seg];
    HIPCHK(hipMemcpyAsync(s.d_np, now.data(), sizeof(int32_t) * g.nseg, hipMemcpyHostToDevice,
                          s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    s.h_np          = now;
    s.np_host_valid = true;
    s.generic_kernels = generic;
  }
};

// PicChunk::sort_particle on one chunk: count and sort its segments with the
// others' counts masked, then put back the other segments' records (the sort
// writes the other record set), their bin starts and their counts.
void sort_one_chunk(Sim& s, int c)
{
  const Geom&  g   = s.g;
  const size_t ncs = (size_t)g.nseg * (g.nbin + 1);
  if (!s.d_cs_saved)
    s.d_cs_saved = dalloc<int32_t>(ncs);
  HIPCHK(hipMemcpyAsync(s.d_cs_saved, s.ss.cellstart, ncs * sizeof(int32_t),
                        hipMemcpyDeviceToDevice, s.stream));
  const bool  was_sorted = s.sorted;
  ChunkWindow win(s, c);
  launch_sort(s); // records of chunk c -> the other set, which becomes current
  const size_t cap  = (size_t)g.cap;
  const size_t lo   = (size_t)c * g.Ns * cap;        // first record of chunk c
  const size_t hi   = (size_t)(c + 1) * g.Ns * cap;  // one past its last
  const size_t tot  = (size_t)g.nseg * cap;
  auto         keep = [&](auto* dst, const auto* src) {
    using T = std::remove_pointer_t<decltype(dst)>;
    if (lo > 0)
      HIPCHK(hipMemcpyAsync(dst, src, lo * sizeof(T), hipMemcpyDeviceToDevice, s.stream));
    if (hi < tot)
      HIPCHK(hipMemcpyAsync(dst + hi, src + hi, (tot - hi) * sizeof(T), hipMemcpyDeviceToDevice,
                            s.stream));
  };
  keep(s.pa.x, s.pb.x);
  keep(s.pa.y, s.pb.y);
  keep(s.pa.z, s.pb.z);
  keep(s.pa.ux, s.pb.ux);
  keep(s.pa.uy, s.pb.uy);
  keep(s.pa.uz, s.pb.uz);
  keep(s.pa.id, s.pb.id);
  const size_t clo = (size_t)c * g.Ns * (g.nbin + 1), chi = (size_t)(c + 1) * g.Ns * (g.nbin + 1);
  win.restore();
  // the sort clears every segment's bin starts it does not own: put them back
  if (clo > 0)
    HIPCHK(hipMemcpyAsync(s.ss.cellstart, s.d_cs_saved, clo * sizeof(int32_t),
                          hipMemcpyDeviceToDevice, s.stream));
  if (chi < ncs)
    HIPCHK(hipMemcpyAsync(s.ss.cellstart + chi, s.d_cs_saved + chi, (ncs - chi) * sizeof(int32_t),
                          hipMemcpyDeviceToDevice, s.stream));
  if (s.chunk_sorted.empty())
    s.chunk_sorted.assign(g.nchunk, was_sorted ? 1 : 0);
  s.chunk_sorted[c] = 1;
  bool all = true;
  for (char v : s.chunk_sorted)
    all = all && v;
  s.sorted  = all;
  s.counted = false;
}

int check_seg(Sim& s, int chunk, int species)
{
  if (chunk < 0 || chunk >= s.g.nchunk || species < 0 || species >= s.g.Ns)
    return fail(s, PN_ERR_ARG, "chunk or species index out of range");
  return PN_OK;
}

} // namespace

extern "C" {

int pn_create(const char* toml, int load, pn_sim** out)
{
  if (!toml || !out)
    return PN_ERR_ARG;
  *out   = nullptr;
  auto h = new pn_sim();
  Sim& s = h->s;
  try {
    TomlMap kv;
    s.toml = toml;
    if (!parse_toml(toml, kv, s.err) || !build_config(kv, s.cfg, s.err)) {
      // keep the handle so that pn_last_error can report the reason
      *out = h;
      return PN_ERR_CONFIG;
    }
    if ((s.cfg.precision != "fp64" && s.cfg.precision != "fp32") || s.cfg.Ns > kMaxSpecies) {
      s.err = "configuration uses an option this build does not implement yet "
              "(supported: precision = fp64 or fp32, Ns <= 8)";
      *out  = h;
      return PN_ERR_UNSUPPORTED;
    }
    s.fp32 = s.cfg.precision == "fp32";
    if (s.cfg.problem == 1 && s.cfg.friedman != 0.0) {
      s.err = "the shock problem is built without the Friedman filter";
      *out  = h;
      return PN_ERR_UNSUPPORTED;
    }
    setup_geometry(s);
    if (const char* k = std::getenv("PICNIX_KERNELS"))
      s.generic_kernels = std::string(k) == "generic";
    if (const char* k = std::getenv("PICNIX_MARCH_ROWS"))
      s.march_rows = std::max(0, std::atoi(k));
    if (const char* k = std::getenv("PICNIX_WORK_QUEUE")) // persistent row kernels
      s.dep_queue = s.vel_queue = std::atoi(k) != 0 ? 1 : 0;
    if (const char* k = std::getenv("PICNIX_FUSED_PUSH")) // gather + position push in one kernel
      s.fuse_push = std::atoi(k) != 0 ? 1 : 0;
    if (const char* k = std::getenv("PICNIX_FP32_TILE")) { // FP32 deposit tile form
      const std::string v(k);
      s.tile32 = v == "staged" ? 0 : (v == "plain" ? 2 : (v == "comp" ? 3 : 1));
    }
    if (const char* k = std::getenv("PICNIX_STAGE_CELLS")) // FP32 staged tile: cells per phase
      s.stage_cells = std::max(0, std::atoi(k));
    if (const char* k = std::getenv("PICNIX_DEFER_SORT")) // 0: pn_step sorts in place
      s.defer_sort = std::atoi(k) != 0;
    if (const char* k = std::getenv("PICNIX_RHO")) // "deposit": rho from particles every step
      s.rho_deposit = std::string(k) == "deposit";
    allocate(s);
    energy_static_init(s);
    if (s.cfg.problem == 1) {
      shock_setup(s);
      shock_init_fields(s);
    } else {
      init_fields(s);
    }
    launch_init_friedman(s); // PicChunk::init_friedman in the examples' setup
    if (load) {
      if (s.cfg.problem == 1)
        shock_load_particles(s);
      else
        load_particles(s);
      // initial sort, pic/pic_chunk.cpp:447-453 (count + sort per species)
      if (s.fp32)
        launch32_sort(s);
      else
        launch_sort(s);
    }
    HIPCHK(hipStreamSynchronize(s.stream));
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    s.err = e.msg;
    *out  = h;
    return PN_ERR_HIP;
  }
  *out = h;
  return PN_OK;
}

int pn_destroy(pn_sim* h)
{
  if (!h)
    return PN_ERR_ARG;
  release(h->s);
  delete h;
  return PN_OK;
}

const char* pn_last_error(const pn_sim* h) { return h ? h->s.err.c_str() : "null handle"; }

int pn_num_chunks(const pn_sim* h, int* n)
{
  if (!h || !n)
    return PN_ERR_ARG;
  *n = h->s.g.nchunk;
  return PN_OK;
}

int pn_num_species(const pn_sim* h, int* ns)
{
  if (!h || !ns)
    return PN_ERR_ARG;
  *ns = h->s.g.Ns;
  return PN_OK;
}

int pn_chunk_info(const pn_sim* h, int chunk, int* info)
{
  if (!h || !info || chunk < 0 || chunk >= h->s.g.nchunk)
    return PN_ERR_ARG;
  const Sim& s = h->s;
  const Geom& g = s.g;
  int ng = 1;
  for (int d = 0; d < 3; d++) {
    info[d]     = s.coord[chunk * 3 + d];
    info[3 + d] = s.offset[chunk * 3 + d];
    info[6 + d] = g.dims[d] + 2 * g.Nb;
    ng *= g.dims[d] + 2 * g.Nb;
  }
  info[9]  = g.Nb;
  info[10] = ng;
  return PN_OK;
}

int pn_get_np(const pn_sim* hc, int chunk, int species, int64_t* np)
{
  pn_sim* h = const_cast<pn_sim*>(hc);
  return guarded(h, [&](Sim& s) {
    if (!np)
      return PN_ERR_ARG;
    int rc = check_seg(s, chunk, species);
    if (rc)
      return rc;
    refresh_np(s);
    *np = s.h_np[chunk * s.g.Ns + species];
    return PN_OK;
  });
}

int pn_get_particles(pn_sim* h, int chunk, int species, double* xu7, int64_t cap, int64_t* np)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_seg(s, chunk, species);
    if (rc)
      return rc;
    if ((rc = check_device_flags(s)))
      return rc;
    refresh_np(s);
    int     seg = chunk * s.g.Ns + species;
    int64_t n   = s.h_np[seg];
    if (np)
      *np = n;
    if (!xu7 || n == 0)
      return PN_OK;
    if (cap < n)
      return fail(s, PN_ERR_ARG, "output buffer smaller than the particle count");
    ensure_stage(s, (size_t)n * 7);
    if (s.fp32)
      launch32_to_f64(s, seg, 1, false);
    launch_soa_to_aos(s, seg, n);
    HIPCHK(hipMemcpyAsync(xu7, s.d_stage, n * 7 * sizeof(double), hipMemcpyDeviceToHost,
                          s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    return PN_OK;
  });
}

int pn_set_particles(pn_sim* h, int chunk, int species, const double* xu7, int64_t np)
{
  return guarded(h, [&](Sim& s) {
    s.rho_valid = false; // particles move or J is written outside pn_step
    int rc = check_seg(s, chunk, species);
    if (rc)
      return rc;
    if (np < 0 || np > s.g.cap || (np > 0 && !xu7))
      return fail(s, PN_ERR_CAPACITY, "particle count exceeds the segment capacity");
    refresh_np(s);
    int seg = chunk * s.g.Ns + species;
    if (np > 0) {
      ensure_stage(s, (size_t)np * 7);
      HIPCHK(hipMemcpyAsync(s.d_stage, xu7, np * 7 * sizeof(double), hipMemcpyHostToDevice,
                            s.stream));
      launch_aos_to_soa(s, seg, np);
    }
    s.h_np[seg] = (int32_t)np;
    HIPCHK(hipMemcpyAsync(s.d_np + seg, &s.h_np[seg], sizeof(int32_t), hipMemcpyHostToDevice,
                          s.stream));
    if (s.fp32 && np > 0)
      launch32_from_f64(s, seg, 1);
    HIPCHK(hipStreamSynchronize(s.stream));
    s.sorted  = false;
    s.chunk_sorted.clear();
    s.counted = false;
    return PN_OK;
  });
}

int pn_get_particles_old(pn_sim* h, int chunk, int species, double* xv3, int64_t cap)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_seg(s, chunk, species);
    if (rc)
      return rc;
    refresh_np(s);
    int     seg = chunk * s.g.Ns + species;
    int64_t n   = s.h_np[seg];
    if (!xv3 || cap < n)
      return fail(s, PN_ERR_ARG, "output buffer smaller than the particle count");
    std::vector<double> t(n);
    if (s.fp32) {
      launch32_to_f64(s, seg, 1, true);
      HIPCHK(hipStreamSynchronize(s.stream));
    }
    for (int d = 0; d < 3; d++) {
      HIPCHK(hipMemcpy(t.data(), s.d_xo[d] + (size_t)seg * s.g.cap, n * sizeof(double),
                       hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; i++)
        xv3[i * 3 + d] = t[i];
    }
    return PN_OK;
  });
}

int pn_get_gindex(pn_sim* h, int chunk, int species, int32_t* gindex, int64_t cap)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_seg(s, chunk, species);
    if (rc)
      return rc;
    if (!s.counted)
      return fail(s, PN_ERR_STATE, "no fresh count: call pn_push_position first");
    refresh_np(s);
    int     seg = chunk * s.g.Ns + species;
    int64_t n   = s.h_np[seg];
    if (!gindex || cap < n)
      return fail(s, PN_ERR_ARG, "output buffer smaller than the particle count");
    HIPCHK(hipMemcpy(gindex, s.ss.key + (size_t)seg * s.g.cap, n * sizeof(int32_t),
                     hipMemcpyDeviceToHost));
    int ng = (s.g.dims[0] + 2 * s.g.Nb) * (s.g.dims[1] + 2 * s.g.Nb) * (s.g.dims[2] + 2 * s.g.Nb);
    for (int64_t i = 0; i < n; i++)
      if (gindex[i] < 0)
        gindex[i] = ng;
    return PN_OK;
  });
}

int pn_get_pindex(pn_sim* h, int chunk, int species, int32_t* pindex)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_seg(s, chunk, species);
    if (rc)
      return rc;
    if (!s.sorted)
      return fail(s, PN_ERR_STATE, "particles are not sorted");
    if (!pindex)
      return PN_ERR_ARG;
    int                  seg = chunk * s.g.Ns + species;
    std::vector<int32_t> cs(s.g.nbin + 1);
    HIPCHK(hipMemcpy(cs.data(), s.ss.cellstart + (size_t)seg * (s.g.nbin + 1),
                     cs.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    int ng = (s.g.dims[0] + 2 * s.g.Nb) * (s.g.dims[1] + 2 * s.g.Nb) * (s.g.dims[2] + 2 * s.g.Nb);
    for (int i = 0; i <= ng; i++)
      pindex[i] = i < s.g.nbin ? cs[i] : cs[s.g.nbin];
    return PN_OK;
  });
}

int pn_get_fields(pn_sim* h, int chunk, int which, double* data)
{
  return guarded(h, [&](Sim& s) {
    if (chunk < 0 || chunk >= s.g.nchunk || !data ||
        (which != PN_FIELD_UF && which != PN_FIELD_UJ && which != PN_FIELD_UM))
      return fail(s, PN_ERR_ARG, "invalid chunk, field selector or buffer");
    if (which == PN_FIELD_UM && !s.d_M)
      return fail(s, PN_ERR_STATE, "no moments: call pn_deposit_moment first");
    const Geom& g  = s.g;
    size_t      nn = (size_t)(g.dims[0] + 2 * g.Nb) * (g.dims[1] + 2 * g.Nb) * (g.dims[2] + 2 * g.Nb);
    int         nc = which == PN_FIELD_UF ? 6 : (which == PN_FIELD_UJ ? 4 : 14 * g.Ns);
    ensure_stage(s, nn * nc);
    if (s.fp32 && which != PN_FIELD_UM)
      launch32_fields_to_f64(s, which);
    launch_chunk_fields_get(s, chunk, which, s.d_stage);
    HIPCHK(hipMemcpyAsync(data, s.d_stage, nn * nc * sizeof(double), hipMemcpyDeviceToHost,
                          s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    return PN_OK;
  });
}

int pn_set_fields(pn_sim* h, int chunk, int which, const double* data)
{
  return guarded(h, [&](Sim& s) {
    s.rho_valid = false; // particles move or J is written outside pn_step
    if (chunk < 0 || chunk >= s.g.nchunk || !data || (which != PN_FIELD_UF && which != PN_FIELD_UJ))
      return fail(s, PN_ERR_ARG, "invalid chunk, field selector or buffer");
    const Geom& g  = s.g;
    size_t      nn = (size_t)(g.dims[0] + 2 * g.Nb) * (g.dims[1] + 2 * g.Nb) * (g.dims[2] + 2 * g.Nb);
    int         nc = which == PN_FIELD_UF ? 6 : 4;
    ensure_stage(s, nn * nc);
    HIPCHK(hipMemcpyAsync(s.d_stage, data, nn * nc * sizeof(double), hipMemcpyHostToDevice,
                          s.stream));
    if (s.fp32)
      launch32_fields_to_f64(s, which); // the mirror holds the other chunks' values
    launch_chunk_fields_set(s, chunk, which, s.d_stage);
    if (which == PN_FIELD_UF) {
      launch_fill_ghost(s, s.d_E, 3);
      launch_fill_ghost(s, s.d_B, 3);
    }
    if (s.fp32)
      launch32_fields_from_f64(s, which);
    HIPCHK(hipStreamSynchronize(s.stream));
    return PN_OK;
  });
}

int pn_push_bfd(pn_sim* h, int chunk, double delt)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_chunk_arg(s, chunk);
    if (rc)
      return rc;
    launch_push_bfd(s, delt, chunk);
    return PN_OK;
  });
}

int pn_push_efd(pn_sim* h, int chunk, double delt)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_chunk_arg(s, chunk);
    if (rc)
      return rc;
    if (chunk >= 0 && s.cfg.friedman != 0.0)
      return fail(s, PN_ERR_UNSUPPORTED,
                  "per-chunk push_efd with the Friedman filter is not built; pass chunk = -1");
    launch_push_efd(s, delt, chunk);
    return PN_OK;
  });
}

int pn_push_velocity(pn_sim* h, int chunk, double delt)
{
  return guarded(h, [&](Sim& s) {
    int rc = check_chunk_arg(s, chunk);
    if (rc)
      return rc;
    if (chunk >= 0) {
      ChunkWindow win(s, chunk);
      launch_push_velocity(s, delt);
      win.restore();
    } else if (s.fp32) {
      launch32_push_velocity(s, delt);
    } else {
      launch_push_velocity(s, delt);
    }
    return PN_OK;
  });
}

int pn_push_position(pn_sim* h, int chunk, double delt)
{
  return guarded(h, [&](Sim& s) {
    s.rho_valid = false; // particles move or J is written outside pn_step
    int rc = check_chunk_arg(s, chunk);
    if (rc)
      return rc;
    bool was_sorted = s.sorted;
    if (s.fp32)
      launch32_push_position(s, delt);
    else
      launch_push_position(s, delt, chunk);
    if (s.g.xwall)
      shock_position_boundary(s);
    // the particle order is unchanged, so cellstart still describes the old
    // positions the deposit walks
    s.sorted = was_sorted;
    return PN_OK;
  });
}

int pn_deposit_current(pn_sim* h, int chunk, double delt)
{
  return guarded(h, [&](Sim& s) {
    s.rho_valid = false; // particles move or J is written outside pn_step
    int rc = check_chunk_arg(s, chunk);
    if (rc)
      return rc;
    if (chunk >= 0) {
      // The reference clears each chunk's uj (ghosts included) and the Cur
      // exchange adds the ghosts into the neighbours.  Here every chunk adds
      // into one global J, so J is cleared once per round of per-chunk
      // deposits: at the first call, or when this chunk has already deposited
      // since the last clear.
      const Geom& g = s.g;
      if (s.dep_round.empty() || s.dep_round[chunk]) {
        launch_zero(s, s.d_J, 4 * g.vol);
        s.dep_round.assign(g.nchunk, 0);
      }
      s.dep_round[chunk] = 1;
      ChunkWindow win(s, chunk);
      s.dep_accumulate = true;
      launch_deposit(s, delt);
      s.dep_accumulate = false;
      win.restore();
    } else if (s.fp32) {
      s.dep_round.clear();
      launch32_deposit(s, delt, RhoMode::Now);
    } else {
      s.dep_round.clear();
      launch_deposit(s, delt);
    }
    return PN_OK;
  });
}

int pn_init_friedman(pn_sim* h)
{
  return guarded(h, [&](Sim& s) {
    launch_init_friedman(s);
    return PN_OK;
  });
}

int pn_deposit_moment(pn_sim* h, int chunk)
{
  return guarded(h, [&](Sim& s) {
    if (chunk != -1)
      return fail(s, PN_ERR_UNSUPPORTED, "the moment deposit runs on every chunk; pass chunk = -1");
    if (s.fp32)
      return fail(s, PN_ERR_UNSUPPORTED, "moments are not built for precision = fp32");
    if (!s.d_M)
      s.d_M = dalloc<double>(14 * (size_t)s.g.Ns * s.g.vol);
    launch_deposit_moment(s);
    return PN_OK;
  });
}

int pn_sort_particle(pn_sim* h, int chunk)
{
  return guarded(h, [&](Sim& s) {
    s.rho_valid = false; // particles move or J is written outside pn_step
    int rc = check_chunk_arg(s, chunk);
    if (rc)
      return rc;
    if (chunk >= 0) {
      sort_one_chunk(s, chunk);
    } else {
      s.chunk_sorted.clear();
      if (s.fp32)
        launch32_sort(s);
      else
        launch_sort(s);
    }
    return PN_OK;
  });
}

int pn_exchange(pn_sim* h, int mode)
{
  return guarded(h, [&](Sim& s) {
    switch (mode) {
    case PN_BOUNDARY_EMF:
      if (s.fp32) {
        launch_fill_ghost(s, s.f_E, 3);
        launch_fill_ghost(s, s.f_B, 3);
      } else {
        launch_fill_ghost(s, s.d_E, 3);
        launch_fill_ghost(s, s.d_B, 3);
      }
      launch_wall_emf(s); // set_boundary_field(BoundaryEmf) after the unpack
      return PN_OK;
    case PN_BOUNDARY_CUR:
      s.rho_valid = false;
      s.dep_round.clear(); // the fold ends a round of per-chunk deposits
      launch_fold_current(s);
      return PN_OK;
    case PN_BOUNDARY_PARTICLE:
      s.rho_valid = false;
      if (!s.counted)
        return fail(s, PN_ERR_STATE, "particle exchange needs the count of pn_push_position");
      if (s.g.xwall)
        shock_inject(s); // set_boundary_pack(BoundaryParticle), pic/pic_chunk.cpp:305-309
      s.chunk_sorted.clear();
      if (s.fp32)
        launch32_migrate(s);
      else
        launch_migrate(s);
      return PN_OK;
    case PN_BOUNDARY_MOM:
      if (s.fp32)
        return fail(s, PN_ERR_UNSUPPORTED, "moments are not built for precision = fp32");
      if (!s.d_M)
        return fail(s, PN_ERR_STATE, "moment exchange needs pn_deposit_moment first");
      launch_moment_exchange(s, false);
      return PN_OK;
    default:
      return fail(s, PN_ERR_ARG, "invalid boundary mode");
    }
  });
}

// PicApplication::push_openmp (pic/pic_application.cpp:219-292): per chunk
// bfd(dt/2), velocity, position(+count), current, bfd(dt/2); then the current
// fold, efd, particle migration + sort, EMF exchange (done by the ghost refresh
// that follows every field update).
namespace {

// rho: after a step, rho is the deposit of the current positions, so the next
// step can carry it by the continuity equation the Esirkepov current satisfies
// (launch_rho_continuity) instead of depositing it again; the first step after
// anything else moved particles or wrote J deposits it (after the sort, where
// the new positions are grouped by cell).
// whether this step carries rho by continuity (PICNIX_RHO=deposit: never; every
// step then deposits rho from the particles, as the reference does)
bool carry_rho(const Sim& s) { return s.rho_valid && !s.rho_deposit; }

// pn_step leaves the record move of its sort to the next step's row gather
// (periodic problems whose gather takes the row form)
bool defers_sort(const Sim& s)
{
  return s.defer_sort && gather_takes_perm(s);
}

void step_launches32(Sim& s, double delt)
{
  launch_push_bfd(s, 0.5 * delt);
  bool was_sorted = s.sorted;
  if (!launch32_push_velocity(s, delt, s.fuse_push != 0))
    launch32_push_position(s, delt);
  s.sorted = was_sorted;
  const RhoDone rho = launch32_deposit(s, delt, carry_rho(s) ? RhoMode::Keep : RhoMode::Defer);
  launch_push_bfd(s, 0.5 * delt);
  launch_fold_current(s, 1, 4);
  if (rho == RhoDone::Kept)
    launch_rho_continuity(s, delt);
  launch_push_efd(s, delt);
  launch32_migrate(s, defers_sort(s) && rho != RhoDone::Deferred);
  if (rho == RhoDone::Deferred) {
    launch32_rho(s);
    launch_fold_current(s, 0, 1);
  }
  s.rho_valid = true;
}

// The shock problem: the same operator order with the physical x boundaries --
// the wall reflection and the upstream redraw in the position push, the EMF
// boundary after the field update, the injection before the migration.  The
// injected particles carry charge that no current brought, so rho is deposited
// every step rather than carried by continuity: where the current kernel leaves
// it, after the sort (grouped by the new cells) minus the particles the
// injection has just appended, which the reference's deposit does not see.
void step_launches_wall(Sim& s, double delt)
{
  launch_push_bfd(s, 0.5 * delt);
  bool was_sorted = s.sorted;
  if (s.fp32) {
    if (!launch32_push_velocity(s, delt, s.fuse_push != 0))
      launch32_push_position(s, delt);
  } else {
    if (!launch_push_velocity(s, delt, s.fuse_push != 0))
      launch_push_position(s, delt);
  }
  shock_position_boundary(s);
  s.sorted = was_sorted;
  const RhoDone rho =
      s.fp32 ? launch32_deposit(s, delt, RhoMode::Defer) : launch_deposit(s, delt, RhoMode::Defer);
  launch_push_bfd(s, 0.5 * delt);
  launch_fold_current(s, rho == RhoDone::Deposited ? 0 : 1, 4);
  launch_push_efd(s, delt);
  launch_wall_emf(s);
  shock_inject(s);
  // the rho deposit below reads the particles through a pending sort, so the
  // record move is left to the next row gather
  if (s.fp32)
    launch32_migrate(s, defers_sort(s));
  else
    launch_migrate(s, defers_sort(s));
  if (rho == RhoDone::Deferred) {
    if (s.fp32)
      launch32_rho(s);
    else
      launch_rho(s, true);
    launch_rho_unject(s, s.shk.last_ninj);
    launch_fold_current(s, 0, 1);
  }
  s.rho_valid = false;
}

void step_launches(Sim& s, double delt)
{
  if (s.g.xwall) {
    step_launches_wall(s, delt);
    return;
  }
  if (s.fp32) {
    step_launches32(s, delt);
    return;
  }
  launch_push_bfd(s, 0.5 * delt);
  bool was_sorted = s.sorted;
  if (!launch_push_velocity(s, delt, s.fuse_push != 0))
    launch_push_position(s, delt);
  s.sorted = was_sorted;
  const RhoDone rho = launch_deposit(s, delt, carry_rho(s) ? RhoMode::Keep : RhoMode::Defer);
  launch_push_bfd(s, 0.5 * delt);
  launch_fold_current(s, rho == RhoDone::Deposited ? 0 : 1, 4);
  if (rho == RhoDone::Kept)
    launch_rho_continuity(s, delt);
  launch_push_efd(s, delt);
  // a deferred rho deposit reads the particles in their new order, so the
  // sort's record move is left to the next row gather only when rho is carried
  launch_migrate(s, defers_sort(s) && rho != RhoDone::Deferred);
  if (rho == RhoDone::Deferred) {
    launch_rho(s, true);
    launch_fold_current(s, 0, 1);
  }
  s.rho_valid = true;
}

void drop_graphs(Sim& s)
{
  for (int k = 0; k < 2; k++)
    if (s.gexec[k]) {
      (void)hipGraphExecDestroy(s.gexec[k]);
      s.gexec[k] = nullptr;
    }
}

} // namespace

namespace picnix {
void tune_step(Sim& s, double delt);
}

namespace {

// A capacity overflow of an earlier step, as far as the host can see it without
// waiting: the flags copied after the last step, once that copy has completed.
int step_flags_seen(Sim& s)
{
  if (s.flag_queued && hipEventQuery(s.flag_ev) == hipSuccess && s.h_flag[0])
    return fail(s, PN_ERR_CAPACITY,
                "particle segment capacity exceeded (raise [application.option] buffer_ratio)");
  return PN_OK;
}

// Queue the copy of the device flags that the next pn_step reads.
void queue_step_flags(Sim& s)
{
  HIPCHK(hipMemcpyAsync(s.h_flag, s.ss.errflag, 4 * sizeof(int32_t), hipMemcpyDeviceToHost,
                        s.stream));
  HIPCHK(hipEventRecord(s.flag_ev, s.stream));
  s.flag_queued = true;
}

int step_impl(Sim& s, double delt);

} // namespace

// The kernels keep every access inside the particle segments when a segment
// overflows (the sort clamps its bin starts and counts to the capacity), and
// pn_step reports the overflow without a synchronisation: PN_ERR_CAPACITY comes
// back from the first pn_step whose start sees the flags of the step that
// overflowed, and from pn_synchronize / pn_get_energy / pn_get_particles.
int pn_step(pn_sim* h, double delt)
{
  return guarded(h, [&](Sim& s) {
    int rc = step_flags_seen(s);
    if (rc)
      return rc;
    s.dep_round.clear();
    s.chunk_sorted.clear();
    rc = step_impl(s, delt);
    if (rc == PN_OK)
      queue_step_flags(s);
    return rc;
  }, /*keep_perm=*/true);
}

namespace {

int step_impl(Sim& s, double delt)
{
  {
    if (s.tune.state == 1) {
      tune_step(s, delt); // a measured step of an autotuning session
      return PN_OK;
    }
    // a replayed graph needs the state the capture saw: sorted particles and
    // no phase timing (event records are host-side bookkeeping)
    // (and rho carried from the previous step: the recording uses the continuity update)
    // (the shock problem's boundary draws run on the host between launches)
    // (and, where pn_step defers its sort, the deferred sort of the previous
    // step pending, as the recorded step leaves it)
    if (!s.step_graph || !s.sorted || s.timing || !s.rho_valid || s.g.xwall ||
        (defers_sort(s) && carry_rho(s) && !s.perm_pending)) {
      step_launches(s, delt);
      return PN_OK;
    }
    void* cur = s.fp32 ? (void*)s.qa.rx : (void*)s.pa.x; // buffer parity
    if (!s.buf0_x)
      s.buf0_x = cur;
    const int k = cur == s.buf0_x ? 0 : 1;
    if (!s.gexec[k] || s.graph_dt[k] != delt) {
      if (s.gexec[k]) {
        (void)hipGraphExecDestroy(s.gexec[k]);
        s.gexec[k] = nullptr;
      }
      // record the launches (nothing runs), then run the recording; the host
      // state (buffer swap, flags) has advanced exactly as an eager step would
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
      step_launches(s, delt);
      HIPCHK(hipStreamEndCapture(s.stream, &graph));
      hipError_t e = hipGraphInstantiate(&s.gexec[k], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(e);
      s.graph_dt[k] = delt;
    } else {
      // replay: advance the host-side state as the recorded step does
      if (s.fp32)
        std::swap(s.qa, s.qb);
      else
        std::swap(s.pa, s.pb);
      s.np_host_valid = false;
      s.counted       = false;
      s.sorted        = true;
      s.perm_pending  = defers_sort(s) && carry_rho(s);
    }
    HIPCHK(hipGraphLaunch(s.gexec[k], s.stream));
    return PN_OK;
  }
}

} // namespace

int pn_set_step_graph(pn_sim* h, int enable)
{
  return guarded(h, [&](Sim& s) {
    s.step_graph = enable != 0;
    if (!s.step_graph)
      drop_graphs(s);
    return PN_OK;
  });
}

int pn_get_energy(pn_sim* h, double* efd, double* bfd, double* particle)
{
  return guarded(h, [&](Sim& s) {
    std::vector<double> out(4 + s.g.Ns);
    launch_energy(s, out.data());
    if (efd)
      *efd = out[0];
    if (bfd)
      *bfd = out[1];
    if (particle)
      for (int sp = 0; sp < s.g.Ns; sp++)
        particle[sp] = out[4 + sp];
    return check_device_flags(s);
  });
}

int pn_get_energy_static(pn_sim* h, double* efd, double* bfd, double* particle)
{
  return guarded(h, [&](Sim& s) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev != s.device)
      return fail(s, PN_ERR_STATE, "pn_get_energy_static: the handle's device is not current");
    double out[4 + kMaxSpecies];
    const int rc = launch_energy_static(s, out);
    if (rc == PN_ERR_UNSUPPORTED)
      return fail(s, rc, "pn_get_energy_static: periodic problems only (a physical boundary "
                         "forms the kinetic energy from the moment deposit, which allocates)");
    if (rc)
      return fail(s, rc, "pn_get_energy_static: no scratch address for this device");
    if (efd)
      *efd = out[0];
    if (bfd)
      *bfd = out[1];
    if (particle)
      for (int sp = 0; sp < s.g.Ns; sp++)
        particle[sp] = out[4 + sp];
    return check_device_flags(s);
  });
}

int pn_get_diverror(pn_sim* h, double* efd, double* bfd)
{
  return guarded(h, [&](Sim& s) {
    std::vector<double> out(4 + s.g.Ns);
    launch_energy(s, out.data());
    if (efd)
      *efd = out[2];
    if (bfd)
      *bfd = out[3];
    return PN_OK;
  });
}

int pn_get_boundary_state(pn_sim* h, int chunk, int64_t* out6)
{
  return guarded(h, [&](Sim& s) {
    if (!out6 || chunk < 0 || chunk >= s.g.nchunk)
      return fail(s, PN_ERR_ARG, "invalid chunk or buffer");
    if (!s.g.xwall)
      return fail(s, PN_ERR_STATE, "no physical boundary state: the problem is periodic");
    out6[0] = s.shk.random_seed[chunk];
    out6[1] = s.shk.lastid[chunk];
    out6[2] = s.shk.influx[2 * chunk];
    out6[3] = s.shk.influx[2 * chunk + 1];
    out6[4] = s.shk.efflux[2 * chunk];
    out6[5] = s.shk.efflux[2 * chunk + 1];
    return PN_OK;
  });
}

int pn_synchronize(pn_sim* h)
{
  return guarded(h, [&](Sim& s) {
    HIPCHK(hipStreamSynchronize(s.stream));
    return check_device_flags(s);
  });
}

} // extern "C"

namespace picnix {

static hipEvent_t take_event(Sim& s)
{
  if (!s.ev_pool.empty()) {
    hipEvent_t e = s.ev_pool.back();
    s.ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  return e;
}

PhaseTimer::PhaseTimer(Sim& s_, int ph) : s(s_), phase(ph)
{
  if (!s.timing)
    return;
  a = take_event(s);
  b = take_event(s);
  HIPCHK(hipEventRecord(a, s.stream));
}

PhaseTimer::~PhaseTimer()
{
  if (!a)
    return;
  (void)hipEventRecord(b, s.stream);
  s.ev_pending.push_back({phase, a, b});
}

static void drain_timing(Sim& s)
{
  HIPCHK(hipStreamSynchronize(s.stream));
  for (auto& t : s.ev_pending) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, t.start, t.stop));
    s.ph_ms[t.phase] += ms;
    s.ph_count[t.phase] += 1;
    s.ev_pool.push_back(t.start);
    s.ev_pool.push_back(t.stop);
  }
  s.ev_pending.clear();
}

// ---------------------------------------------------------------------------
// Autotuner (pn_autotune): the launch shapes of the particle kernels are picked
// on the run's own steps.  A session of 2 x max(#candidates, 2) steps runs them
// eagerly with per-phase HIP-event timing; step k tries candidate k mod n of
// every knob (bin rows per march block for the current and rho march kernels;
// 1, 2 or 4 bin rows per gather block for all species or 1 or 2 per species,
// or one thread per particle), each candidate is measured twice
// and keeps its faster time, and the fastest candidate of every knob stays.
// Every candidate computes the same step (the current is summed in another
// order), so the tuning steps are ordinary steps of the run.
// ---------------------------------------------------------------------------
// gather shapes: bin rows per block (0: one thread per particle) and whether a
// block serves one species (1) or all of them (0)
static constexpr int kVelShapes           = 6;
static constexpr int kVelRows[kVelShapes]  = {1, 2, 4, 1, 2, 0};
static constexpr int kVelSplit[kVelShapes] = {0, 0, 0, 1, 1, 0};

static int tune_cycle(const Sim& s)
{
  return std::max<int>((int)s.tune.cand.size(), s.tune.tune_vel ? kVelShapes : 2);
}

// steps of a session: two per candidate of the longest knob, then kQueueSteps
// over the block-per-item and persistent work-queue launches, with and without
// the FP32 gather running the position push, each combination twice
static constexpr int kQueueSteps = 8;
static int tune_steps(const Sim& s) { return 2 * tune_cycle(s) + kQueueSteps; }

// rows > 0: that many bin rows per gather block; 0: one thread per particle;
// -1: the built-in heuristic.  split: one block per species.
static void set_vel_shape(Sim& s, int rows, int split = 0)
{
  s.vel_generic = rows == 0 ? 1 : 0;
  s.vel_rows    = std::max(rows, 0);
  s.vel_split   = split;
}

void tune_begin(Sim& s)
{
  Sim::AutoTune& t = s.tune;
  t                = Sim::AutoTune{};
  const Geom& g    = s.g;
  s.dep_rows = s.rho_rows = 0;
  s.dep_queue = s.vel_queue = s.fuse_push = 0;
  set_vel_shape(s, -1);
  if (g.dim == 3 && !s.generic_kernels) {
    const int nbz = g.nbin / g.bsz;
    int       hz, nzr;
    march_split(g, 0, hz, nzr);
    for (int v : {1, 2, 3, 4, 5, 6, 7, 8, 9, 12, nbz, hz})
      if (v >= 1 && v <= nbz && std::find(t.cand.begin(), t.cand.end(), v) == t.cand.end())
        t.cand.push_back(v);
    std::sort(t.cand.begin(), t.cand.end());
  }
  t.tune_vel = g.dim >= 2 && !s.generic_kernels && s.sorted;
  if (t.cand.empty() && !t.tune_vel) {
    t.state = 2;
    return;
  }
  const double inf = std::numeric_limits<double>::infinity();
  t.t_dep.assign(t.cand.size(), inf);
  t.t_rho.assign(t.cand.size(), inf);
  t.t_vel.assign(kVelShapes, inf);
  t.state                 = 1;
  drop_graphs(s);
}

// end of the shape phase: the fastest shape of every knob stays
static void tune_pick_shapes(Sim& s)
{
  Sim::AutoTune& t    = s.tune;
  // a knob keeps the heuristic unless at least two of its candidates ran (rho is
  // deposited only on steps that cannot carry it by continuity)
  auto           best = [&](const std::vector<double>& v) {
    const long ran = std::count_if(v.begin(), v.end(), [](double x) { return std::isfinite(x); });
    size_t     k   = std::min_element(v.begin(), v.end()) - v.begin();
    return (ran >= 2 && k < v.size()) ? t.cand[k] : 0;
  };
  s.dep_rows    = best(t.t_dep);
  s.rho_rows    = best(t.t_rho);
  if (t.tune_vel) {
    const size_t k = std::min_element(t.t_vel.begin(), t.t_vel.end()) - t.t_vel.begin();
    if (std::isfinite(t.t_vel[k]))
      set_vel_shape(s, kVelRows[k], kVelSplit[k]);
    else
      set_vel_shape(s, -1);
  } else {
    set_vel_shape(s, -1);
  }
}

void tune_step(Sim& s, double delt)
{
  Sim::AutoTune& t     = s.tune;
  const int      n     = tune_cycle(s);
  const int      c     = t.step % n;
  const bool     shape = t.step < 2 * n;
  const int      qv    = (t.step - 2 * n) & 1;        // 0 block per item, 1 persistent
  const int      fv    = ((t.step - 2 * n) >> 1) & 1; // 1: fused gather + position push
  if (shape) {
    if (!t.cand.empty())
      s.dep_rows = s.rho_rows = t.cand[c % t.cand.size()];
    if (t.tune_vel)
      set_vel_shape(s, kVelRows[c % kVelShapes], kVelSplit[c % kVelShapes]);
  } else {
    s.dep_queue = s.vel_queue = qv;
    s.fuse_push               = fv;
  }
  // time this step's phases alone; the caller's accumulators are set aside
  drain_timing(s);
  double saved_ms[PN_NPHASE + 1];
  int    saved_n[PN_NPHASE + 1];
  for (int i = 0; i <= PN_NPHASE; i++) {
    saved_ms[i]   = s.ph_ms[i];
    saved_n[i]    = s.ph_count[i];
    s.ph_ms[i]    = 0;
    s.ph_count[i] = 0;
  }
  const bool saved_timing = s.timing;
  s.timing                = true;
  step_launches(s, delt);
  drain_timing(s);
  const double dep = s.ph_count[3] ? s.ph_ms[3] : -1.0;
  const double rho = s.ph_count[kPhaseRhoDeposit] ? s.ph_ms[kPhaseRhoDeposit] : -1.0;
  // gather + position push: the row gather may run the position push too
  const double vel = s.ph_count[1] ? s.ph_ms[1] + s.ph_ms[2] : -1.0;
  for (int i = 0; i <= PN_NPHASE; i++) {
    s.ph_ms[i]    = saved_ms[i];
    s.ph_count[i] = saved_n[i];
  }
  s.timing = saved_timing;
  if (!shape) {
    if (dep >= 0)
      t.t_q[0][qv] = std::min(t.t_q[0][qv], dep);
    if (vel >= 0)
      t.t_v[2 * fv + qv] = std::min(t.t_v[2 * fv + qv], vel);
  } else {
    if (!t.cand.empty()) {
      const size_t k = c % t.cand.size();
      if (dep >= 0)
        t.t_dep[k] = std::min(t.t_dep[k], dep);
      if (rho >= 0)
        t.t_rho[k] = std::min(t.t_rho[k], rho);
    }
    if (t.tune_vel && vel >= 0)
      t.t_vel[c % kVelShapes] = std::min(t.t_vel[c % kVelShapes], vel);
  }
  t.step++;
  if (t.step == 2 * n)
    tune_pick_shapes(s);
  if (t.step >= tune_steps(s)) {
    s.dep_queue  = t.t_q[0][1] < t.t_q[0][0] ? 1 : 0;
    const int kv = (int)(std::min_element(t.t_v, t.t_v + 4) - t.t_v);
    s.vel_queue  = kv & 1;
    s.fuse_push  = kv >> 1;
    t.state     = 2;
    drop_graphs(s);
  }
}

} // namespace picnix

extern "C" {

int pn_autotune(pn_sim* h, int enable)
{
  return guarded(h, [&](Sim& s) {
    if (enable) {
      tune_begin(s);
    } else {
      s.tune     = Sim::AutoTune{};
      s.dep_rows = s.rho_rows = 0;
      s.dep_queue = s.vel_queue = s.fuse_push = 0;
      set_vel_shape(s, -1);
      drop_graphs(s);
    }
    return PN_OK;
  });
}

int pn_get_tuning(pn_sim* h, int* out9)
{
  return guarded(h, [&](Sim& s) {
    if (!out9)
      return PN_ERR_ARG;
    out9[0] = s.tune.state;
    out9[1] = s.tune.state == 1 ? tune_steps(s) - s.tune.step : 0;
    out9[2] = s.dep_rows;
    out9[3] = s.rho_rows;
    out9[4] = s.vel_generic ? -1 : s.vel_rows;
    out9[5] = s.vel_split;
    out9[6] = s.dep_queue;
    out9[7] = s.vel_queue;
    out9[8] = s.fuse_push;
    return PN_OK;
  });
}

int pn_timing(pn_sim* h, int enable)
{
  return guarded(h, [&](Sim& s) {
    drain_timing(s);
    for (int i = 0; i <= PN_NPHASE; i++) {
      s.ph_ms[i]    = 0;
      s.ph_count[i] = 0;
    }
    s.timing = enable != 0;
    return PN_OK;
  });
}

int pn_timing_report(pn_sim* h, double* ms_total, int* launches)
{
  return guarded(h, [&](Sim& s) {
    drain_timing(s);
    for (int i = 0; i < PN_NPHASE; i++) {
      if (ms_total)
        ms_total[i] = s.ph_ms[i];
      if (launches)
        launches[i] = s.ph_count[i];
    }
    return PN_OK;
  });
}

} // extern "C"
