// pn_shock.cpp -- host side of the shock problem (example/shock/main.cpp): the
// set-up constants, the field profile, the particle loader and the random draws
// of the upstream boundary (MainChunk::set_boundary_particle, inject_particle).
//
// The draws use the generators the example uses -- std::mt19937_64 seeded from
// option["random_seed"], libstdc++'s uniform_real_distribution,
// gamma_distribution and poisson_distribution (aliased at nix/random.hpp:9-12)
// and the Maxwell-Juttner sampler below -- consumed in the example's order, so a
// run reproduces the reference's random stream.  They are sequential per chunk
// and touch only the chunks at x = Lx; the device does everything else.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "../../include/picnix.h"
#include "pn_sim.hpp"

namespace picnix {

namespace {

// nix::MaxwellJuttner (nix/random.hpp:27-163; S. Zenitani, Phys. Plasmas 22,
// 042116 (2015)): the Lorentz factor by the modified Canfield method from four
// gamma distributions, an isotropic direction, then the flipping-method boost
// along x.  Same operations in the same order as the reference, so that the
// libstdc++ draws map to the same velocities.
class MaxwellJuttner {
public:
  MaxwellJuttner(double temperature, double drift) : temperature_(temperature), drift_(drift)
  {
    const double T = temperature_;
    table_ = {0, std::sqrt(M_PI), kA * std::sqrt(2 * T), kB * 1.5 * std::sqrt(M_PI) * T,
              std::sqrt(2 * T) * (2 * T)};
    for (size_t i = 1; i < table_.size(); i++)
      table_[i] = table_[i - 1] + table_[i];
    const double total = table_.back();
    for (double& v : table_)
      v = v / total;
    gamma_ = {std::gamma_distribution<double>(1.5, T), std::gamma_distribution<double>(2.0, T),
              std::gamma_distribution<double>(2.5, T), std::gamma_distribution<double>(3.0, T)};
  }

  void set_drift(double drift) { drift_ = drift; }

  void operator()(std::mt19937_64& r, double& ux, double& uy, double& uz)
  {
    double xx     = -1;
    bool   reject = false;
    do {
      double r1   = uniform_(r);
      double r2   = uniform_(r);
      int    k    = (int)(std::upper_bound(table_.begin(), table_.end(), r1) - table_.begin()) - 1;
      xx          = gamma_[k](r);
      double R1   = (xx + 1) * std::sqrt(xx + 2);
      double R2   = std::sqrt(2) + kA * std::sqrt(xx) + kB * std::sqrt(2) * xx + std::pow(xx, 1.5);
      reject      = (r2 > R1 / R2);
    } while (reject);
    double uu = std::sqrt(xx * (xx + 2));
    double r3 = uniform_(r);
    double r4 = uniform_(r);
    ux        = uu * (2 * r3 - 1);
    uy        = uu * 2 * std::sqrt(r3 * (1 - r3)) * std::cos(2 * M_PI * r4);
    uz        = uu * 2 * std::sqrt(r3 * (1 - r3)) * std::sin(2 * M_PI * r4);
    // boost: flip the sign of ux with probability -V vx, then transform
    double rr = uniform_(r);
    double Gm = std::sqrt(1 + drift_ * drift_);
    double Vx = drift_ / Gm;
    double gm = std::sqrt(1 + ux * ux + uy * uy + uz * uz);
    double vx = ux / gm;
    ux        = (-Vx * vx > rr) ? -ux : +ux;
    ux        = Gm * (ux + Vx * gm);
  }

private:
  static constexpr double                      kA = 0.56;
  static constexpr double                      kB = 0.35;
  double                                       temperature_, drift_;
  std::uniform_real_distribution<double>       uniform_;
  std::vector<std::gamma_distribution<double>> gamma_;
  std::vector<double>                          table_;
};

// One chunk's generator state (MainChunk::mt and its stateless `uniform`).
struct ChunkRng {
  std::mt19937_64                        mt;
  std::uniform_real_distribution<double> uniform;
  // MainChunk::reset_random_number (example/shock/main.cpp:19-24)
  void reset(int& seed)
  {
    seed = seed + 1;
    mt.seed(seed);
  }
};

// MainChunk::generate_injection_particle (example/shock/main.cpp:26-48) with c = 1:
// a velocity and a start in [xmax, xmax + c dt), kept once the particle has moved
// inside within one step.
void injection_particle(ChunkRng& g, MaxwellJuttner& mj, double delt, double xmax, double& x,
                        double& ux, double& uy, double& uz)
{
  const double cc = 1.0;
  const double rc = 1.0 / cc;
  mj(g.mt, ux, uy, uz);
  double xx = xmax + cc * delt * g.uniform(g.mt);
  double dx = ux * delt / std::sqrt(1 + (ux * ux + uy * uy + uz * uz) * rc * rc);
  while (xx + dx >= xmax) {
    mj(g.mt, ux, uy, uz);
    xx = xmax + cc * delt * g.uniform(g.mt);
    dx = ux * delt / std::sqrt(1 + (ux * ux + uy * uy + uz * uz) * rc * rc);
  }
  x = xx + dx;
}

// example/shock/main.cpp:50-54
double flow_profile(const Geom& g, double x, double L)
{
  double xmin = g.glim[2][0] + g.Nb * g.del[2];
  return -(x - xmin < L ? std::sin(0.5 * M_PI * (x - xmin) / L) : 1.0);
}

// density extension: n(x) = 1 below ramp_x, a linear ramp over ramp_w cells,
// `density` above; particles per cell = floor(nppc n(x) + 1/2)
double shock_density(const ShockConfig& c, double x, double delh)
{
  if (c.density == 1.0)
    return 1.0;
  double xc = x / delh;
  if (xc < c.ramp_x)
    return 1.0;
  if (xc < c.ramp_x + c.ramp_w)
    return 1.0 + (c.density - 1.0) * (xc - c.ramp_x) / c.ramp_w;
  return c.density;
}

int shock_npc(const ShockConfig& c, double x, double delh)
{
  return (int)std::floor(c.nppc * shock_density(c, x, delh) + 0.5);
}

// f(k, ids[k]) for every k on up to 16 host threads; threads = false runs them in
// order on the calling thread (small amounts of work, where starting the threads
// costs more than the work while the GPU waits)
template <typename F>
void parallel_chunks(const std::vector<int>& ids, F f, bool threads = true)
{
  const unsigned n    = (unsigned)ids.size();
  unsigned       nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nthr                = std::min(nthr, std::max(1u, n));
  if (!threads) {
    for (unsigned k = 0; k < n; k++)
      f(k, ids[k]);
    return;
  }
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthr; t++)
    pool.emplace_back([&, t] {
      for (unsigned k = t; k < n; k += nthr)
        f(k, ids[k]);
    });
  for (auto& th : pool)
    th.join();
}

void ensure_vals(Sim& s, size_t elems)
{
  ShockState& h = s.shk;
  if (elems <= h.vals_cap)
    return;
  if (h.d_vals)
    HIPCHK(hipFree(h.d_vals));
  size_t cap = std::max<size_t>(elems, 4096);
  HIPCHK(hipMalloc((void**)&h.d_vals, cap * sizeof(double)));
  h.vals_cap = cap;
}

// upper chunks in chunk-id order (one entry per species segment in shk.useg)
std::vector<int> upper_chunks(const Sim& s)
{
  std::vector<int> out;
  for (size_t k = 0; k < s.shk.useg.size(); k += s.g.Ns)
    out.push_back(s.shk.useg[k] / s.g.Ns);
  return out;
}

// The next pinned staging buffer, at least `bytes` long, once its previous copy
// has finished.
char* stage_buffer(Sim& s, size_t bytes, int& slot)
{
  ShockState& h = s.shk;
  slot          = h.stage_next;
  h.stage_next  = (slot + 1) % ShockState::kStages;
  if (h.stage_ev[slot])
    HIPCHK(hipEventSynchronize(h.stage_ev[slot]));
  else
    HIPCHK(hipEventCreateWithFlags(&h.stage_ev[slot], hipEventDisableTiming));
  if (bytes > h.stage_cap[slot]) {
    if (h.h_stage[slot])
      HIPCHK(hipHostFree(h.h_stage[slot]));
    h.h_stage[slot]   = nullptr;
    h.stage_cap[slot] = 0;
    const size_t cap  = std::max<size_t>(bytes + bytes / 2, 1 << 16);
    HIPCHK(hipHostMalloc(&h.h_stage[slot], cap, hipHostMallocDefault));
    h.stage_cap[slot] = cap;
  }
  return static_cast<char*>(h.h_stage[slot]);
}

// Values per upper segment (and, for an injection, the counts) to the device,
// through a pinned staging buffer: the copies run in stream order with no wait on
// the host.
void upload_vals(Sim& s, const std::vector<std::vector<double>>& vals,
                 const std::vector<int32_t>* counts)
{
  ShockState&  h  = s.shk;
  const size_t nk = h.useg.size();
  size_t       tot = 0;
  for (size_t k = 0; k < nk; k++)
    tot += vals[k].size();
  ensure_vals(s, tot);
  // staging layout: offsets (int64) [nk], counts (int32) [nk, padded], values
  const size_t boff = 0, bcnt = nk * sizeof(int64_t);
  const size_t bval = bcnt + ((nk * sizeof(int32_t) + 7) & ~(size_t)7);
  int          slot;
  char*        st   = stage_buffer(s, bval + tot * sizeof(double), slot);
  int64_t*     off  = reinterpret_cast<int64_t*>(st + boff);
  double*      flat = reinterpret_cast<double*>(st + bval);
  size_t       at   = 0;
  for (size_t k = 0; k < nk; k++) {
    off[k] = (int64_t)at;
    std::copy(vals[k].begin(), vals[k].end(), flat + at);
    at += vals[k].size();
  }
  if (tot)
    HIPCHK(hipMemcpyAsync(h.d_vals, flat, tot * sizeof(double), hipMemcpyHostToDevice,
                          s.stream));
  HIPCHK(hipMemcpyAsync(h.d_voff, off, nk * sizeof(int64_t), hipMemcpyHostToDevice, s.stream));
  if (counts) {
    int32_t* cnt = reinterpret_cast<int32_t*>(st + bcnt);
    std::copy(counts->begin(), counts->end(), cnt);
    HIPCHK(hipMemcpyAsync(h.d_count, cnt, nk * sizeof(int32_t), hipMemcpyHostToDevice,
                          s.stream));
  }
  HIPCHK(hipEventRecord(h.stage_ev[slot], s.stream));
}

} // namespace

// MainChunk::setup constants (example/shock/main.cpp:56-92); cc = 1, Ns = 2.
void shock_constants(Sim& s)
{
  const ShockConfig& c  = s.cfg.shock;
  ShockState&        h  = s.shk;
  Geom&              g  = s.g;
  const double       cc = 1.0;
  const double       u0 = c.u0;
  double gamma = std::sqrt(1.0 + (u0 * u0) / (cc * cc));
  double me    = 1.0 / c.nppc;
  double qe    = -c.wp / c.nppc * std::sqrt(gamma);
  double mi    = me * c.mime;
  double qi    = -qe;
  double b0    = cc * std::sqrt(c.sigma) / std::abs(qe / me);
  double vae   = cc * std::sqrt(c.sigma);
  double vai   = cc * std::sqrt(c.sigma / c.mime);
  h.delt       = s.cfg.delt;
  h.u0         = u0;
  h.taper      = c.taper;
  h.vte        = vae * std::sqrt(0.5 * c.betae);
  h.vti        = vai * std::sqrt(0.5 * c.betai);
  h.B0[0]      = b0 * std::cos(c.theta / 180 * M_PI);
  h.B0[1]      = b0 * std::sin(c.theta / 180 * M_PI) * std::cos(c.phi / 180 * M_PI);
  h.B0[2]      = b0 * std::sin(c.theta / 180 * M_PI) * std::sin(c.phi / 180 * M_PI);
  h.E0[0]      = 0;
  h.E0[1]      = +(-u0) * h.B0[2] / (gamma * cc);
  h.E0[2]      = -(-u0) * h.B0[1] / (gamma * cc);
  g.m[0]       = me;
  g.q[0]       = qe;
  g.m[1]       = mi;
  g.q[1]       = qi;
  h.npc_max    = c.density == 1.0 ? c.nppc
                                  : (int)std::floor(c.nppc * std::max(1.0, c.density) + 0.5);
  // MainChunk::inject_particle (example/shock/main.cpp:453-461): particles per
  // (z, y) column and step
  const double rc  = 1.0 / cc;
  const double npu = c.density == 1.0 ? (double)c.nppc
                                      : (double)shock_npc(c, g.glim[2][1], g.del[2]);
  h.flux = npu * std::abs(u0) / std::sqrt(1 + u0 * u0 * rc * rc) * h.delt / g.del[2];
}

void shock_setup(Sim& s)
{
  ShockState& h  = s.shk;
  const Geom& g  = s.g;
  const int   nc = g.nchunk;
  h.random_seed.assign(nc, 0);
  h.lastid.assign(nc, 0);
  h.influx.assign(2 * (size_t)nc, 0);
  h.efflux.assign(2 * (size_t)nc, 0);
  std::random_device rd;
  for (int c = 0; c < nc; c++) // pic/pic_chunk.cpp:243-254
    h.random_seed[c] = s.cfg.seed_type == "fixed" ? c + s.cfg.seed_offset : (int)rd();
  h.useg.clear();
  for (int c = 0; c < nc; c++)
    if (s.coord[c * 3 + 2] == g.C[2] - 1)
      for (int sp = 0; sp < g.Ns; sp++)
        h.useg.push_back(c * g.Ns + sp);
  const size_t nk = h.useg.size();
  HIPCHK(hipMalloc((void**)&h.d_useg, std::max<size_t>(nk, 1) * sizeof(int32_t)));
  HIPCHK(hipMalloc((void**)&h.d_list, std::max<size_t>(nk * g.cap, 1) * sizeof(int32_t)));
  HIPCHK(hipMalloc((void**)&h.d_count, std::max<size_t>(nk, 1) * sizeof(int32_t)));
  HIPCHK(hipMalloc((void**)&h.d_voff, std::max<size_t>(nk, 1) * sizeof(int64_t)));
  HIPCHK(hipMemcpy(h.d_useg, h.useg.data(), nk * sizeof(int32_t), hipMemcpyHostToDevice));
  ensure_vals(s, 4096);
}

void shock_release(Sim& s)
{
  ShockState& h = s.shk;
  h.ahead.reset(); // waits for draws still running on the host thread
  if (s.stream)
    (void)hipStreamSynchronize(s.stream);
  for (void* p : {(void*)h.d_useg, (void*)h.d_list, (void*)h.d_count, (void*)h.d_voff,
                  (void*)h.d_vals})
    if (p)
      (void)hipFree(p);
  for (int k = 0; k < ShockState::kStages; k++) {
    if (h.h_stage[k])
      (void)hipHostFree(h.h_stage[k]);
    if (h.stage_ev[k])
      (void)hipEventDestroy(h.stage_ev[k]);
  }
  h = ShockState{};
}

// MainChunk::setup field profile (example/shock/main.cpp:94-137), uniform in y and
// z.  Each x column takes the value its owner chunk computes, from the chunk's
// array index: xi = (ix + 0.5) dx + xmin (the wall chunks own the x padding).
// The initial EMF exchange and set_boundary_emf follow.
void shock_init_fields(Sim& s)
{
  const Geom&       g  = s.g;
  const ShockState& h  = s.shk;
  const double      cc = 1.0;
  const int         nx = g.dims[2];
  std::vector<double> col(6 * (size_t)g.P[2]);
  for (int X = 0; X < g.P[2]; X++) {
    int cx = std::min(std::max((X - g.pad[2]) / nx, 0), g.C[2] - 1);
    if (X < g.pad[2])
      cx = 0;
    // any chunk of that x slab: all share lim[x]
    const int    id   = s.id_of_coord[(size_t)cx];
    const double xmin = s.lim[id * 6 + 4];
    const int    ix   = X - cx * nx;
    double       xi   = (ix + 0.5) * g.del[2] + xmin;
    double       Ux   = h.u0 * flow_profile(g, xi, h.taper);
    double       gm   = std::sqrt(1 + Ux * Ux / (cc * cc));
    col[6 * X + 0]    = h.E0[0];
    col[6 * X + 1]    = +Ux * h.B0[2] / (gm * cc);
    col[6 * X + 2]    = -Ux * h.B0[1] / (gm * cc);
    col[6 * X + 3]    = h.B0[0];
    col[6 * X + 4]    = h.B0[1];
    col[6 * X + 5]    = h.B0[2];
  }
  std::vector<double> he(3 * (size_t)g.vol), hb(3 * (size_t)g.vol);
  for (long r = 0; r < (long)g.P[0] * g.P[1]; r++)
    for (int X = 0; X < g.P[2]; X++)
      for (int k = 0; k < 3; k++) {
        he[k * g.vol + r * g.P[2] + X] = col[6 * X + k];
        hb[k * g.vol + r * g.P[2] + X] = col[6 * X + 3 + k];
      }
  HIPCHK(hipMemcpy(s.d_E, he.data(), he.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s.d_B, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice));
  launch_fill_ghost(s, s.d_E, 3);
  launch_fill_ghost(s, s.d_B, 3);
  if (s.fp32)
    launch32_fields_from_f64(s, PN_FIELD_UF); // the wall reset below acts on the FP32 fields
  launch_wall_emf(s);
}

// MainChunk::setup particle loader (example/shock/main.cpp:139-210): nppc
// particles per cell at uniform positions, electron and ion at the same place,
// velocities from a Maxwell-Juttner distribution drifting with the flow profile;
// ids (ciy + ciz cy) maxid + cix mp + ip.
void shock_load_particles(Sim& s)
{
  const Geom&        g  = s.g;
  const ShockConfig& c  = s.cfg.shock;
  ShockState&        h  = s.shk;
  const int          nc = g.nchunk;
  const size_t       tot = (size_t)g.nseg * g.cap;
  std::vector<double>  hx(tot), hy(tot), hz(tot), hux(tot), huy(tot), huz(tot);
  std::vector<int64_t> hid(tot);
  std::vector<int32_t> hnp(g.nseg, 0);
  std::vector<int>     all(nc);
  for (int k = 0; k < nc; k++)
    all[k] = k;
  bool overflow = false;
  parallel_chunks(all, [&](unsigned, int id) {
    ChunkRng rng;
    rng.mt.seed(h.random_seed[id]);
    MaxwellJuttner mj[2] = {MaxwellJuttner(h.vte * h.vte, 0.0), MaxwellJuttner(h.vti * h.vti, 0.0)};
    const int*    off = &s.offset[id * 3];
    const double* lim = &s.lim[id * 6];
    int           mp  = h.npc_max * g.dims[0] * g.dims[1] * g.dims[2];
    int           cz  = g.N[0] / g.dims[0];
    int           cy  = g.N[1] / g.dims[1];
    int64_t maxid     = std::numeric_limits<int64_t>::max() / (cy * cz);
    int64_t ciz       = off[0] / g.dims[0];
    int64_t ciy       = off[1] / g.dims[1];
    int64_t cix       = off[2] / g.dims[2];
    int64_t lastid    = (ciy + ciz * cy) * maxid + cix * mp;
    int     ip        = 0;
    for (int iz = 0; iz < g.dims[0]; iz++)
      for (int iy = 0; iy < g.dims[1]; iy++)
        for (int ix = 0; ix < g.dims[2]; ix++) {
          int npc = c.nppc;
          if (c.density != 1.0)
            npc = shock_npc(c, lim[4] + (ix + 0.5) * g.del[2], g.del[2]);
          for (int jp = 0; jp < npc; ip++, jp++) {
            double x = lim[4] + (ix + rng.uniform(rng.mt)) * g.del[2];
            double y = lim[2] + (iy + rng.uniform(rng.mt)) * g.del[1];
            double z = lim[0] + (iz + rng.uniform(rng.mt)) * g.del[0];
            double U = h.u0 * flow_profile(g, x, h.taper);
            for (int sp = 0; sp < 2; sp++) {
              mj[sp].set_drift(U);
              double ux, uy, uz;
              mj[sp](rng.mt, ux, uy, uz);
              if (ip >= g.cap) {
                overflow = true;
                continue;
              }
              size_t i = ((size_t)id * g.Ns + sp) * g.cap + ip;
              hx[i]    = x;
              hy[i]    = y;
              hz[i]    = z;
              hux[i]   = ux;
              huy[i]   = uy;
              huz[i]   = uz;
              hid[i]   = lastid + ip;
            }
          }
        }
    for (int sp = 0; sp < 2; sp++)
      hnp[id * g.Ns + sp] = std::min(ip, (int)g.cap);
    h.lastid[id] = lastid + mp;
  });
  if (overflow)
    throw std::runtime_error("shock loader: particles per chunk exceed the segment capacity");
  auto up = [&](void* d, const void* p, size_t bytes) {
    HIPCHK(hipMemcpy(d, p, bytes, hipMemcpyHostToDevice));
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

// MainChunk::set_boundary_particle at x = Lx (example/shock/main.cpp:418-442):
// after the position push has reflected the particles at the wall and marked
// those past x = Lx, every upper chunk redraws the marked particles' x and u
// (their y, z stay) in particle order, species by species.
void shock_position_boundary(Sim& s)
{
  ShockState& h  = s.shk;
  const Geom& g  = s.g;
  const size_t nk = h.useg.size();
  if (nk == 0)
    return;
  launch_wall_list(s);
  std::vector<int32_t> cnt(nk);
  HIPCHK(hipMemcpyAsync(cnt.data(), h.d_count, nk * sizeof(int32_t), hipMemcpyDeviceToHost,
                        s.stream));
  HIPCHK(hipStreamSynchronize(s.stream));
  std::vector<std::vector<double>> vals(nk);
  const std::vector<int>           up    = upper_chunks(s);
  long                             total = 0;
  for (int32_t n : cnt)
    total += n;
  parallel_chunks(up, [&](unsigned u, int id) {
    ChunkRng rng;
    rng.reset(h.random_seed[id]);
    MaxwellJuttner mj[2] = {MaxwellJuttner(h.vte * h.vte, -h.u0),
                            MaxwellJuttner(h.vti * h.vti, -h.u0)};
    const double   xmax  = s.lim[id * 6 + 5];
    for (int sp = 0; sp < g.Ns; sp++) {
      const size_t k = (size_t)u * g.Ns + sp;
      vals[k].resize(4 * (size_t)cnt[k]);
      for (int t = 0; t < cnt[k]; t++) {
        double x, ux, uy, uz;
        injection_particle(rng, mj[sp], h.delt, xmax, x, ux, uy, uz);
        vals[k][4 * t + 0] = x;
        vals[k][4 * t + 1] = ux;
        vals[k][4 * t + 2] = uy;
        vals[k][4 * t + 3] = uz;
      }
      h.efflux[2 * id + sp] = cnt[k];
    }
  }, total >= 16384); // the redraws of a step are usually few
  upload_vals(s, vals, nullptr);
  if (s.fp32)
    launch32_wall_apply(s);
  else
    launch_wall_apply(s);
}

namespace {

// What one injection call needs: the upper chunks, their random_seed and lastid
// when the call starts, and the constants of the draws.
struct InjectInput {
  std::vector<int>     up;     // upper chunk ids, chunk-id order
  std::vector<int>     seed;   // random_seed[id] before the call's reset
  std::vector<int64_t> lastid; // lastid[id] before the call
  std::vector<double>  lim;    // [u][6] chunk ranges
  int    nz = 0, ny = 0, ns = 0;
  double dy = 0, dz = 0, flux = 0, vte = 0, vti = 0, u0 = 0, delt = 0;
};

// Its results: the values per upper segment and the particles per chunk.
struct InjectDraws {
  std::vector<std::vector<double>> vals;  // [u][sp] x, y, z, ux, uy, uz, id bits
  std::vector<int>                 total; // [u]
};

// MainChunk::inject_particle (example/shock/main.cpp:445-535): a Poisson number of
// particles per (z, y) column of every upper chunk, the same number for both
// species, appended column by column with ids from lastid.  The chunk's
// generator is reseeded at the start of the call, so the draws depend on the
// seed and lastid only, not on the particles.
InjectDraws draw_injection(const InjectInput& in, bool threads)
{
  const size_t nu = in.up.size();
  InjectDraws  d;
  d.vals.resize(nu * in.ns);
  d.total.assign(nu, 0);
  parallel_chunks(in.up, [&](unsigned u, int) {
    ChunkRng rng;
    int      seed = in.seed[u];
    rng.reset(seed);
    std::poisson_distribution<int> poisson(in.flux);
    MaxwellJuttner mj[2] = {MaxwellJuttner(in.vte * in.vte, -in.u0),
                            MaxwellJuttner(in.vti * in.vti, -in.u0)};
    const double*  lim   = &in.lim[6 * u];
    const int      ncol  = in.nz * in.ny;
    std::vector<int> col(ncol);
    int              total = 0;
    for (int k = 0; k < ncol; k++) {
      col[k] = poisson(rng.mt);
      total += col[k];
    }
    std::vector<int64_t> id64(in.ns, in.lastid[u]);
    for (int sp = 0; sp < in.ns; sp++)
      d.vals[(size_t)u * in.ns + sp].reserve(7 * (size_t)total);
    for (int iz = 0; iz < in.nz; iz++)
      for (int iy = 0; iy < in.ny; iy++)
        for (int sp = 0; sp < in.ns; sp++) {
          std::vector<double>& v = d.vals[(size_t)u * in.ns + sp];
          for (int t = 0; t < col[iz * in.ny + iy]; t++) {
            double x, ux, uy, uz;
            injection_particle(rng, mj[sp], in.delt, lim[5], x, ux, uy, uz);
            double y = lim[2] + (iy + rng.uniform(rng.mt)) * in.dy;
            double z = lim[0] + (iz + rng.uniform(rng.mt)) * in.dz;
            double idbits;
            std::memcpy(&idbits, &id64[sp], sizeof(double));
            id64[sp]++;
            v.insert(v.end(), {x, y, z, ux, uy, uz, idbits});
          }
        }
    d.total[u] = total;
  }, threads);
  return d;
}

} // namespace

// The next injection's draws, started on a host thread when the previous
// injection is done; `in` holds the seed and lastid they were drawn for.
struct ShockAhead {
  InjectInput              in;
  std::future<InjectDraws> draws;
};

namespace {

InjectInput inject_input(const Sim& s, const std::vector<int>& up)
{
  const ShockState& h = s.shk;
  const Geom&       g = s.g;
  InjectInput       in;
  in.up = up;
  for (int id : up) {
    in.seed.push_back(h.random_seed[id]);
    in.lastid.push_back(h.lastid[id]);
    in.lim.insert(in.lim.end(), &s.lim[id * 6], &s.lim[id * 6] + 6);
  }
  in.nz   = g.dims[0];
  in.ny   = g.dims[1];
  in.ns   = g.Ns;
  in.dy   = g.del[1];
  in.dz   = g.del[0];
  in.flux = h.flux;
  in.vte  = h.vte;
  in.vti  = h.vti;
  in.u0   = h.u0;
  in.delt = h.delt;
  return in;
}

} // namespace

// Injection at x = Lx.  The draws of this call were usually made ahead, on a host
// thread, while the device ran the step; they are used when they were drawn for
// the seed and lastid the chunks hold now (otherwise -- the first call, a restart,
// another operator order -- they are drawn here).  Then the draws of the next
// call start: every step resets each upper chunk's generator once in the
// position boundary (seed + 1) before the injection resets it again.
void shock_inject(Sim& s)
{
  ShockState&  h  = s.shk;
  const Geom&  g  = s.g;
  const size_t nk = h.useg.size();
  if (nk == 0)
    return;
  const std::vector<int> up = upper_chunks(s);
  InjectInput            in = inject_input(s, up);
  InjectDraws            d;
  bool                   have = false;
  if (h.ahead) {
    InjectDraws r = h.ahead->draws.get();
    if (h.ahead->in.seed == in.seed && h.ahead->in.lastid == in.lastid) {
      d    = std::move(r);
      have = true;
    }
    h.ahead.reset();
  }
  if (!have)
    d = draw_injection(in, true);
  std::vector<int32_t> ninj(nk, 0);
  for (size_t u = 0; u < up.size(); u++) {
    const int id = up[u];
    h.random_seed[id] += 1; // the call's reset
    h.lastid[id] += d.total[u];
    for (int sp = 0; sp < g.Ns; sp++) {
      ninj[u * g.Ns + sp]   = d.total[u];
      h.influx[2 * id + sp] = d.total[u];
    }
  }
  upload_vals(s, d.vals, &ninj);
  h.last_ninj = ninj.empty() ? 0 : *std::max_element(ninj.begin(), ninj.end());
  if (s.fp32)
    launch32_inject_append(s);
  else
    launch_inject_append(s);
  s.np_host_valid = false;
  // the next call's draws
  InjectInput next = inject_input(s, up);
  for (int& sd : next.seed)
    sd += 1; // the next position boundary's reset
  auto a   = std::make_shared<ShockAhead>();
  a->in    = next;
  a->draws = std::async(std::launch::async, [next] { return draw_injection(next, true); });
  h.ahead  = std::move(a);
}

} // namespace picnix
