// pn_sim.hpp -- host-side state of one simulation handle.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "pn_config.hpp"
#include "pn_device.hpp"

struct pn_sim; // the C-ABI handle (include/picnix.h)

namespace picnix {

// A failed HIP runtime call inside a C-ABI entry point (reported as PN_ERR_HIP).
struct HipError {
  std::string msg;
};

inline void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}

#define HIPCHK(call) ::picnix::hip_check((call), #call)

// Per-step scratch of the counting sort and particle migration.
// spare entries after the member list: the rank kernels read up to 15 keys past
// the end of a bin in 16-key groups
constexpr int kMemberPad = 16;

struct SortScratch {
  int32_t* cellcnt   = nullptr; // [nseg][nbin]   particles per bin (atomic slots)
  int32_t* cellstart = nullptr; // [nseg][nbin+1] first particle of every bin
  int32_t* key       = nullptr; // [nseg*cap]     bin of the particle in its destination
  int32_t* slot      = nullptr; // [nseg*cap]     arbitrary slot inside that bin
  int32_t* dseg      = nullptr; // [nseg*cap]     destination segment (-1 = dropped)
  int32_t* didx      = nullptr; // [nseg*cap]     index in the destination's pre-sort array
  int32_t* mkey      = nullptr; // [nseg*cap+pad] member list of every bin (pre-sort indices)
  int32_t* msrc      = nullptr; // [nseg*cap]     where each member's record currently is
  int32_t* perm      = nullptr; // [nseg*cap]     deferred sort: source record of each sorted slot
  int32_t* tilecnt   = nullptr; // [nseg][ntile][27] leavers per 256-particle tile and direction
  int32_t* nout      = nullptr; // [nseg][27]     leavers per direction
  int32_t* recvoff   = nullptr; // [nseg][27]     first pre-sort index of each receive slot
  int32_t* npnew     = nullptr; // [nseg]         particle count after the sort
  int32_t* errflag   = nullptr; // [4]            device-side error flags
  int      ntile     = 0;
};

// example/shock (Config::problem == 1): the constants of MainChunk::setup
// (example/shock/main.cpp:56-92), the per-chunk state the example keeps in
// option["random_seed"] and option["boundary"], and the device buffers of the
// upstream boundary.
struct ShockAhead; // the next injection's draws, computed ahead (pn_shock.cpp)

struct ShockState {
  double delt = 0, u0 = 0, vte = 0, vti = 0, taper = 0;
  double E0[3] = {0, 0, 0}, B0[3] = {0, 0, 0};
  double flux = 0;                 // Poisson mean of injected particles per (z, y) column
  int    npc_max = 0;              // particles per cell at the highest density
  int    last_ninj = 0;            // most particles the last injection appended to a segment
  std::vector<int>     random_seed; // [chunk]
  std::vector<int64_t> lastid;      // [chunk]
  std::vector<int>     influx;      // [chunk][2]
  std::vector<int>     efflux;      // [chunk][2]
  std::vector<int>     useg;        // segments of the chunks at x = Lx, chunk-id order
  int32_t* d_useg  = nullptr;       // [nk]
  int32_t* d_list  = nullptr;       // [nk][cap] particles past x = Lx, in particle order
  int32_t* d_count = nullptr;       // [nk] list lengths, then injected counts
  int64_t* d_voff  = nullptr;       // [nk] offsets into d_vals
  double*  d_vals  = nullptr;       // replacement (4 per particle) / injected (7) values
  size_t   vals_cap = 0;
  // pinned host staging of the uploads, used in turn; each buffer's event marks
  // the end of its last copy
  static constexpr int kStages = 3;
  void*      h_stage[kStages]   = {};
  size_t     stage_cap[kStages] = {};
  hipEvent_t stage_ev[kStages]  = {};
  int        stage_next         = 0;
  std::shared_ptr<ShockAhead> ahead; // draws of the next shock_inject, or null
};

struct Sim;
// the state behind a C-ABI handle (for host-side modules of the library)
Sim* sim_state(::pn_sim* h);

struct Sim {
  Config      cfg;
  std::string toml; // the configuration text given to pn_create
  Geom        g{};

  // chunk tables (indexed by chunk id)
  std::vector<int>    id_of_coord; // [Cz][Cy][Cx] -> id
  std::vector<int>    coord;       // [id][3]
  std::vector<int>    offset;      // [id][3] interior cell offset (z, y, x)
  std::vector<double> lim;         // [id][3][2] physical range (z, y, x)
  std::vector<double> rng;         // [id][3][2] range used for out-of-chunk tests
  std::vector<int>    nb;          // [id][27] neighbour ids

  int*    d_off = nullptr;
  double* d_lim = nullptr;
  double* d_rng = nullptr;
  int*    d_nb  = nullptr;

  hipStream_t stream = nullptr;

  // global padded periodic field arrays, component-major
  double* d_E = nullptr; // [3][vol]
  double* d_B = nullptr; // [3][vol]
  double* d_J = nullptr; // [4][vol]
  double* d_M = nullptr; // [Ns][14][vol] moments (PicChunk::um), allocated on first use
  double* d_ff = nullptr; // [3 levels][3][vol] Friedman history, only for theta != 0

  // precision = "fp32": FP32 fields and cell-relative FP32 particles.  The FP64
  // arrays above stay as the mirror used at the C-ABI boundary and for the FP64
  // energy and charge reductions.
  bool   fp32 = false;
  float* f_E  = nullptr; // [3][vol]
  float* f_B  = nullptr; // [3][vol]
  float* f_J  = nullptr; // [4][vol]
  float* f_ff = nullptr; // [3][3][vol], theta != 0 only
  P32    qa{}, qb{};     // current / sort target
  float* f_ro[3] = {nullptr, nullptr, nullptr}; // old cell-relative positions

  // particles: A = current, Bv = sort destination
  PArr     pa{}, pb{};
  double*  d_xo[3] = {nullptr, nullptr, nullptr}; // old positions (XtensorParticle::xv)
  int32_t* d_np    = nullptr;                     // [nseg]
  std::vector<int32_t> h_np;                      // host mirror, valid when np_host_valid
  bool     np_host_valid = false;

  SortScratch ss{};
  // device error flags as last seen by the host without a synchronisation:
  // every pn_step queues a copy of ss.errflag into pinned h_flag and records
  // flag_ev; the next pn_step reads h_flag once flag_ev has completed
  int32_t*    h_flag  = nullptr;
  hipEvent_t  flag_ev = nullptr;
  bool        flag_queued = false;
  ShockState  shk{};
  bool        counted = false; // keys hold a fresh count (push_position / count)
  bool        sorted  = false; // particles are in sorted order and cellstart is valid
  // the last sort (in pn_step) left its record move to the next row gather:
  // sorted slot i of the buffer pb-to-be holds record perm[i] of pa.  Every
  // other reader of the particles applies it first (materialize).
  bool        perm_pending = false;
  bool        defer_sort   = true; // PICNIX_DEFER_SORT=0: pn_step sorts in place
  // rho (component 0 of J) is the deposit of the current positions, left by
  // pn_step; any other call that moves particles or writes J clears it
  bool        rho_valid = false;
  // PICNIX_RHO=deposit: pn_step deposits rho from the particles every step
  // instead of carrying it by continuity (an independent charge check)
  bool        rho_deposit = false;
  bool        generic_kernels = false; // PICNIX_KERNELS=generic: one-thread-per-particle forms
  // per-chunk operator calls (chunk >= 0): chunks that deposited since J was
  // last cleared (the first deposit of a new round clears it; the Cur
  // exchange and whole-domain deposits end a round), the deposit's
  // "accumulate into J" switch, which chunks a per-chunk sort has sorted,
  // and device scratch for the saved counts and bin starts
  std::vector<char> dep_round;
  bool              dep_accumulate = false;
  std::vector<char> chunk_sorted;
  int32_t*          d_cs_saved     = nullptr;
  int         march_rows      = 0;     // PICNIX_MARCH_ROWS: bin rows per march block (0: auto)
  // launch shapes picked by the autotuner (pn_autotune); 0 = the heuristic
  int         dep_rows    = 0; // bin rows per block of the current march kernels
  int         rho_rows    = 0; // bin rows per block of the rho march kernels
  int         vel_generic = 0; // 1: one thread per particle instead of the bin-row gather
  int         vel_rows    = 0; // bin rows of one z plane per gather block (0: heuristic)
  int         vel_split   = 0; // 1: one gather block per species (0: all species)
  int         dep_queue   = 0; // 1: the march deposit runs persistent over a work queue
  int         vel_queue   = 0; // 1: the row gather runs persistent over a work queue
  int         fuse_push   = 0; // pn_step: the row gather also runs the position push
  int         tile32      = 1; // FP32 march deposit tile: 1 FP64 (default), 0 staged
                               // compensated pairs, 2 plain FP32, 3 compensated pairs by
                               // LDS atomics
  int         stage_cells = 0; // FP32 staged tile: cells per phase (0: 4)
  int*        d_wq        = nullptr; // work-queue counters, 2 per launch site (kWq*)
  int         ncu         = 0;       // compute units of the device
  int         device      = 0;       // HIP device the handle was created on
  double*     energy_scratch = nullptr; // this device's address of the energy scratch array
  struct AutoTune {
    int                 state = 0; // 0 off, 1 measuring, 2 done
    int                 step  = 0; // tuning steps taken
    std::vector<int>    cand;      // candidate bin rows per march block
    std::vector<double> t_dep, t_rho;
    std::vector<double> t_vel; // per gather shape (kVelRows / kVelSplit)
    double              t_q[2][2] = {{1e300, 1e300}, {1e300, 1e300}}; // [deposit][queue]
    double              t_v[4] = {1e300, 1e300, 1e300, 1e300}; // gather [2 fused + queue]
    bool                tune_vel = false;
  } tune;

  // pn_step as a replayed hipGraph (pn_set_step_graph).  The particle arrays
  // swap roles every step, so there is one graph per buffer parity.
  bool            step_graph  = false;
  hipGraphExec_t  gexec[2]    = {nullptr, nullptr};
  double          graph_dt[2] = {0, 0};
  void*           buf0_x      = nullptr; // first position array of parity 0

  // staging for host <-> device particle transfers
  double* d_stage     = nullptr;
  size_t  stage_elems = 0;

  // phase timing (pn_timing); events are recorded on `stream`
  bool                         timing = false;
  std::vector<hipEvent_t>      ev_pool;
  struct TimedPhase {
    int        phase;
    hipEvent_t start, stop;
  };
  std::vector<TimedPhase> ev_pending;
  // phases 0..7 as pn_timing_report; slot 8 (internal) times the rho march
  // deposit alone, for the autotuner
  double                  ph_ms[9]    = {0};
  int                     ph_count[9] = {0};

  std::string err;
};

// Records HIP events around one phase when timing is enabled.
constexpr int kPhaseRhoDeposit = 8;

struct PhaseTimer {
  Sim&       s;
  int        phase;
  hipEvent_t a = nullptr, b = nullptr;
  PhaseTimer(Sim& s_, int ph);
  ~PhaseTimer();
};

// bin rows per block of the FP64 march kernels (fixed > 0) or the heuristic
void march_split(const Geom& g, int fixed, int& zrows, int& nzr);

// kernels / launchers (pn_field.hip, pn_particle.hip, pn_sort.hip, pn_deposit.hip, pn_diag.hip)
void launch_push_bfd(Sim& s, double delt, int chunk = -1);
void launch_push_efd(Sim& s, double delt, int chunk = -1);
void launch_fill_ghost(Sim& s, double* f, int ncomp);
void launch_fill_ghost(Sim& s, float* f, int ncomp);
// zero fills and copies as kernels (no memset/memcpy nodes under graph capture)
// work-queue slots of the persistent row kernels (Sim::d_wq, 2 ints each)
enum { kWqDeposit = 0, kWqDeposit32 = 1, kWqGather = 2, kWqGather32 = 3, kWqSlots = 8 };
// persistent blocks for a queue of nitems: resident blocks per CU x CUs, at most nitems
int queue_blocks(Sim& s, const void* kernel, size_t lds, long nitems);
void launch_zero(Sim& s, double* p, long n);
void launch_zero(Sim& s, float* p, long n);
void launch_zero(Sim& s, int32_t* p, long n);
void launch_copy(Sim& s, int32_t* dst, const int32_t* src, long n);
void launch_init_friedman(Sim& s); // every Friedman level = E (no-op for theta = 0)
// fold the padding of components [k0, k1) of a [ncomp][vol] global array onto
// the periodic interior (the Cur / Mom exchanges); phase: timing slot
void launch_fold(Sim& s, double* F, int k0, int k1, int phase);
void launch_fold(Sim& s, float* F, int k0, int k1, int phase);
void launch_fold_current(Sim& s, int k0 = 0, int k1 = 4);
// moment deposit into d_M (allocated on first use), without the Mom exchange
void launch_deposit_moment(Sim& s);
bool launch_push_velocity(Sim& s, double delt, bool with_position = false);
void launch_push_position(Sim& s, double delt, int chunk = -1);
void launch_count(Sim& s);
void launch_migrate(Sim& s, bool defer = false);
void launch_sort(Sim& s);
// apply a pending deferred sort (record move pa -> pb, swap); no-op otherwise
void materialize(Sim& s);
// whether the next row gather can read the particles through a pending sort
bool gather_takes_perm(const Sim& s);
// What the Esirkepov deposit does with rho (component 0 of J):
//   Now   -- deposits it from the new positions, as the reference does;
//   Defer -- where the kernel allows, leaves it for launch_rho after the sort;
//   Keep  -- where the kernel allows, keeps rho^n for launch_rho_continuity.
// The result says what happened.
enum class RhoMode { Now, Defer, Keep };
enum class RhoDone { Deposited, Deferred, Kept };
RhoDone launch_deposit(Sim& s, double delt, RhoMode mode = RhoMode::Now);
// rho^{n+1} = rho^n - dt div J, after the current fold
void launch_rho_continuity(Sim& s, double delt);
// rho from the current positions; sorted_by_new_cell: cellstart describes them
void launch_rho(Sim& s, bool sorted_by_new_cell);
void launch_energy(Sim& s, double* out /* host: efd, bfd, divE, divB, particle[Ns] */);
// the same sums with no allocation (scratch array of the GPU, see pn_diag.hip)
void energy_static_init(Sim& s); // once per handle, at creation (device current)
int  launch_energy_static(Sim& s, double* out /* host, 4 + Ns */);
// moments (PicChunk::deposit_moment, when `deposit`) and the Mom exchange
void launch_moment_exchange(Sim& s, bool deposit);

// shock problem (pn_boundary.hip device side, pn_shock.cpp host side)
void launch_wall_emf(Sim& s);     // MainChunk::set_boundary_emf after an EMF exchange
void launch_wall_list(Sim& s);    // particles past x = Lx -> shk.d_list / d_count
void launch_wall_apply(Sim& s);   // their replacement values -> particles, bins, slots
void launch_inject_append(Sim& s);
void launch32_wall_apply(Sim& s);    // the FP32 form's versions (pn_fp32.hip)
void launch32_inject_append(Sim& s);
void launch_moment_energy(Sim& s, double* out /* host: particle[Ns] */);
// minus the rho of the particles the last shock_inject appended (nmax: the most
// injected into one segment)
void launch_rho_unject(Sim& s, int nmax);
void shock_setup(Sim& s);         // constants, per-chunk state, device buffers
void shock_release(Sim& s);
void shock_init_fields(Sim& s);   // MainChunk::setup field profile (host -> device)
void shock_load_particles(Sim& s);
void shock_position_boundary(Sim& s); // set_boundary_particle after the position push
void shock_inject(Sim& s);            // inject_particle (before the migration)

// FP32 form (pn_fp32.hip, pn_sort.hip)
void launch32_from_f64(Sim& s, int seg0, int nseg);        // pa (FP64 SoA) -> qa
void launch32_to_f64(Sim& s, int seg0, int nseg, bool old); // qa -> pa (and f_ro -> d_xo)
void launch32_fields_to_f64(Sim& s, int which);            // f_* -> d_* mirror
void launch32_fields_from_f64(Sim& s, int which);          // d_* mirror -> f_*
bool launch32_push_velocity(Sim& s, double delt, bool with_position = false);
void launch32_push_position(Sim& s, double delt);
RhoDone launch32_deposit(Sim& s, double delt, RhoMode mode);
void launch32_rho(Sim& s);
void launch32_migrate(Sim& s, bool defer = false);
void materialize32(Sim& s);
bool gather32_takes_perm(const Sim& s);
void launch32_sort(Sim& s);
void launch_aos_to_soa(Sim& s, int seg, int64_t np);
void launch_soa_to_aos(Sim& s, int seg, int64_t np);
// which: PN_FIELD_UF, PN_FIELD_UJ or PN_FIELD_UM
void launch_chunk_fields_get(Sim& s, int chunk, int which, double* d_out);
void launch_chunk_fields_set(Sim& s, int chunk, int which, const double* d_in);

} // namespace picnix
