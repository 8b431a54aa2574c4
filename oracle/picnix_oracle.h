/*
 * picnix_oracle.h -- C ABI of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * It restates, in plain C++17 (compiled with -ffp-contract=off), the
 * per-timestep hot path of amanotk/pic-nix in the reference's own data
 * layout (per-chunk ghost-inclusive arrays, AoS [Np][7] particle records,
 * per-cell counting sort with 8 lane sub-bins) and the reference's operator
 * order (pic/pic_application.cpp:219-292).  Halo exchanges are in-memory
 * copies that follow the reference's 26-direction order.
 */
#ifndef PICNIX_ORACLE_H
#define PICNIX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNO_MAX_SPECIES 8

typedef struct pno_params {
  int    Nx, Ny, Nz;     /* global cells (reference [parameter] Nx/Ny/Nz) */
  int    Cx, Cy, Cz;     /* chunk counts (reference Cx/Cy/Cz)             */
  int    Ns;             /* number of species                             */
  int    order;          /* shape order 1..4 ([application.option] order)  */
  int    pusher;         /* 0 Boris, 1 Vay, 2 Higuera-Cary                 */
  int    load;           /* 1: run the seeded loader, 0: empty particles   */
  int    loader;         /* 0 thermal/beam (uniform), 1 density profile    */
  int    seed_offset;    /* seed = chunk id + seed_offset (0 = 'fixed')    */
  double cc, delt, delh;
  double friedman;       /* Friedman filter theta                          */
  double buffer_ratio;   /* particle buffer headroom                       */
  double Ex, Ey, Ez, Bx, By, Bz;
  int    np[PNO_MAX_SPECIES];
  double qm[PNO_MAX_SPECIES], ro[PNO_MAX_SPECIES], vt[PNO_MAX_SPECIES];
  double vx[PNO_MAX_SPECIES], vy[PNO_MAX_SPECIES], vz[PNO_MAX_SPECIES];
  int    interp;         /* 0 MC, 1 WT ([application.option] interpolation) */
  /* problem = 1: example/shock (non-periodic x, reflecting wall at x = 0, upstream
     injection at x = Lx); the [parameter] keys of example/shock/config.toml */
  int    problem;
  int    emf_order;      /* shock: 0 the reference's order (set_boundary_field after the
                            EMF unpack), 1 also before the pack (split-independent) */
  int    nppc;
  double wp, mime, u0, theta, phi, sigma, betae, betai, taper;
  double density, ramp_x, ramp_w; /* density n(x): 1 below ramp_x, linear ramp over
                                     ramp_w cells, `density` above (1 = the reference) */
} pno_params;

typedef struct pno_sim pno_sim;

pno_sim* pno_create(const pno_params* p);
void     pno_destroy(pno_sim* s);
int      pno_num_chunks(const pno_sim* s);
/* info[0..2] = chunk coord (cz,cy,cx); [3..5] = cell offset (z,y,x);
 * [6..8] = dims incl. ghosts (nz,ny,nx); [9] = Nb; [10] = Ng (sort bins) */
void     pno_chunk_info(const pno_sim* s, int id, int* info);
int64_t  pno_get_np(const pno_sim* s, int chunk, int species);
void     pno_get_particles(const pno_sim* s, int chunk, int species, double* xu7);
void     pno_get_particles_old(const pno_sim* s, int chunk, int species, double* xv7);
void     pno_set_particles(pno_sim* s, int chunk, int species, const double* xu7, int64_t np);
void     pno_get_gindex(const pno_sim* s, int chunk, int species, int32_t* out);
void     pno_get_pindex(const pno_sim* s, int chunk, int species, int32_t* out);
/* which: 0 = uf [z][y][x][6], 1 = uj [z][y][x][4], 2 = ff [z][y][x][3][3] (the
 * reference allocates [3][6] and uses components 0..2 only), 3 = um [z][y][x][Ns][14] */
void     pno_get_field(const pno_sim* s, int chunk, int which, double* out);
void     pno_set_field(pno_sim* s, int chunk, int which, const double* in);

void pno_push_bfd(pno_sim* s, double delt);
void pno_push_efd(pno_sim* s, double delt);
void pno_push_velocity(pno_sim* s, double delt);
void pno_push_position(pno_sim* s, double delt); /* includes count(reset) */
void pno_deposit_current(pno_sim* s, double delt);
void pno_sort_particle(pno_sim* s);
/* mode: 0 Emf, 1 Cur, 2 Mom, 3 Particle (pack+send+recv+unpack+sort) */
void pno_exchange(pno_sim* s, int mode);
void pno_step(pno_sim* s);
void pno_deposit_moment(pno_sim* s);
/* out[0] div(E)-rho, [1] div(B), [2] E^2/2, [3] B^2/2, [4..4+Ns) particle */
void pno_history(pno_sim* s, double* out);
/* set OpenMP threads used by the chunk loops (<=0: leave default) */
void pno_set_threads(int n);

/* ---- primitive entry points (for closed-form and invariant tests) ---- */
void   pno_shape_mc(int order, double x, double X, double rdx, double* s);
/* WT weights (nix/primitives.hpp:331-566); dt = c dt / dx, rdt = 1 / dt */
void   pno_shape_wt(int order, double x, double X, double rdx, double dt, double rdt, double* s);
int    pno_digitize(double x, double xmin, double rdx);
double pno_lorentz_factor(double ux, double uy, double uz, double rc);
void   pno_push_boris(double* u, const double* eb, double cc);
void   pno_push_vay(double* u, const double* eb, double cc);
void   pno_push_higuera_cary(double* u, const double* eb, double cc);
/* PicChunk::init_friedman on every chunk: all Friedman levels = E */
void   pno_init_friedman(pno_sim* s);
/* single-particle Esirkepov: ss_old/new from positions, returns tile (O+3)^3*4 */
void   pno_esirkepov3d(int order, const double* xv, const double* xu, double q, double delt,
                       double dx, double dy, double dz, double* tile, int* origin);
void   pno_esirkepov2d(int order, const double* xv, const double* xu, double q, double delt,
                       double dx, double dy, double* tile, int* origin);
/* one particle's 1D Esirkepov tile [order+3][4] (local1d, pic/engine/current.hpp:207-258) */
void pno_esirkepov1d(int order, const double* xv, const double* xu, double q, double delt,
                     double dx, double* tile, int* origin);

/* ---- random number generation restated from libstdc++ (GCC 11.4) ---- */
uint64_t pno_mt19937_64(uint64_t seed, int64_t skip);
void     pno_uniform(uint64_t seed, int64_t n, double* out);
void     pno_normal(uint64_t seed, int64_t n, double* out);

/* libstdc++ gamma_distribution<double>(alpha, beta) and poisson_distribution<int>(mean < 12)
   driven by mt19937_64(seed); nix::MaxwellJuttner(temperature, drift) -> n x (ux, uy, uz) */
void pno_gamma(double alpha, double beta, uint64_t seed, int64_t n, double* out);
int  pno_poisson(double mean, uint64_t seed, int64_t n, int* out);
void pno_maxwell_juttner(double temperature, double drift, uint64_t seed, int64_t n,
                         double* out3);
/* shock problem state of a chunk: [random_seed, lastid, influx[0..1], efflux[0..1]] */
void pno_shock_state(const pno_sim* s, int chunk, int64_t* out6);

/* Tangent-linear reference of the adjoint form (include/picnix_adjoint.h): nsteps
   2D periodic order-1 steps from particles xu6 [np][6] with E = B = 0, in
   forward-mode dual numbers along the initial-state direction dxu6; returns
   L = 1/2 sum B^2 and its directional derivative dL. */
void pno_adj_tangent(int nx, int ny, double delh, double delt, double cc, double q, double m,
                     int64_t np, const double* xu6, const double* dxu6, int nsteps, double* loss,
                     double* dloss);

/* ---- chunk map (generalized Hilbert curve) ---- */
void pno_chunk_map(int Cz, int Cy, int Cx, int* id_of_coord);

#ifdef __cplusplus
}
#endif
#endif
