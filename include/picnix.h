/*
 * picnix.h -- C ABI of the MI355X (gfx950) implementation of pic-nix's
 * per-timestep particle-in-cell hot path.
 *
 * One opaque handle owns every chunk of the domain (all on one GPU), the
 * device buffers, one HIP stream and the options parsed from the same
 * config.toml keys the reference reads.  Every call returns 0 on success or a
 * negative PN_ERR_* code; pn_last_error() then describes the failure.  No C++
 * exception crosses this boundary.  Host buffers are borrowed only for the
 * duration of a call.  A handle is not re-entrant: calls on one handle must not
 * overlap; the chunk loop lives inside each call.
 *
 * Each entry point names the reference interface it replaces.  The reference's
 * PicChunk operators act on one chunk; these act on all chunks at once
 * (chunk = -1), which is what PicApplication::push_openmp does in its
 * `omp for` loops (pic/pic_application.cpp:219-292).
 *
 * Layouts at this boundary are the reference's:
 *   particles : AoS [np][7] float64 (x, y, z, ux, uy, uz, id as int64 bits)
 *               (nix/particle.hpp:19, nix/xtensor_particle.hpp:14-15)
 *   uf        : [nz+2Nb][ny+2Nb][nx+2Nb][6] (Ex,Ey,Ez,Bx,By,Bz) per chunk
 *   uj        : [nz+2Nb][ny+2Nb][nx+2Nb][4] (rho,Jx,Jy,Jz) per chunk
 *               (pic/pic_chunk.cpp:106-122)
 * Inside the library fields live in one global periodic array per component
 * and particles in per-(chunk, species) SoA segments; see DESIGN.md.
 */
#ifndef PICNIX_H
#define PICNIX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PN_OK 0
#define PN_ERR_ARG -1        /* invalid argument                          */
#define PN_ERR_CONFIG -2     /* invalid or unsupported configuration      */
#define PN_ERR_HIP -3        /* HIP runtime error                         */
#define PN_ERR_CAPACITY -4   /* particle segment capacity exceeded        */
#define PN_ERR_UNSUPPORTED -5 /* option recognised but not implemented    */
#define PN_ERR_STATE -6      /* call sequence invalid for current state   */

/* boundary modes, pic/pic.hpp:40-46 */
#define PN_BOUNDARY_EMF 0
#define PN_BOUNDARY_CUR 1
#define PN_BOUNDARY_MOM 2
#define PN_BOUNDARY_PARTICLE 3

/* field selectors for pn_get_fields / pn_set_fields */
#define PN_FIELD_UF 0
#define PN_FIELD_UJ 1
#define PN_FIELD_UM 3 /* moments [z][y][x][Ns][14] (PicChunk::um); read only */

typedef struct pn_sim pn_sim;

/* Create a simulation from config.toml text (the reference's [application.option]
 * and [parameter] keys, nix/cfgparser.hpp:145-217, pic/pic_chunk.cpp:135-262),
 * allocate every chunk on the current HIP device and run the example loader
 * (example/thermal/main.cpp:15-113, example/beam/main.cpp:15-117) followed by the
 * initial sort (PicChunk::sort_particle, pic/pic_chunk.cpp:447-453) and the
 * initial EMF exchange (PicApplication::setup_chunks, pic/pic_application.cpp:106-130).
 * load = 0 skips the particle loader (empty chunks). */
int pn_create(const char* toml, int load, pn_sim** out);
/* Release every resource (reference: chunk destructors). */
int pn_destroy(pn_sim* sim);
/* Text of the last error on this handle ("" if none). */
const char* pn_last_error(const pn_sim* sim);

/* Geometry queries (nix::Chunk accessors, nix/chunk.hpp:96-352).
 * info[0..2] chunk coord (cz,cy,cx); [3..5] cell offset (z,y,x);
 * [6..8] ghost-inclusive extent (nz,ny,nx); [9] Nb; [10] Ng. */
int pn_num_chunks(const pn_sim* sim, int* nchunk);
int pn_num_species(const pn_sim* sim, int* ns);
int pn_chunk_info(const pn_sim* sim, int chunk, int* info);

/* Particles of one (chunk, species) as AoS [np][7] (XtensorParticle::xu,
 * nix/xtensor_particle.hpp:14).  pn_get_particles writes at most cap records. */
int pn_get_np(const pn_sim* sim, int chunk, int species, int64_t* np);
int pn_get_particles(pn_sim* sim, int chunk, int species, double* xu7, int64_t cap,
                     int64_t* np);
int pn_set_particles(pn_sim* sim, int chunk, int species, const double* xu7, int64_t np);
/* Old positions written by pn_push_position (XtensorParticle::xv, columns 0..2). */
int pn_get_particles_old(pn_sim* sim, int chunk, int species, double* xv3, int64_t cap);
/* Cell index of every particle from the last count (XtensorParticle::gindex,
 * nix/xtensor_particle.hpp:16); out-of-chunk particles report Ng. */
int pn_get_gindex(pn_sim* sim, int chunk, int species, int32_t* gindex, int64_t cap);
/* First particle of every cell after the last sort (XtensorParticle::pindex, Ng+1 entries). */
int pn_get_pindex(pn_sim* sim, int chunk, int species, int32_t* pindex);

/* Ghost-inclusive per-chunk field arrays (PicChunk::uf / uj).  Ghost cells of
 * uf read back as the periodic images held after an EMF exchange. */
int pn_get_fields(pn_sim* sim, int chunk, int which, double* data);
int pn_set_fields(pn_sim* sim, int chunk, int which, const double* data);

/* Per-operator calls; chunk must be -1 (all chunks). */
int pn_push_bfd(pn_sim* sim, int chunk, double delt);        /* PicChunk::push_bfd, pic/pic_chunk.cpp:544-550 */
int pn_push_efd(pn_sim* sim, int chunk, double delt);        /* PicChunk::push_efd, pic/pic_chunk.cpp:536-542 */
int pn_push_velocity(pn_sim* sim, int chunk, double delt);   /* PicChunk::push_velocity, pic/pic_chunk.cpp:501-512 */
int pn_push_position(pn_sim* sim, int chunk, double delt);   /* PicChunk::push_position + count, pic/pic_chunk.cpp:491-499 */
int pn_deposit_current(pn_sim* sim, int chunk, double delt); /* PicChunk::deposit_current, pic/pic_chunk.cpp:514-523 */
int pn_sort_particle(pn_sim* sim, int chunk);                /* PicChunk::sort_particle, pic/pic_chunk.cpp:447-453 */
int pn_deposit_moment(pn_sim* sim, int chunk);               /* PicChunk::deposit_moment, pic/pic_chunk.cpp:525-534 */
/* Reset the Friedman-filter history to the current E (PicChunk::init_friedman,
 * pic/pic_chunk.cpp:264-268; pn_create already does it once).  No-op for
 * friedman = 0. */
int pn_init_friedman(pn_sim* sim);
/* Boundary exchange of one mode: pack+begin+end+unpack
 * (PicChunk::set_boundary_{pack,begin,end,unpack}, pic/pic_chunk.cpp:287-405).
 * PARTICLE migrates leavers to their neighbour chunk and sorts; MOM adds the
 * moment ghosts into the neighbours (after pn_deposit_moment). */
int pn_exchange(pn_sim* sim, int mode);
/* One full timestep in the reference operator order (PicApplication::push_openmp,
 * pic/pic_application.cpp:219-292). */
int pn_step(pn_sim* sim, double delt);
/* Replay pn_step as a captured hipGraph (enable = 1) or launch it eagerly (0,
 * the default).  No counterpart in the reference (its step is a loop of OpenMP
 * phases); for callers that run many steps back to back.  The graph is recorded
 * at the first step that starts from sorted particles, once per buffer parity
 * and time step; phase timing (pn_timing) runs steps eagerly. */
int pn_set_step_graph(pn_sim* sim, int enable);
/* Field energy E^2/2, B^2/2 and per-species kinetic energy (PicChunk::get_energy,
 * pic/pic_chunk.cpp:407-439, summed over chunks as HistoryDiag does). */
int pn_get_energy(pn_sim* sim, double* efd, double* bfd, double* particle);
/* pn_get_energy without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (tight serving loops that read
 * the history every few steps, harnesses that forbid hipMalloc).  It
 * synchronises the handle's stream, so it is not for graph capture.
 * pn_get_energy takes a
 * partial-sum buffer with hipMallocAsync on every call; this variant keeps its
 * partial sums (at most 2048 rows of 12 doubles, 197 KB) in a file-scope device
 * array that each GPU holds until the process ends, and calls no allocating
 * routine.  Every call on one GPU uses the same array (each call writes every
 * entry it reads, so nothing carries over), so calls that overlap on one GPU
 * race.  Periodic problems only (PN_ERR_UNSUPPORTED for a physical boundary);
 * the handle's device must be current (PN_ERR_STATE otherwise).  Same outputs as
 * pn_get_energy, summed in a fixed order. */
int pn_get_energy_static(pn_sim* sim, double* efd, double* bfd, double* particle);
/* Signed sums of div(E) - rho and div(B) (PicChunk::get_diverror, pic/pic_chunk.cpp:441-445). */
int pn_get_diverror(pn_sim* sim, double* efd, double* bfd);
/* Block until queued device work on the handle's stream is done. */
int pn_synchronize(pn_sim* sim);

/* Autotuner of the particle kernels' launch shapes (the bin rows per block of the
 * current and rho march kernels, the bin rows and species per gather block or
 * one thread per particle, whether the march deposit and the row gather run one
 * block per work item or persistent blocks walking a device work queue, and
 * whether the FP32 row gather also runs the position push).
 * No reference counterpart (the reference picks its loop forms by the
 * [application.option] vectorization table, pic/pic_chunk.cpp:140-189).
 * enable = 1 starts a session: the next pn_get_tuning()[1] pn_step calls run
 * eagerly, each timing candidate shapes on that step, then the fastest shape
 * of every kernel stays for the rest of the run (step graphs are re-captured).
 * The tuning steps are ordinary steps of the run: every candidate computes the
 * same step, up to the order of the current sums.  enable = 0 returns to the
 * built-in heuristics.  pn_get_tuning: out9 = {state (0 off, 1 measuring,
 * 2 done), steps left, bin rows of the current march (0 = heuristic), bin rows of
 * the rho march, bin rows per gather block (0 = heuristic, -1 = one thread per
 * particle), 1 if a gather block serves one species (0: all species), 1 if the
 * march deposit runs persistent over the work queue, the same for the gather, 1
 * if the FP32 row gather also runs the position push}. */
int pn_autotune(pn_sim* sim, int enable);
int pn_get_tuning(pn_sim* sim, int* out9);

/* Upstream boundary state of one chunk for [application] problem = 'shock' (the
 * values MainChunk keeps in option["random_seed"] and option["boundary"],
 * example/shock/main.cpp:19-24,212-226,440-442,528-534): out6 = {random_seed,
 * lastid, influx[0], influx[1], efflux[0], efflux[1]}.  PN_ERR_STATE for other
 * problems. */
int pn_get_boundary_state(pn_sim* sim, int chunk, int64_t* out6);

/* Snapshot outputs in the reference's on-disk format (one process, "mpiio"
 * layout): <dir>/<step %08d>.data (raw native-endian f8, chunk after chunk in
 * chunk-id order) + <dir>/<step %08d>.json (nixio metadata per dataset, meta
 * record).
 * pn_write_field_diag: FieldDiag (pic/diag/field.hpp:66-185): "uf" colocated
 * at cell centres and block-averaged by decimate (>= 1), or the raw interior
 * (decimate < 0); then "um" after the moment deposit + Mom exchange.  FP64 only.
 * pn_write_particle_diag: ParticleDiag (pic/diag/particle.hpp:56-130): per
 * species "up%02d", the records of a random fraction of every chunk's particles
 * (std::mt19937_64(seed) shuffle, kept in index order).
 * pn_write_profile: the application profile (nix/application.cpp:11-26 saved by
 * nix/statehandler.hpp:44-63, e.g. <basedir>/profile.msgpack): configuration,
 * chunk map and step/time, as msgpack. */
int pn_write_field_diag(pn_sim* sim, const char* dir, int step, double time, int decimate);
int pn_write_particle_diag(pn_sim* sim, const char* dir, int step, double time, double fraction,
                           int seed);
int pn_write_profile(pn_sim* sim, const char* path, int step, double time);

/* Checkpoint and restart (StateHandler::save / load, nix/statehandler.hpp:36-147,
 * with PicChunk::pack / unpack, pic/pic_chunk.cpp:59-104): <prefix>.msgpack holds
 * the application record (as pn_write_profile, plus the library's state flags
 * and the block table) and <prefix>.data the raw device state: E, B, J, the
 * Friedman history, the bin starts and every segment's particle records.  Load
 * into a handle created from the same configuration (pn_create with load = 0
 * or 1); the grid, chunks, [parameter] and [application.option] tables and the
 * precision must match (PN_ERR_CONFIG otherwise, as Application::from_json,
 * nix/application.cpp:28-48).  step/time return the saved values. */
int pn_save_state(pn_sim* sim, const char* prefix, int step, double time);
int pn_load_state(pn_sim* sim, const char* prefix, int* step, double* time);

/* Phase timing with HIP events recorded on the handle's stream.  No counterpart
 * in the reference beyond its per-step wall-clock "push" time
 * (PicApplication::push, pic/pic_application.cpp:196-216); it exists so that a
 * benchmark can report each kernel's launch duration.  pn_timing(sim, 1) clears
 * the accumulators and starts recording; pn_timing_report() synchronises and
 * returns the summed milliseconds and launch count of every phase. */
#define PN_PHASE_BFD 0      /* B curl update + ghost refresh            */
#define PN_PHASE_VELOCITY 1 /* gather + Boris                           */
#define PN_PHASE_POSITION 2 /* position push + count                    */
#define PN_PHASE_DEPOSIT 3  /* Esirkepov deposit kernel                 */
#define PN_PHASE_FOLD 4     /* current zero + periodic fold             */
#define PN_PHASE_EFD 5      /* E curl update + ghost refresh            */
#define PN_PHASE_SORT 6     /* migration + counting sort                */
#define PN_PHASE_RHO 7      /* charge density deposit + fold            */
#define PN_NPHASE 8
int pn_timing(pn_sim* sim, int enable);
int pn_timing_report(pn_sim* sim, double* ms_total, int* launches);

#ifdef __cplusplus
}
#endif
#endif
