/*
 * picnix_adjoint.h -- C ABI of the adjoint form of the PIC step (libpicnix_amd.so).
 *
 * No counterpart in the reference: pic-nix has no differentiable form.  The form
 * is a 2D periodic run (FP64 or FP32, pn_adj_params.precision) of one species (electrons, immobile neutralising
 * ions) with the hot path's operators at order 1 -- push_bfd / push_efd
 * (pic/engine/maxwell.hpp:231-348), the MC gather and the Boris push
 * (pic/engine/velocity.hpp:390-418, nix/primitives.hpp:163-189), the position
 * push (pic/engine/position.hpp:117-130) and the Esirkepov current
 * (nix/esirkepov.hpp:76-143, pic/engine/current.hpp:275-324) -- with the
 * particles kept in a fixed order, and its hand-written discrete adjoint: the
 * deposit's adjoint is a gather, the gather's adjoint a scatter, and the
 * per-particle Jacobians of the push come from dual numbers.
 *
 * The loss is the magnetic energy 1/2 sum B^2 after nsteps steps.  The parameters
 * theta = (r, eps_0 .. eps_{M-1}) set the initial state of a seeded Weibel set-up:
 *   u = (r vth N_x, vth N_y, vth N_z)            (N_* standard normal draws)
 *   x = x0 - sum_k eps_k k / |k|^2 sin(k . x0 + phi_k)   (density 1 + sum eps_k cos)
 * with x0 uniform, seeded modes k = 2 pi (m_x / Lx, m_y / Ly), phases phi_k.
 * pn_adj_forward runs the steps and keeps a checkpoint every K steps;
 * pn_adj_backward recomputes each K-step segment from its checkpoint and runs
 * the adjoint steps backwards, returning dL/dtheta.
 */
#ifndef PICNIX_ADJOINT_H
#define PICNIX_ADJOINT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PN_ADJ_MAX_MODES 16

typedef struct pn_adj pn_adj;

typedef struct pn_adj_params {
  int      Nx, Ny;      /* cells (periodic) */
  double   delh, delt;  /* cell size, time step */
  double   cc;          /* speed of light */
  int      nppc;        /* electrons per cell */
  double   qm, ro;      /* charge / mass and density: m = ro / nppc, q = qm m */
  double   vth;         /* thermal velocity along y and z (x: theta[0] * vth) */
  int      nmodes;      /* density modes M (<= PN_ADJ_MAX_MODES) */
  int      kmax;        /* mode numbers m_x, m_y drawn from [-kmax, kmax] */
  uint64_t seed;        /* positions, velocities, modes and phases */
  int      precision;   /* 0: FP64; 1: FP32 particles, fields and co-states (theta, the
                           loss and the gradient stay FP64; the loss and gradient sums
                           are reduced in FP64) */
} pn_adj_params;

int         pn_adj_create(const pn_adj_params* p, pn_adj** out);
int         pn_adj_destroy(pn_adj* a);
const char* pn_adj_last_error(const pn_adj* a);
/* modes: out[3 k + 0..2] = (m_x, m_y, phi) */
int pn_adj_modes(pn_adj* a, double* out);
/* loss L(theta) after nsteps steps; records a checkpoint every K steps */
int pn_adj_forward(pn_adj* a, const double* theta, int nsteps, int checkpoint_every, double* loss);
/* dL/dtheta of the last pn_adj_forward ([1 + nmodes] values) */
int pn_adj_backward(pn_adj* a, double* grad);
/* plain state access and steps, for checking the forward form against the hot path:
   particles [np][6] (x, y, z, ux, uy, uz), fields [3][Ny][Nx] */
int pn_adj_set_state(pn_adj* a, const double* xu6, int64_t np, const double* E, const double* B);
int pn_adj_get_state(pn_adj* a, double* xu6, double* E, double* B);
int pn_adj_step(pn_adj* a, int nsteps);
int pn_adj_num_particles(pn_adj* a, int64_t* np);

#ifdef __cplusplus
}
#endif
#endif
