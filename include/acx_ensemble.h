/*
 * acx_ensemble.h -- a batch of independent 3D acoustic problems stepped by
 * one persistent kernel over a shared tile work queue.
 *
 * No counterpart in the reference (each acoustic_petsc program runs one grid,
 * 3d/weno4/plane_wave/hype.c:102-183).  Every problem is the ACOUSTIC4
 * operator with the shipped linear CWENO4 weights on its own periodic grid
 * (edge lengths may differ per problem and axis, common cell size h), stepped
 * with the compensated FP32 SSP-RK3 of acx_ssp_rks3_step_f32c.  One stage is
 * one launch of a persistent kernel whose workgroups pull (problem, tile,
 * z-chunk) items from an atomic counter; a step (counter resets + 4 stage
 * launches) can be captured once in a hipGraph and replayed.  The result of
 * every problem is bitwise the result of the single-problem FP32 stepper.
 */
#ifndef ACX_ENSEMBLE_H
#define ACX_ENSEMBLE_H

#include "acx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct acx_ensemble acx_ensemble;

/* nprob problems with n[3*p + d] cells along axis d; all share h, c0 = rho0 = 1.
 * Allocates the FP32 state (hi, lo and work arrays) on the current device. */
int acx_ensemble_create(int nprob, const int32_t *n, double h, acx_ensemble **out);
int acx_ensemble_destroy(acx_ensemble *e);

/* cell offset of every problem in the packed state (nprob + 1 entries) */
int acx_ensemble_offsets(const acx_ensemble *e, size_t *offsets);

/* device pointers of the current state (hi, lo: 4 floats per cell, problems
 * packed one after another in offset order) */
int acx_ensemble_state(acx_ensemble *e, float **hi, float **lo);

/* device-to-device copies of the whole packed state into / out of the ensemble */
int acx_ensemble_copy_in(acx_ensemble *e, const float *hi, const float *lo, void *stream);
int acx_ensemble_copy_out(const acx_ensemble *e, float *hi, float *lo, void *stream);

/* tile configurations (acx_ensemble_num_configs() of them): 0..5 = (rows per
 * tile 8 or 16) x (z-chunk 16, 32, 64), 6 = (8, 128), 7 = (16, 128) */
int acx_ensemble_set_config(acx_ensemble *e, int config);
int acx_ensemble_get_config(const acx_ensemble *e);
int acx_ensemble_num_configs(void);

/* One step: 4 persistent-kernel stages on `stream`. */
int acx_ensemble_step(acx_ensemble *e, double dt, void *stream);

/* nsteps steps (even) by replaying a captured two-step hipGraph; the graph is
 * captured on first use and again whenever dt or the configuration changes. */
int acx_ensemble_graph_steps(acx_ensemble *e, double dt, int nsteps, void *stream);

/* Times every configuration on this ensemble (a few steps each, then the
 * state is restored) and keeps the fastest; returns it in *best. */
int acx_ensemble_autotune(acx_ensemble *e, double dt, void *stream, int *best, float *ms_per_step);

#ifdef __cplusplus
}
#endif
#endif
