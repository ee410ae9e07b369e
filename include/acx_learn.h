/*
 * acx_learn.h -- learned troubled-cell indicator and selective stepping.
 *
 * No counterpart in the reference: acoustic_petsc always runs the full
 * CWENO4 weights (3d/weno4/plane_wave/reconstruction.c:38-259 and its copies).  This adds a small
 * MLP that predicts, per cell and component, whether the CWENO weights leave
 * the linear weights: the L1 distance sum_k |w_k - gamma_k| over the eight
 * normalised weights (R4 and the seven R3 stencils,
 * reconstruction.c:205-218) exceeds a threshold.  With w = gamma the CWENO
 * combination (reconstruction.c:223-257) is exactly the R4 polynomial, and its
 * change is bounded by that distance, so cells below it can take the linear
 * 4th-order polynomial while full CWENO stays the fallback everywhere else.
 *
 * Network: 19 inputs (18 normalised undivided differences, 6 per axis, plus
 * a log-amplitude term), padded to 32 -> 32 -> 32 -> 1 sigmoid, hidden
 * activation ReLU (default) or tanh (ACX_MLP_TANH).
 * Inference and the training forward run on MFMA (bf16 operands, fp32
 * accumulation); the backward pass and Adam run in fp32.
 */
#ifndef ACX_LEARN_H
#define ACX_LEARN_H

#include "acx.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ACX_MLP_IN   32   /* padded input width (19 used) */
#define ACX_MLP_H    32
#define ACX_MLP_NFEAT 19

/* fp32 master parameters, row-major [out][in]; one contiguous block of
 * ACX_MLP_NPARAM floats in this order: w1, b1, w2, b2, w3, b3. */
#define ACX_MLP_NPARAM (32 * 32 + 32 + 32 * 32 + 32 + 32 + 1)

#define ACX_MLP_RELU 0
#define ACX_MLP_TANH 1

/* Packs the fp32 parameters into the MFMA operand layout (device to device);
 * `packed` needs acx_mlp_packed_bytes() bytes and carries the hidden
 * activation (acx_mlp_pack: ReLU).  Forward and inference use it. */
size_t acx_mlp_packed_bytes(void);
int acx_mlp_pack(const float *params, void *packed, void *stream);
int acx_mlp_pack_act(const float *params, int activation, void *packed, void *stream);

/* Feature vectors of every (cell, component): X[(cell*4 + v)*32 + f] (fp32,
 * zero padded).  U is an ACOUSTIC4 state (fp32, interlaced). */
int acx_features_f32(const acx_grid *g, const float *U, float *X, void *stream);

/* Training labels from the solver's own weights: for every (cell, component)
 * of the FP64 state U, dev = sum_k |w_k - gamma_k| of the CWENO4 weights with
 * the grid's gamma_hi/gamma_lo, label = dev > thr.  dev may be NULL. */
int acx_weno_labels_f64(const acx_grid *g, const double *U, double thr, uint8_t *labels, float *dev, void *stream);

/* MFMA forward of n samples X[n][32]: Y[n] = sigmoid output; H1, H2 (n x 32,
 * may be NULL) keep the activations for the backward pass. */
int acx_mlp_forward(const void *packed, const float *X, int n, float *Y, float *H1, float *H2, void *stream);

/* Gradient of the class-weighted binary cross entropy
 *   L = -(1/n) sum_i c_i [t_i log y_i + (1 - t_i) log(1 - y_i)],
 *   c_i = pos_weight if t_i = 1 else 1,
 * with respect to the parameters (grad: ACX_MLP_NPARAM floats, overwritten).
 * Uses the bf16-rounded weights and inputs the forward pass used. */
int acx_mlp_backward(const float *params, const float *X, const float *H1, const float *H2, const float *Y,
                     const uint8_t *T, int n, float pos_weight, float *grad, void *stream);
/* The same for either activation (acx_mlp_backward: ReLU); the derivative
 * is taken from the activation H the forward pass kept (ReLU: H > 0, tanh:
 * 1 - H^2). */
int acx_mlp_backward_act(const float *params, int activation, const float *X, const float *H1, const float *H2,
                         const float *Y, const uint8_t *T, int n, float pos_weight, float *grad, void *stream);

/* One Adam step (fp32): params -= lr * mhat / (sqrt(vhat) + eps). */
int acx_mlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr, float beta1,
                 float beta2, float eps, void *stream);

/* Fused inference over a grid: features of HI (+ D if not NULL), MLP,
 * flag bit v of flags[cell] set when y(cell, v) > thr (full CWENO). */
int acx_indicator_flags(const acx_grid *g, const float *HI, const float *D, const void *packed, float thr,
                        uint8_t *flags, void *stream);

/* Full-WENO judge of the selective step.  Reconstructs every (cell,
 * component) of the FP32 state HI with the full CWENO4 (grid gammas, FP32
 * ratio-form weights).  Where `used` -- the flags a run of selective steps
 * ran with -- has the component's bit clear, that run used the linear
 * polynomial: the judge measures the largest difference between the CWENO4
 * and the linear face averages of the cell (six faces) and counts a miss
 * when it exceeds tol (absolute).  *missed (host; synchronises the stream)
 * receives the number of misses, *worst (may be NULL) the largest
 * difference.  Independently, label bit v of out[cell] is set when the eight
 * normalised weights leave the linear weights, sum_k |w_k - gamma_k| > thr
 * (the training label of acx_weno_labels_f64), and out[cell] = base[cell] |
 * labels (base may be NULL).  out or missed may be NULL, not both; missed
 * needs used. */
int acx_weno_judge_f32(const acx_grid *g, const float *HI, const uint8_t *used, const uint8_t *base, double thr,
                       double tol, uint8_t *out, long long *missed, double *worst, void *stream);

/* flags_out[cell] = all four component bits (0xF) if any bit of flags_in is
 * set over the cell and its 6 face neighbours, else 0.  The components are
 * spread together because the system couples them: within one step a steep
 * pressure front makes the velocity steep where it was zero (a flag taken at
 * the step start from u = 0 would miss it). */
int acx_flags_dilate(const acx_grid *g, const uint8_t *flags_in, uint8_t *flags_out, void *stream);

/* Number of set bits over all cells (host result). */
int acx_flags_count(const uint8_t *flags, size_t ncells, long long *bits, void *stream);

/* One compensated FP32 SSP-RK3 step with the selective reconstruction:
 * full CWENO4 (grid gammas) where the flag bit of a component is set, the
 * linear 4th-order polynomial elsewhere.  Each stage runs the linear-scheme
 * kernel over every cell, then the CWENO kernel over the tiles (30 x 14 x 8
 * cells; ACX_ZC_PATCH sets the 8, read once per process) that hold a flagged
 * cell or a face neighbour of one, visiting only those tiles.  The tile mask
 * (acx_sel_mask) is made from the flags on every call, into a buffer taken
 * from the stream-ordered pool of `stream` for the call.  ACOUSTIC4. */
int acx_ssp_rks3_step_f32c_sel(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                               float *LO_OUT, float *D0, float *D1, const uint8_t *flags, double dt, void *stream);

/* acx_ssp_rks3_step_f32c_sel without any allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The tile
 * mask is kept in a fixed static device array sized for
 * ACX_SEL_NOALLOC_MAX_TILES patch tiles (1.25 MiB per GPU, held until the
 * process ends): grids with more patch tiles are rejected with ACX_EINVAL.  Calls on one GPU share
 * that array, so calls that overlap on one GPU (other streams or threads)
 * race; the mask is rewritten on every call.  The non-finite watch applies
 * once another entry point has created the device's watch word.  Results
 * equal acx_ssp_rks3_step_f32c_sel bit for bit. */
#define ACX_SEL_NOALLOC_MAX_TILES 262144
int acx_ssp_rks3_step_f32c_sel_noalloc(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                                       float *LO_OUT, float *D0, float *D1, const uint8_t *flags, double dt,
                                       void *stream);

/* The tile mask of the selective step for these flags: a buffer of
 * acx_sel_mask_bytes(g) bytes (0 for an invalid grid), 4-byte aligned.  It
 * holds one byte per patch tile, 1 where the tile holds a flagged cell or a
 * face neighbour of one, then, from the next 4-byte boundary, the marked
 * tiles as int32: their number, then their indices in ascending order (the
 * list the CWENO pass walks).  Masks are made only by acx_sel_mask. */
size_t acx_sel_mask_bytes(const acx_grid *g);
int acx_sel_mask(const acx_grid *g, const uint8_t *flags, uint8_t *mask, void *stream);

/* acx_ssp_rks3_step_f32c_sel with the tile mask made by acx_sel_mask from the
 * same flags (flags change only when the indicator runs, so a caller that
 * keeps flags over several steps makes the mask once).  No library scratch:
 * steps on different states may run concurrently. */
int acx_ssp_rks3_step_f32c_selm(const acx_grid *g, const float *HI, const float *LO, float *HI_OUT,
                                float *LO_OUT, float *D0, float *D1, const uint8_t *flags, const uint8_t *mask,
                                double dt, void *stream);

#ifdef __cplusplus
}
#endif
#endif
