/* C-ABI of the MI355X-native IGM ice-flow hot path (libigm_amd.so).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative code on error;
 *    igm_last_error() then describes the failure (per host thread).
 *  - Pointers are device pointers unless a parameter says "host".  The caller
 *    owns every buffer; no entry point allocates device memory.
 *  - All work is enqueued on the given HIP stream (hipStream_t passed as
 *    void*; NULL = the default stream).  Nothing synchronises the host.
 *  - Velocity-shaped buffers (U, V, gradients, Adam moments) share one layout,
 *    described by element strides: batch, level (k), row (j), column (i).
 *    Planar state layout (Nz, Ny, Nx): kstride = Ny*Nx, rstride = Nx,
 *    cstride = 1 with separate U and V pointers.  Emulator layout
 *    Y (N, Ny, Nx, 2Nz) (emulate/emulate.py:139, iceflow/utils.py:133-139):
 *    U = Y, V = Y + Nz, kstride = 1, cstride = 2Nz, rstride = Nx*2Nz.
 *
 * The reference (IGM, TensorFlow) has no native interface for this path;
 * each entry point names the reference function(s) it replaces.
 */
#ifndef IGM_AMD_H
#define IGM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IGM_MAX_NZ 32

/* Physics parameters, conf/processes/iceflow.yaml:4-22. */
typedef struct IgmPhysics {
  double exp_glen;       /* n (Glen) */
  double exp_weertman;   /* m (Weertman) */
  double regu_glen;
  double regu_weertman;
  double thr_ice_thk;
  double min_sr;
  double max_sr;
  double ice_density;
  double gravity_cst;
  int32_t force_negative_gravitational_energy; /* must be 0 (not supported) */
} IgmPhysics;

/* Legacy-Keras Adam step constants (solve/solve.py:11-20, 84).
 * alpha = lr * sqrt(1 - beta2^t) / (1 - beta1^t), computed by the caller. */
typedef struct IgmAdam {
  double alpha;
  double beta1;
  double beta2;
  double epsilon;
} IgmAdam;

/* Device-side EarlyStopping state (iceflow/utils.py:80-111). */
typedef struct IgmEarlyStop {
  double best_loss;
  double relative_min_delta;
  int32_t has_best;
  int32_t wait;
  int32_t patience;
  int32_t stopped;     /* set to 1 when should_stop() returned True */
  int32_t stop_iter;   /* iteration index at which it stopped */
  int32_t iters_done;  /* number of iterations whose update was applied */
  int32_t passes;      /* igm_energy_solve_pass_f32 launches that applied an update */
} IgmEarlyStop;

const char* igm_last_error(void);

/* ---------------------------------------------------------------- energy -- */

/* Number of double words of the `partials` workspace that
 * igm_energy_pass_* and igm_energy_pass2_f32 need for this problem size. */
int64_t igm_energy_workspace_doubles(int32_t nbatch, int32_t Ny, int32_t Nx);

/* Per-cell, velocity-independent coefficients of the staggered energy:
 * stag4h of thk, B = 2 A^(-1/n) and slidingco, the staggered slopes of usurf
 * and of the bed (usurf - thk) from the dX field, the floating damping and
 * the 1/S mean weights.  Replaces the U-independent half of
 * energy/cost_shear.py:104-139, cost_gravity.py:28-33,
 * cost_sliding_weertman.py:27-36 and utils/gradient/compute_gradient.py:3-16.
 * Fields are read with element strides (fb, fr, fc); `extents` (device,
 * nullable) holds per-batch (ny, nx) valid sizes: cells outside get zero
 * coefficients, which masks them out of the energy.
 * Outputs: coef [nbatch][7][Ny][Nx], hscale [nbatch] = 1/(4 dX[b,0,0]). */
int igm_energy_coeffs_f32(const float* thk, const float* usurf, const float* arrhenius,
                          const float* slidingco, const float* dX, int64_t fb, int64_t fr,
                          int64_t fc, const int32_t* extents, int32_t nbatch, int32_t Ny,
                          int32_t Nx, const IgmPhysics* phys, float* coef, float* hscale,
                          void* stream);
int igm_energy_coeffs_f64(const double* thk, const double* usurf, const double* arrhenius,
                          const double* slidingco, const double* dX, int64_t fb, int64_t fr,
                          int64_t fc, const int32_t* extents, int32_t nbatch, int32_t Ny,
                          int32_t Nx, const IgmPhysics* phys, double* coef, double* hscale,
                          void* stream);

/* One fused pass of the energy over (U, V).
 * mode 0: energy only (iceflow_energy + reduce_mean, energy/energy.py:12-22,
 *         solve/solve.py:57-59), per-block float64 partial sums.
 * mode 1: + gradient dJ/dU, dJ/dV into gU, gV (the GradientTape of
 *         solve/solve.py:77 and emulate/emulate.py:252).
 * mode 2: + legacy Adam update (solve/solve.py:84): reads U, V and the
 *         moments; writes Uout, Vout (must not alias U, V) and the moments in
 *         place.  Planar velocity layout only.  moments: [nbatch][nz][Ny][Nx][4]
 *         = (m_U, v_U, m_V, v_V) per node and level, contiguous (the state of
 *         the reference's Adam slots, one 16-B record per node-level).
 * dzeta: host array of nz-1 layer thicknesses (vert_disc.py:26-35).
 * stop_flag (device, nullable): when *stop_flag != 0 the pass does nothing.
 * partials: igm_energy_workspace_doubles() doubles (per-block sums).
 * prev_partials (device, nullable; nbatch 1): the partials of the previous
 *         iteration's pass (another buffer).  Every block first does what
 *         igm_energy_finalize does for that iteration: COST (solve.py:57-59)
 *         and the EarlyStopping rule (solve.py:112) applied to a copy of
 *         *es_in, and skips this iteration if it says stop; block 0 writes
 *         the new state to *es_out (!= es_in), the energies to out[4] and
 *         the COST to cost_hist[prev_iter].  One launch per solver
 *         iteration; the last iteration's finalize stays a separate call. */
int igm_energy_pass_f32(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                        const float* U, const float* V, int64_t bstride, int64_t kstride,
                        int64_t rstride, int64_t cstride, const float* coef, const float* hscale,
                        const double* dzeta, const IgmPhysics* phys, float* gU, float* gV,
                        float* Uout, float* Vout, float* moments, const IgmAdam* adam,
                        double* partials, const int32_t* stop_flag, const double* prev_partials,
                        const IgmEarlyStop* es_in, IgmEarlyStop* es_out, double* out,
                        double* cost_hist, int32_t prev_iter, void* stream);
int igm_energy_pass_f64(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                        const double* U, const double* V, int64_t bstride, int64_t kstride,
                        int64_t rstride, int64_t cstride, const double* coef,
                        const double* hscale, const double* dzeta, const IgmPhysics* phys,
                        double* gU, double* gV, double* Uout, double* Vout, double* moments,
                        const IgmAdam* adam, double* partials, const int32_t* stop_flag,
                        const double* prev_partials, const IgmEarlyStop* es_in,
                        IgmEarlyStop* es_out, double* out, double* cost_hist, int32_t prev_iter,
                        void* stream);

/* Deterministic float64 reduction of the pass's partial sums:
 * out[b] = {mean shear, mean sliding, mean gravity, COST} per batch entry
 * (tf.reduce_mean + tf.add_n, solve/solve.py:57-59).
 * With `es` (device, nullable; nbatch must be 1): no-op once es->stopped;
 * otherwise stores COST in cost_hist[iter] (device, nullable) and applies
 * EarlyStopping.should_stop (solve/solve.py:112, iceflow/utils.py:80-111). */
int igm_energy_finalize(const double* partials, int32_t nbatch, int32_t Ny, int32_t Nx,
                        double* out, IgmEarlyStop* es, double* cost_hist, int32_t iter,
                        void* stream);

/* Two consecutive iterations of solve_iceflow's loop (solve/solve.py:45-84)
 * in one pass: energy, adjoint and legacy Adam at U (iteration i, step
 * adam1), then again at the updated field (iteration i+1, step adam2), with
 * no EarlyStopping check between the two.  Float32, planar (Nz, Ny, Nx), one
 * field.  The result equals two igm_energy_pass_f32 mode-2 calls.
 * Writes Uout, Vout (must not alias U, V).  V = NULL: U holds (U, V)
 * interleaved, [nz][Ny][Nx][2]; Vout = NULL: Uout is written that way (one
 * 8-B access per node and level instead of two 4-B ones; a solve keeps its
 * intermediate fields so).  Moments [nz][Ny][Nx][4] are read
 * from moments_in and written to moments_out (another buffer: neighbouring
 * waves read rows this pass writes); moments_in = NULL stands for all-zero
 * moments (a fresh optimizer: nothing is read), moments_out = NULL drops the
 * updated moments (the last pass of a solve).  adam1, adam2: same beta1,
 * beta2 and epsilon; alpha may differ.
 * partials: 6 doubles per block, the (shear, sliding, gravity) sums of
 * iteration i then i+1 (igm_energy_workspace_doubles() covers them).
 * prev_partials, es_in, es_out, out, cost_hist, prev_iter: as for
 * igm_energy_pass_f32; prev_count = 1 when the previous pass was an
 * igm_energy_pass_f32 call, 2 when it was an igm_energy_pass2_f32 call (its
 * two iterations prev_iter and prev_iter+1 are accounted in order).
 * nz: 2, 3, 4, 5, 6, 8, 10 or 12. */
int igm_energy_pass2_f32(int32_t nz, int32_t Ny, int32_t Nx, const float* U, const float* V,
                         const float* coef, const float* hscale, const double* dzeta,
                         const IgmPhysics* phys, float* Uout, float* Vout,
                         const float* moments_in, float* moments_out, const IgmAdam* adam1,
                         const IgmAdam* adam2, double* partials, const int32_t* stop_flag,
                         const double* prev_partials, int32_t prev_count,
                         const IgmEarlyStop* es_in, IgmEarlyStop* es_out, double* out,
                         double* cost_hist, int32_t prev_iter, void* stream);

/* One pass of a solve whose iteration count is decided on the device
 * (solve/solve.py:45-84 with the EarlyStopping check of solve.py:112): the
 * pass first accounts the previous pass's iterations (their COSTs into
 * cost_hist, the EarlyStopping rule applied to a copy of *es_in, the new
 * state to *es_out), then runs two iterations when the first of them cannot
 * end the solve (the rule has no best COST yet, or one more non-improving
 * COST leaves wait below patience), one when it can, none once stopped or
 * when iters_done has reached nbitmax.  Results equal one-iteration passes.
 * Launch k (from 0) reads U, V and moments from buffer k % 2 and writes
 * buffer (k + 1) % 2 of each pair, partials to a buffer of its own (two in
 * turn), es_in = state k % 2, es_out = state (k + 1) % 2; launch 0 gets no
 * prev_partials, and moments_in = NULL for zero moments; moments_out = NULL
 * drops the moments, and V or Vout = NULL selects the interleaved layout (as
 * igm_energy_pass2_f32).  A last launch with finalize_only = 1 accounts the final
 * pass; the result is then in buffer passes % 2 (IgmEarlyStop.passes).
 * alpha_tab: device array of the Adam step alpha of iterations 0..nbitmax-1;
 * adam: beta1, beta2 and epsilon.  Float32, planar (Nz, Ny, Nx), one field;
 * nz: 2, 3, 4, 5, 6, 8, 10 or 12. */
int igm_energy_solve_pass_f32(int32_t nz, int32_t Ny, int32_t Nx, const float* U, const float* V,
                              const float* coef, const float* hscale, const double* dzeta,
                              const IgmPhysics* phys, float* Uout, float* Vout,
                              const float* moments_in, float* moments_out, const IgmAdam* adam,
                              const float* alpha_tab, int32_t nbitmax, double* partials,
                              const double* prev_partials, const IgmEarlyStop* es_in,
                              IgmEarlyStop* es_out, double* out, double* cost_hist,
                              int32_t finalize_only, void* stream);

/* Energy (mode 0) or energy and gradient (mode 1) of (U, V), reduced into
 * out[nbatch][4] = {mean shear, mean sliding, mean gravity, COST}: the same
 * kernels and result, bit for bit, as igm_energy_pass_f32 (same mode, no
 * Adam, no early stopping) followed by igm_energy_finalize, but it
 * allocates nothing and takes no workspace.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Its per-block partial sums live in a static 512 KiB device array per GPU
 * (65,536 doubles); a batch that needs more (3 * blocks per entry, about
 * 800 doubles for one 1000 x 1000 field) is rejected with an error.  Calls
 * that overlap on one GPU share that array and race. */
int igm_energy_eval_f32_noalloc(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                                const float* U, const float* V, int64_t bstride, int64_t kstride,
                                int64_t rstride, int64_t cstride, const float* coef,
                                const float* hscale, const double* dzeta, const IgmPhysics* phys,
                                float* gU, float* gV, double* out, void* stream);

/* igm_energy_finalize for the partials of an igm_energy_pass2_f32 call:
 * COST and EarlyStopping (solve.py:57-59, 112) of iterations iter and
 * iter+1, in order; out[4] holds the energies of the second. */
int igm_energy_finalize_pair(const double* partials, int32_t Ny, int32_t Nx, double* out,
                             IgmEarlyStop* es, double* cost_hist, int32_t iter, void* stream);

/* ---------------------------------------------------- velocity outputs -- */

/* where(thk > 0, U, 0) and the derived 2-D fields (solve/solve.py:116-117,
 * 307-309; emulate/emulate.py:143-152; iceflow/utils.py:34-64):
 * Uo, Vo planar (nz, Ny, Nx) (nullable), uvelbase, vvelbase, uvelsurf,
 * vvelsurf, ubar, vbar (Ny, Nx) (each nullable).  Input layout by strides.
 * vert_weight: host array of nz weights (vert_disc.py:43-47). */
int igm_velocity_outputs_f32(int32_t nz, int32_t Ny, int32_t Nx, const float* U, const float* V,
                             int64_t kstride, int64_t rstride, int64_t cstride, const float* thk,
                             const double* vert_weight, float* Uo, float* Vo, float* uvelbase,
                             float* vvelbase, float* uvelsurf, float* vvelsurf, float* ubar,
                             float* vbar, void* stream);
int igm_velocity_outputs_f64(int32_t nz, int32_t Ny, int32_t Nx, const double* U,
                             const double* V, int64_t kstride, int64_t rstride, int64_t cstride,
                             const double* thk, const double* vert_weight, double* Uo, double* Vo,
                             double* uvelbase, double* vvelbase, double* uvelsurf,
                             double* vvelsurf, double* ubar, double* vbar, void* stream);

/* ------------------------------------------------------------ emulator -- */
/* The "cnn" emulator (emulate/neural_network.py:9-66): nb_layers x
 * [Conv2D 3x3 "same" + bias -> LeakyReLU(0.01)] -> Conv2D 1x1 + bias, with
 * 5 input fields and 32 filters.  Activations: NHWC bf16 (B, H, W, 32);
 * output Y: NHWC fp32 (B, H, W, cout).  `extents` (device, nullable): per
 * batch entry (ny, nx); outputs outside are written as zeros, so a padded
 * batch equals per-entry "same" zero padding.  Weights are Keras HWIO fp32
 * masters; MFMA layers take packed bf16 operands (igm_pack_conv_weights). */

/* Layer 1 (Conv2D 3x3, 5 -> 32, neural_network.py:23-34), fp32 arithmetic on
 * the raw fields [thk, usurf, arrhenius, slidingco, dX] (iceflow/utils.py:147-158)
 * read with element strides (fb, fr, fc); bias + LeakyReLU; bf16 output. */
int igm_conv_in_fwd(const float* f0, const float* f1, const float* f2, const float* f3,
                    const float* f4, int64_t fb, int64_t fr, int64_t fc, const int32_t* extents,
                    int32_t B, int32_t H, int32_t W, const float* w_hwio, const float* bias,
                    void* out_bf16, void* stream);

/* Layers 2..nb_layers (Conv2D 3x3, 32 -> 32) + bias + LeakyReLU on bf16 MFMA
 * (fp32 accumulate).  wpacked: [9][32 out][32 in] bf16. */
int igm_conv3x3_fwd_bf16(const void* in, const void* wpacked, const float* bias,
                         const int32_t* extents, int32_t B, int32_t H, int32_t W, void* out,
                         void* stream);

/* Input gradient of a 3x3 layer (the tape through Conv2D, emulate.py:252):
 * dz_below = conv(dz, flipped W) * LeakyReLU'(act_below).
 * wpacked_dgrad: [9][32 in][32 out] bf16 with the taps reversed. */
int igm_conv3x3_dgrad_bf16(const void* dz, const void* wpacked_dgrad, const void* act_below,
                           const int32_t* extents, int32_t B, int32_t H, int32_t W, void* dz_below,
                           void* stream);

/* Output layer Conv2D 1x1, 32 -> cout (<= 32), no activation
 * (neural_network.py:56-64): Y = A W + b, fp32.  wpacked: [32][32] bf16. */
int igm_conv1x1_head_fwd(const void* in, const void* wpacked, const float* bias,
                         const int32_t* extents, int32_t B, int32_t H, int32_t W, int32_t cout,
                         float* Y, void* stream);

/* Backward of the output layer: dz16 = (dY W^T) * LeakyReLU'(a16) (bf16) and
 * dY_bf16 = bf16(dY) zero-padded to 32 channels (input of its wgrad).
 * w_hwio: [32][cout] fp32. */
int igm_conv1x1_head_bwd(const float* dY, const void* a16, const float* w_hwio,
                         const int32_t* extents, int32_t B, int32_t H, int32_t W, int32_t cout,
                         void* dz16, void* dY_bf16, void* stream);

/* Floats of workspace for igm_conv_wgrad_bf16 (taps 9 or 1) or
 * igm_conv_in_wgrad (taps 0) with `nblocks` partial-sum blocks. */
int64_t igm_wgrad_workspace_floats(int32_t taps, int32_t nblocks);

/* Weight and bias gradient of a 3x3 (taps=9) or 1x1 (taps=1) layer on bf16
 * MFMA: dW_hwio [taps][32][cout] and db [cout] (fp32), summed over all
 * pixels of the batch in a fixed order (per-block partials + reduction);
 * accumulate != 0 adds to dW_hwio, db instead of overwriting them. */
int igm_conv_wgrad_bf16(int32_t taps, const void* act, const void* dz, const int32_t* extents,
                        int32_t B, int32_t H, int32_t W, int32_t cout, int32_t nblocks,
                        float* workspace, int32_t accumulate, float* dW_hwio, float* db,
                        void* stream);

/* Weight and bias gradient of layer 1 (fp32): dW_hwio [3][3][5][32], db [32]
 * (accumulate != 0 adds to them). */
int igm_conv_in_wgrad(const float* f0, const float* f1, const float* f2, const float* f3,
                      const float* f4, int64_t fb, int64_t fr, int64_t fc, const int32_t* extents,
                      int32_t B, int32_t H, int32_t W, const void* dz1, int32_t nblocks,
                      float* workspace, int32_t accumulate, float* dW_hwio, float* db,
                      void* stream);

/* The two halves of igm_conv_wgrad_bf16 / igm_conv_in_wgrad, for a batch
 * split into sub-batches (the gradient of one training batch,
 * pretraining/pretraining.py:198-493, over patches of uneven sizes padded per
 * sub-batch): the *_partials calls write (accumulate == 0) or add to
 * (accumulate != 0) the per-block partial sums in `partials`
 * (igm_wgrad_workspace_floats(taps, nblocks) floats, the same nblocks on every
 * call); igm_conv_wgrad_reduce_f32 (taps 9, 1, or 0 for layer 1 with cout 32)
 * then sums them into dW_hwio, db once, in a fixed order, and consumes
 * `partials` (its group sums overwrite rows).  Sub-batch k adds to the
 * partials of sub-batches 0..k-1 in stream order: deterministic. */
int igm_conv_wgrad_partials_bf16(int32_t taps, const void* act, const void* dz,
                                 const int32_t* extents, int32_t B, int32_t H, int32_t W,
                                 int32_t nblocks, float* partials, int32_t accumulate,
                                 void* stream);
int igm_conv_in_wgrad_partials(const float* f0, const float* f1, const float* f2, const float* f3,
                               const float* f4, int64_t fb, int64_t fr, int64_t fc,
                               const int32_t* extents, int32_t B, int32_t H, int32_t W,
                               const void* dz1, int32_t nblocks, float* partials,
                               int32_t accumulate, void* stream);
int igm_conv_wgrad_reduce_f32(int32_t taps, int32_t cout, int32_t nblocks, float* partials,
                              int32_t accumulate, float* dW_hwio, float* db, void* stream);

/* Gradient of the emulator with respect to its 5 input fields (the tape
 * gradient through the network of data_assimilation/optimize/update.py:29-73),
 * through layer 1: dX (fp32 NHWC [B][H][W][5]) from dz1, the bf16 gradient at
 * layer 1's pre-activation output, and the layer-1 kernel w_hwio [3][3][5][32]. */
int igm_conv_in_dgrad(const void* dz1, const float* w_hwio, const int32_t* extents, int32_t B,
                      int32_t H, int32_t W, float* dX, void* stream);

/* Legacy Keras Adam with per-tensor clipnorm on a flat fp32 parameter vector
 * (emulate.py:24-35, 267-269): tensor t spans [offs[t], offs[t+1]);
 * g_t <- g_t * clipnorm / max(||g_t||, clipnorm) (tf.clip_by_norm; clipnorm
 * <= 0 disables it), then the Adam step of IgmAdam.  sq_workspace: ntens floats. */
int igm_adam_clip_f32(float* w, const float* g, float* m, float* v, int64_t n, const int64_t* offs,
                      int32_t ntens, float* sq_workspace, double clipnorm, const IgmAdam* adam,
                      void* stream);

/* bf16 MFMA operands from the fp32 HWIO masters: for each 3x3 layer at
 * w + offs3x3[L] (device int64 array), fwd [L][9][32][32] and dgrad
 * [L][9][32][32]; head [32][32] from w_head [32][cout]. */
int igm_pack_conv_weights(const float* w, const int64_t* offs3x3, int32_t nlayers3x3,
                          const float* w_head, int32_t cout, void* fwd, void* dgrad, void* head,
                          void* stream);

/* ----------------------------------------------------------- transport -- */

/* Vertical velocity from U, V: the default of vert_flow/vert_flow.py:21-56
 * (Lagrange basis, version 2, kinematic; vert_flow_v2.py:17-37, 61-87,
 * 125-147).  U, V planar (nz, ny, nx) float32; topg, thk (ny, nx); levels,
 * vert_weight: host arrays of nz (vert_disc.py:10-13, 43-47); dx the grid
 * spacing.  Writes W (nz, ny, nx) and, when given, W[0] and W[nz-1] into
 * wvelbase, wvelsurf (vert_flow.py:55-56).  ny, nx >= 2. */
int igm_vert_flow_kinematic_f32(int32_t nz, int32_t ny, int32_t nx, const float* U, const float* V,
                                const float* topg, const float* thk, const double* levels,
                                const double* vert_weight, double dx, float* W, float* wvelbase,
                                float* wvelsurf, void* stream);

/* Fused thickness update of the transport step (thk.py:27-54 with
 * compute_divflux_slope_limiter.py:9-65): slope-limited upwind flux
 * divergence of (ubar, vbar) * thk, thk_out = max(thk + dt (smb - div), 0),
 * lsurf = max(topg, -ratio_density thk_out + sealevel), usurf = lsurf + thk_out.
 * All fields (ny, nx) fp32; smb, lsurf, divflux nullable; thk_out must not
 * alias thk.  slope_type: 0 godunov, 1 minmod, 2 superbee. */
int igm_transport_f32(int32_t ny, int32_t nx, const float* ubar, const float* vbar,
                      const float* thk, const float* smb, const float* topg, double dx, double dy,
                      double dt, int32_t slope_type, double ratio_density, double sealevel,
                      float* thk_out, float* lsurf, float* usurf, float* divflux, void* stream);

/* out[0] = max(max|a|, max|b|) over n values (time.py:34-49, the CFL speed).
 * workspace: 1024 floats. */
int igm_absmax2_f32(const float* a, const float* b, int64_t n, float* workspace, float* out,
                    void* stream);

/* smb = clip((usurf - ela) * (usurf < ela ? gradabl : gradacc), -100, maxacc)
 * (smb_simple.py:40-42). */
int igm_smb_simple_f32(int64_t n, const float* usurf, double gradabl, double gradacc, double ela,
                       double maxacc, float* smb, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* IGM_AMD_H */
