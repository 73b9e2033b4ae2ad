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

/* Legacy-Keras Adam step constants (solve/solve.py:132-141, 205).
 * alpha = lr * sqrt(1 - beta2^t) / (1 - beta1^t), computed by the caller. */
typedef struct IgmAdam {
  double alpha;
  double beta1;
  double beta2;
  double epsilon;
} IgmAdam;

/* Device-side EarlyStopping state (iceflow/utils.py:174-204). */
typedef struct IgmEarlyStop {
  double best_loss;
  double relative_min_delta;
  int32_t has_best;
  int32_t wait;
  int32_t patience;
  int32_t stopped;     /* set to 1 when should_stop() returned True */
  int32_t stop_iter;   /* iteration index at which it stopped */
  int32_t iters_done;  /* number of iterations whose update was applied */
} IgmEarlyStop;

const char* igm_last_error(void);

/* ---------------------------------------------------------------- energy -- */

/* Number of double words of the `partials` workspace that
 * igm_energy_pass_* needs for this problem size. */
int64_t igm_energy_workspace_doubles(int32_t nbatch, int32_t Ny, int32_t Nx);

/* Per-cell, velocity-independent coefficients of the staggered energy:
 * stag4h of thk, B = 2 A^(-1/n) and slidingco, the staggered slopes of usurf
 * and of the bed (usurf - thk) from the dX field, the floating damping and
 * the 1/S mean weights.  Replaces the U-independent half of
 * energy/cost_shear.py:104-139, cost_gravity.py:205-216,
 * cost_sliding_weertman.py:273-285 and utils/gradient/compute_gradient.py:3-16.
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
 *         solve/solve.py:178-180), per-block float64 partial sums.
 * mode 1: + gradient dJ/dU, dJ/dV into gU, gV (the GradientTape of
 *         solve/solve.py:198 and emulate/emulate.py:252).
 * mode 2: + legacy Adam update (solve/solve.py:205): reads U, V, m, v; writes
 *         Uout, Vout (must not alias U, V) and m, v in place.
 * dzeta: host array of nz-1 layer thicknesses (vert_disc.py:26-35).
 * stop_flag (device, nullable): when *stop_flag != 0 the pass does nothing. */
int igm_energy_pass_f32(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                        const float* U, const float* V, int64_t bstride, int64_t kstride,
                        int64_t rstride, int64_t cstride, const float* coef, const float* hscale,
                        const double* dzeta, const IgmPhysics* phys, float* gU, float* gV,
                        float* Uout, float* Vout, float* mU, float* vU, float* mV, float* vV,
                        const IgmAdam* adam, double* partials, const int32_t* stop_flag,
                        void* stream);
int igm_energy_pass_f64(int32_t mode, int32_t nz, int32_t nbatch, int32_t Ny, int32_t Nx,
                        const double* U, const double* V, int64_t bstride, int64_t kstride,
                        int64_t rstride, int64_t cstride, const double* coef,
                        const double* hscale, const double* dzeta, const IgmPhysics* phys,
                        double* gU, double* gV, double* Uout, double* Vout, double* mU,
                        double* vU, double* mV, double* vV, const IgmAdam* adam,
                        double* partials, const int32_t* stop_flag, void* stream);

/* Deterministic float64 reduction of the pass's partial sums:
 * out[b] = {mean shear, mean sliding, mean gravity, COST} per batch entry
 * (tf.reduce_mean + tf.add_n, solve/solve.py:178-180).
 * With `es` (device, nullable; nbatch must be 1): no-op once es->stopped;
 * otherwise stores COST in cost_hist[iter] (device, nullable) and applies
 * EarlyStopping.should_stop (solve/solve.py:233, iceflow/utils.py:186-204). */
int igm_energy_finalize(const double* partials, int32_t nbatch, int32_t Ny, int32_t Nx,
                        double* out, IgmEarlyStop* es, double* cost_hist, int32_t iter,
                        void* stream);

/* ---------------------------------------------------- velocity outputs -- */

/* where(thk > 0, U, 0) and the derived 2-D fields (solve/solve.py:237-238,
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

#ifdef __cplusplus
}
#endif

#endif /* IGM_AMD_H */
