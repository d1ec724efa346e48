/* ot_tune.h — tile-configuration tuning for the overlapped time-tiled stencil
 * engine: the configuration space, the exhaustive autotuner, the analytic
 * performance model and the learned (MLP) cost model.
 *
 * The reference tunes by running one driver process per configuration and
 * scraping its stdout (scripts/utils.py:36-244 run_experiment,
 * scripts/benchrunner/benchmarks.py:7-30 run_benchmark), then ranks
 * configurations with an analytic model (scripts/run-model.py:9-623) or a
 * WEKA IBk regressor (scripts/run-weka.py:74-149). Here the same steps are
 * in-process calls:
 *   ot_tune_space       — the Cartesian sweep of (T, block, ept) of a sweep YAML
 *                         (e.g. scripts/c2050/j3d.yml:7-13), filtered by the
 *                         generator's consistency checks (computeDerived)
 *   ot_tune_measure     — one driver run's "EventElapsed" (median of k)
 *   ot_tune_exhaustive  — run_experiment: every configuration, a resumable CSV
 *                         with the driver's stdout keys as columns
 *   ot_model_analytic   — run-model.py's per-configuration estimate
 *   ot_model_*          — the learned replacement: 28 features -> 3 x 64 ReLU
 *                         -> log(seconds per point-step), trained and evaluated
 *                         by HIP kernels.
 */
#ifndef OT_OT_TUNE_H
#define OT_OT_TUNE_H

#include "ot/ot.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- configuration space and measurement ------------------------------------ */

/* Every tile configuration (time_tile, block_x/y/z, elems_per_thread) of the
 * engine's compiled kernels that ot_derive accepts for base's stencil, extents,
 * time steps and geometry. Writes min(capacity, total) entries to out (may be
 * NULL with capacity 0) and the total to count.
 * Replaces: the YAML sweep expanded by scripts/utils.py:36-244 and the
 * drivers' consistency checks in computeDerived (jacobi-2d.cpp:21-127). */
ot_status ot_tune_space(const ot_params* base, ot_params* out, int32_t capacity, int32_t* count);

/* Median over reps of the HIP-event time of one whole solve (eager launches on
 * the default stream) after warmup untimed solves, on caller-owned device
 * fields in p's padded layout. registers (may be NULL) receives the plan's
 * "Register Usage". A first solve longer than time_limit_s (0 = no limit)
 * stops the measurement: median_s is that solve's time and the status is
 * OT_OK with *timed_out = 1.
 * Replaces: one driver invocation's EventElapsed (jacobi-2d.cpp:903-952) under
 * the reference's per-run watchdog (scripts/utils.py:104-132). */
ot_status ot_tune_measure(const ot_params* p, float* const* rw_a, float* const* rw_b, const float* const* ro,
                          int32_t warmup, int32_t reps, double time_limit_s, double* median_s,
                          int32_t* registers, int32_t* timed_out);

/* Exhaustive autotune over ot_tune_space(base): allocates and seeds the fields
 * (uniform [0,1)), measures every configuration, appends one CSV row per
 * configuration to csv_path (header written when the file is new) and skips
 * configurations already present in the file (resume). best / best_s receive
 * the fastest configuration over the whole file; measured the number of
 * configurations timed by this call.
 * Replaces: scripts/utils.py:36-244 run_experiment + scripts/benchrunner. */
ot_status ot_tune_exhaustive(const ot_params* base, const char* csv_path, int32_t warmup, int32_t reps,
                             double time_limit_s, ot_params* best, double* best_s, int32_t* measured);

/* ---- analytic model (the method without the learned model) ------------------ */

/* Machine constants of the analytic model; the fields follow the
 * architecture files the reference feeds to run-model.py
 * (scripts/c2050/c2050.arch, scripts/gtx280/gt200.arch). */
typedef struct ot_arch {
  double total_lds;          /* LDS bytes per CU                      (total_shared) */
  double max_waves_per_cu;   /*                                       (max_warps_per_sm) */
  double max_threads_per_cu; /*                                       (max_threads_per_sm) */
  double max_blocks_per_cu;  /*                                       (max_blocks_per_sm) */
  double max_regs_per_cu;    /* VGPR lanes per CU                     (max_regs_per_sm) */
  double num_cu;             /*                                       (num_sm) */
  double clock_hz;           /*                                       (clock) */
  double c_load, c_op;       /* CU issue cycles per wave memory / ALU instruction */
  double b_gmem, b_smem;     /* CU cycles per wave access at the HBM / LDS bandwidth */
  double l_gmem, l_smem;     /* latencies in cycles */
  double launch_penalty;     /* cycles per kernel launch */
  double wave_size;          /*                                       (warp_size) */
  double max_regs_per_thread;
  double l1_bytes;           /* vector L1 capacity used by the cache model */
  double sfu_cycles;         /* issue cycles of one transcendental wave instruction */
} ot_arch;

/* MI355X (gfx950) constants. */
ot_status ot_arch_default(ot_arch* arch);

/* Estimated seconds for the whole solve of p with the given "Register Usage".
 * Restates scripts/run-model.py:9-58 (occupancy) and :198-573 (cache model,
 * global / on-chip phase times, stages, launches). */
ot_status ot_model_analytic(const ot_params* p, const ot_arch* arch, int32_t registers, double* seconds);

/* ---- learned cost model ------------------------------------------------------- */

#define OT_FEATURES 28   /* one-hot stencil (13: stencils 0..12; cs-3d is all
                            zeros there and placed by its descriptors), log2
                            nx/ny/nz, radius, fields,
                            T, log2 block x / rows / z, log2 ept, wave kernel,
                            dims, flops per point, loads per point, and the
                            natural log of ot_model_analytic's seconds per
                            point-step (registers unknown: 0) */
#define OT_MLP_HIDDEN 64 /* three hidden ReLU layers */

/* Feature vector of one configuration (the input of ot_predict). */
ot_status ot_features(const ot_params* p, float* features);

typedef struct ot_model ot_model;

/* Fresh model with seeded He-uniform weights, zero biases and identity feature
 * normalisation. The parameters live in device memory. */
ot_status ot_model_create(uint64_t seed, ot_model** model);
void ot_model_destroy(ot_model* model);

/* Flat parameter vector: W1[28x64] b1[64] W2[64x64] b2[64] W3[64x64] b3[64]
 * W4[64] b4[1] (row-major, input index first) — 10241 floats. */
ot_status ot_model_param_count(const ot_model* model, int32_t* count);
ot_status ot_model_get_params(const ot_model* model, float* host, int32_t count);
ot_status ot_model_set_params(ot_model* model, const float* host, int32_t count);

/* Feature normalisation x' = (x - mean) * inv_std applied inside the model.
 * ot_model_fit_normalization sets it from the rows of X (n x OT_FEATURES). */
ot_status ot_model_set_normalization(ot_model* model, const float* mean, const float* inv_std);
ot_status ot_model_fit_normalization(ot_model* model, const float* X, int32_t n);

/* Mean-squared-error loss of the model on (X, y) and its gradient with respect
 * to the flat parameters (grad may be NULL). X: n x OT_FEATURES, host. */
ot_status ot_model_loss_grad(ot_model* model, const float* X, const float* y, int32_t n, float* loss,
                             float* grad);

/* Minibatch Adam (beta1 0.9, beta2 0.999, eps 1e-8) on (X, y) for `epochs`
 * passes in seeded shuffled order. final_loss (may be NULL) receives the loss
 * over the whole set after training. */
ot_status ot_model_train(ot_model* model, const float* X, const float* y, int32_t n, int32_t epochs,
                         int32_t batch, float lr, uint64_t seed, float* final_loss);

/* Batched inference: y[i] = model(X[i]). */
ot_status ot_model_predict(ot_model* model, const float* X, int32_t n, float* y);

/* Predicted log(seconds per grid point per time step) of one configuration.
 * The learned counterpart of ot_model_analytic. */
ot_status ot_predict(ot_model* model, const ot_params* p, float* log_time);

/* Parameters + normalisation as little-endian binary. */
ot_status ot_model_save(const ot_model* model, const char* path);
ot_status ot_model_load(const char* path, ot_model** model);

#ifdef __cplusplus
}
#endif

#endif /* OT_OT_TUNE_H */
