/*
 * jne.h -- C ABI of the MI355X johansen-null-eigenspectra engine.
 *
 * The reference (Kuan-Lun/johansen-null-eigenspectra v0.8.0, Rust) has no FFI;
 * its hot path is the crate-internal per-walk operator
 *
 *   pub fn calculate_eigenvalues(dim: usize, steps: usize, seed: u32,
 *                                model: JohansenModel) -> Vec<f64>
 *   (src/johansen_statistics.rs:59-85)
 *
 * called once per seed from the rayon loop in
 * src/data_storage/parallel_compute.rs:14-41.  This ABI is what a Rust
 * `extern "C"` block (or any other FFI) binds to replace that operator with a
 * batched device call.  All buffers are caller-owned; every entry point
 * returns JNE_OK (0) or a negative error code and records a message that
 * jne_last_error() returns.  Nothing aborts the process.  A context owns one
 * HIP device and stream and serialises calls made on it; distinct contexts
 * are independent.
 */
#ifndef JNE_H
#define JNE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JNE_OK 0
#define JNE_ERR_INVALID_ARG (-1)
#define JNE_ERR_HIP (-2)
#define JNE_ERR_UNSUPPORTED (-3)
#define JNE_ERR_NOMEM (-4)
#define JNE_ERR_NUMERIC (-5) /* S11 not positive definite for some walk */
#define JNE_ERR_IO (-6)

/* Accumulation precision of the per-walk moment (Gram) pass.  The eigen
 * solve and the model algebra are FP64 in every mode. */
#define JNE_PREC_FP64 0        /* FP64 VALU accumulation */
#define JNE_PREC_FP32_KAHAN 1  /* compensated FP32 VALU accumulation */
#define JNE_PREC_MFMA_FP16X2 2 /* MFMA, fp16 hi+lo operands, fp32 accumulate */
#define JNE_PREC_MFMA_BF16X2 3 /* MFMA, bf16 hi+lo operands, fp32 accumulate */
#define JNE_PREC_MFMA_BF16X3 4 /* MFMA, bf16 hi+mid+lo operands, fp32 accumulate */
#define JNE_PREC_MFMA_FP32 5   /* MFMA, fp32 operands, fp32 accumulate */
/* The MFMA modes keep the state B and the per-64-step segment terms in FP64
 * and put only within-segment corrections on the matrix cores. */

#define JNE_MAX_DIM 12

typedef struct jne_ctx jne_ctx;

/* Create a context on HIP device `device`. */
int jne_create(int device, jne_ctx **out);
void jne_destroy(jne_ctx *ctx);

/* Message of the last failed call on ctx (empty string if none).  With
 * ctx == NULL, the message of the last failed jne_create on this thread. */
const char *jne_last_error(const jne_ctx *ctx);

/* Eigenvalues per walk: dim+1 for models 1 and 3, else dim.
 * Replaces the eigen-count rule of src/data_storage/thread_manager.rs:40-44
 * and JohansenModel ids of src/johansen_models.rs:31-51.
 * Returns a negative code for an invalid model or dim. */
int jne_num_eigenvalues(int model, int dim);

/* Hot path: eigenvalues of n walks given by seed.
 * Replaces calculate_eigenvalues (src/johansen_statistics.rs:59-85) applied
 * to each of seeds[0..n) (src/data_storage/parallel_compute.rs:32-33).
 *   model      0..4 (src/johansen_models.rs:6-26)
 *   dim        1..JNE_MAX_DIM
 *   steps      T >= 2 (model 4 needs T >= 2; the reference panics at T = 1)
 *   rng_stream Philox key word; 0 for the CLI/.dat path, a request's own seed
 *              when serving (walk `seed` of stream s is independent of stream s')
 *   seeds      host array of n walk seeds
 *   out        host array of n * jne_num_eigenvalues(model, dim) doubles,
 *              row-major; row w holds |rho| of walk seeds[w], sorted
 *              descending (src/johansen_statistics.rs:40-45).
 * Returns JNE_ERR_NUMERIC if any walk's S11 was not positive definite (its
 * row is then NaN). */
int jne_eigenvalues_by_seed(jne_ctx *ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
                            const uint32_t *seeds, size_t n, int precision, double *out);

/* Same with device pointers (d_seeds, d_out in device memory of ctx's device);
 * the work is enqueued on `hip_stream` (a hipStream_t; NULL = the context's
 * stream) and the call returns without synchronising.  Non-positive-definite
 * walks yield NaN rows; no host-side check is made. */
int jne_eigenvalues_by_seed_device(jne_ctx *ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
                                   const uint32_t *d_seeds, size_t n, int precision, double *d_out,
                                   void *hip_stream);

/* Parity harness: eigenvalues from caller-supplied N(0,1) increments.
 * z is a host array of n * steps * dim doubles, z[(w*steps + t)*dim + i]
 * (component fastest, like the reference's column-major dim x T draw,
 * src/rng_matrix.rs:11-37); the engine applies the reference's scaling by
 * sqrt(1/steps), cumulative sum and differencing (src/rng_matrix.rs:135-140,
 * src/johansen_statistics.rs:80-82).  out as in jne_eigenvalues_by_seed. */
int jne_eigenvalues_from_normals(jne_ctx *ctx, int model, int dim, uint32_t steps, const double *z, size_t n,
                                 int precision, double *out);

/* The N(0,1) draws the engine uses for walks seeds[0..n): out is a host array
 * of n * steps * dim floats in the layout of jne_eigenvalues_from_normals.
 * Replaces gen_normal_matrix (src/rng_matrix.rs:11-37). */
int jne_walk_normals(jne_ctx *ctx, int dim, uint32_t steps, uint32_t rng_stream, const uint32_t *seeds, size_t n,
                     float *out);

/* Block until all work enqueued on the context's stream has finished. */
int jne_synchronize(jne_ctx *ctx);

/* Kernel timing (no counterpart in the reference, which has no profiling;
 * used by bench.py for the roofline figure).  While enabled, HIP events are
 * recorded on the launch stream around every launch of the walk kernels.
 * jne_profile_enable clears earlier records; jne_profile_read synchronises
 * on the recorded events and returns the summed duration (ms), the number of
 * launches and the number of walks they processed. */
#define JNE_KERNEL_MOMENTS 0 /* one-pass moment (Gram) accumulation */
#define JNE_KERNEL_EIGEN 1   /* model algebra + eigen solve */
int jne_profile_enable(jne_ctx *ctx, int enable);
int jne_profile_read(jne_ctx *ctx, int kernel, double *total_ms, uint64_t *launches, uint64_t *walks);

#ifdef __cplusplus
}
#endif
#endif /* JNE_H */
