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
#define JNE_PREC_AUTO 6        /* per (dim, steps), the faster of FP64 and MFMA_FP16X2 */
/* The MFMA modes keep the state B and the per-64-step segment terms in FP64
 * and put only within-segment corrections on the matrix cores.  Both paths
 * draw the same normals, whose generation bounds them both; measured on
 * MI355X the MFMA pass is faster for dim 12 from 200 steps, dims 10-11 from
 * 300, dim 9 from 500, dim 8 from 1,000 and dim 7 from 3,000 steps, so
 * JNE_PREC_AUTO runs MFMA_FP16X2
 * there and FP64 elsewhere (FP64 results meet the mixed-path tolerance
 * trivially). */

/* Dims 1..JNE_MAX_DIM run on the device.  The reduced-precision moment
 * passes (the MFMA modes and JNE_PREC_FP32_KAHAN) cover dims
 * 1..JNE_MAX_MIXED_DIM; above that every call runs the FP64 path
 * (JNE_PREC_AUTO included) and a reduced precision is JNE_ERR_UNSUPPORTED. */
#define JNE_MAX_DIM 32
#define JNE_MAX_MIXED_DIM 12

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

/* Stream arguments of the device entry points: a hipStream_t.  NULL is HIP's
 * null (default) stream, which HIP orders with every blocking stream of the
 * device, as in any HIP call; JNE_STREAM_CONTEXT selects the context's own
 * (non-blocking) stream. */
#define JNE_STREAM_CONTEXT ((void *)(intptr_t)-1)

/* Same with device pointers (d_seeds, d_out in device memory of ctx's device);
 * the work is enqueued on `hip_stream` (see JNE_STREAM_CONTEXT) and the call
 * returns without synchronising.  Non-positive-definite walks yield NaN rows;
 * no host-side check is made.  The context's device scratch is shared by its
 * calls: a call on another stream than the previous call first waits (on the
 * device) for the previous call's use of it. */
int jne_eigenvalues_by_seed_device(jne_ctx *ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
                                   const uint32_t *d_seeds, size_t n, int precision, double *d_out,
                                   void *hip_stream);

/* Several models of the same walks: out d_outs[k] (n x jne_num_eigenvalues
 * (models[k], dim) doubles, device memory) gets model models[k]'s eigenvalues
 * of the n walks of d_seeds, k < n_models (1..5).  A walk's moments do not
 * depend on the model, so the moment pass runs once per walk and only the
 * model algebra and eigen solve run per model: the result for each model is
 * bit for bit that of jne_eigenvalues_by_seed_device for it.  The reference
 * recomputes the same walks for every model of its sweep
 * (src/main.rs:85-127 runs each model's seeds 1..runs through
 * calculate_eigenvalues, src/johansen_statistics.rs:59-85); this replaces
 * those repeats.  Stream and scratch as jne_eigenvalues_by_seed_device. */
int jne_eigenvalues_by_seed_models_device(jne_ctx *ctx, const int *models, int n_models, int dim, uint32_t steps,
                                          uint32_t rng_stream, const uint32_t *d_seeds, size_t n, int precision,
                                          double *const *d_outs, void *hip_stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same as jne_eigenvalues_by_seed_device, with the
 * same kernels and bit-identical results, but its per-walk scratch (the
 * moments, 8 (2 dim^2 + 6 dim) bytes per walk, and for dims up to
 * JNE_MAX_MIXED_DIM the M matrices of the eigen solve, 8 (dim (dim + 1) / 2 + 1)
 * bytes) lives in a fixed device array of JNE_NOALLOC_SCRATCH_BYTES inside the
 * library, one per GPU, held until the process ends; the call allocates
 * nothing.  n is at most JNE_NOALLOC_SCRATCH_BYTES / (8 (2.5 dim^2 + 6.5 dim + 1))
 * walks up to dim 12 (19,108 at dim 12, 838,860 at dim 1) and
 * JNE_NOALLOC_SCRATCH_BYTES / (8 (2 dim^2 + 6 dim)) above (3,744 at dim 32);
 * larger n is JNE_ERR_INVALID_ARG with a message.
 * Every call on one GPU shares that GPU's array (whatever the context), and
 * the array keeps the last call's contents: calls that overlap on one GPU
 * (other streams, other host threads, other contexts) race on it, so the
 * caller orders them.  Nothing in the array needs to start at zero.
 * Profiling (jne_profile_enable) does not record these launches. */
#define JNE_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int jne_eigenvalues_by_seed_device_noalloc(jne_ctx *ctx, int model, int dim, uint32_t steps, uint32_t rng_stream,
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

/* Same with device pointers (d_seeds: n seeds, d_out: n * steps * dim floats
 * in device memory of ctx's device), enqueued on `hip_stream` (see
 * JNE_STREAM_CONTEXT) without synchronising; steps * dim < 2^32.  For callers
 * that consume the stream on the device.  Also replaces gen_normal_matrix
 * (src/rng_matrix.rs:11-37). */
int jne_walk_normals_device(jne_ctx *ctx, int dim, uint32_t steps, uint32_t rng_stream, const uint32_t *d_seeds,
                            size_t n, float *d_out, void *hip_stream);

/* Percentiles p[0..np) (1 <= np <= 8, each in [0, 1]) of the device array
 * d_values[0..n), n < 2^31, into the host array out[np]: rank p*(n-1),
 * linear interpolation between the neighbouring order statistics, the same
 * values as jne_percentiles, found on the device by bucket selection (a sort
 * only when ties crowd the selected buckets).  Synchronous on hip_stream
 * (see JNE_STREAM_CONTEXT); n = 0 gives NaN as jne_percentiles.  The
 * device counterpart of get_percentile_value (src/simulation_analyzers.rs:4-18),
 * used by the query server; NaN values are JNE_ERR_NUMERIC. */
int jne_percentiles_device(jne_ctx *ctx, const double *d_values, uint64_t n, const double *p, size_t np,
                           double *out, void *hip_stream);

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
#define JNE_KERNEL_NORMALS 2 /* the random stream alone (jne_walk_normals[_device]) */
int jne_profile_enable(jne_ctx *ctx, int enable);
int jne_profile_read(jne_ctx *ctx, int kernel, double *total_ms, uint64_t *launches, uint64_t *walks);

/* ------------------------------------------------------------------------
 * Storage: the append-only .dat V6 file (DATA_FORMAT.md), its reader, resume
 * and percentiles.  Host-only calls; they do not need a context.  On failure
 * they return a negative code and jne_host_last_error() (per thread) holds
 * the message; the messages of the reference's io::Error values are kept.
 * ------------------------------------------------------------------------ */
#define JNE_ERR_FORMAT (-7) /* file content: magic, parameters or eigenvalue counts */

/* Message of the last failed host call on this thread. */
const char *jne_host_last_error(void);

typedef struct jne_dat_writer jne_dat_writer;

/* Open `path` for appending records of (model, dim, steps).
 * Replaces AppendOnlyWriter::with_expected_size (src/data_storage/writer.rs:27-168)
 * and remove_eof_marker (writer.rs:170-195):
 *   - a new file gets the 18-byte header;
 *   - an existing V6 file with the same model, dim and steps is resumed: its
 *     whole records are kept and the EOF trailer is removed.  Unlike the
 *     reference, an incomplete trailing record (a crash mid-write) is
 *     truncated away too, so appended records stay readable;
 *   - a file whose magic differs is deleted and recreated;
 *   - different parameters: JNE_ERR_FORMAT, "Model mismatch: file has model
 *     1, expected 2" (and the Dimension / Steps variants).
 * expected_records > 0 only prints the reference's pre-allocation notice.
 * quiet = 0 prints the reference's messages on stdout. */
int jne_dat_writer_open(const char *path, int model, int dim, uint32_t steps, uint64_t expected_records, int quiet,
                        jne_dat_writer **out);
/* Records in the file (found when opened plus appended). */
uint64_t jne_dat_writer_count(const jne_dat_writer *w);
/* Print the reference's "Simulation progress: ..." line each time the record
 * count crosses a multiple of 10,000 (src/data_storage/thread_manager.rs:59-83),
 * against total_runs; 0 turns it off. */
int jne_dat_writer_set_progress(jne_dat_writer *w, uint64_t total_runs);
/* Append n records: seeds[i] with eigenvalues[i*nf .. i*nf+nf).
 * Replaces append_eigenvalues (writer.rs:198-253); nf must stay the same for
 * the whole file ("Eigenvalue count mismatch: ...") and be <= 255. */
int jne_dat_writer_append(jne_dat_writer *w, const uint32_t *seeds, const double *eigenvalues, size_t n, int nf);
/* Write the trailer (EOF_MARK, record count, nF), close and free w.
 * Replaces AppendOnlyWriter::finish (writer.rs:256-302). */
int jne_dat_writer_finish(jne_dat_writer *w);
/* Flush, close and free w without a trailer (an interrupted run). */
void jne_dat_writer_abandon(jne_dat_writer *w);

/* Header and record count of a .dat file (trailer fast path, or a scan of
 * whole records when the trailer is missing).  Replaces the header/metadata
 * part of read_append_file (src/data_storage/reader.rs:23-104). */
int jne_dat_read_info(const char *path, int *model, int *dim, uint32_t *steps, uint64_t *records, int *nf,
                      int *has_trailer);
/* All records in file order: seeds[n_read], eigenvalues[n_read * nf] (either
 * may be NULL to skip it).  Replaces read_append_file (reader.rs:23-216);
 * capacity is the number of records the arrays hold. */
int jne_dat_read(const char *path, uint32_t *seeds, double *eigenvalues, uint64_t capacity, uint64_t *n_read,
                 int *nf);
/* Seeds 1..=runs that `path` does not hold yet, ascending, into out[capacity]
 * (out may be NULL to count them); *completed = records in the file.
 * Replaces check_append_progress + get_remaining_seeds
 * (src/data_storage/progress.rs:10-61): a missing or unreadable file counts as
 * empty, different parameters are JNE_ERR_FORMAT. */
int jne_dat_remaining_seeds(const char *path, int model, int dim, uint32_t steps, uint64_t runs, uint32_t *out,
                            uint64_t capacity, uint64_t *n_out, uint64_t *completed);

/* Percentiles p[0..np) of values[0..n): rank p*(n-1), linear interpolation
 * between the neighbouring order statistics.  Replaces get_percentile_value
 * (src/simulation_analyzers.rs:4-18); NaN values are JNE_ERR_NUMERIC. */
int jne_percentiles(const double *values, size_t n, const double *p, size_t np, double *out);

/* Per-record statistic of a .dat file: JNE_STAT_TRACE = sum of eigenvalues,
 * JNE_STAT_MAXEIG = largest (src/simulation_analyzers.rs:21-40), for the
 * records with seed <= runs (runs = 0: all).  out may be NULL to count. */
#define JNE_STAT_TRACE 0
#define JNE_STAT_MAXEIG 1
int jne_dat_statistics(const char *path, uint64_t runs, int statistic, double *out, uint64_t capacity,
                       uint64_t *n_out);

/* ------------------------------------------------------------------------
 * Simulation driver: one (model, dim, steps, runs) job streamed to a .dat
 * file with resume.  Replaces run_model_simulation
 * (src/data_storage/parallel_compute.rs:150-232) with its writer thread
 * (src/data_storage/thread_manager.rs:25-92): the remaining seeds are
 * computed on ctx's device in batches while the previous batch is copied to
 * pinned host memory and encoded/written on a host thread.  quiet = 0 prints
 * the reference's progress and status messages.  A file with different
 * parameters is replaced, as the reference does.  Records are written in
 * ascending seed order within each batch.
 * ------------------------------------------------------------------------ */
int jne_run_simulation(jne_ctx *ctx, int model, int dim, uint32_t steps, uint64_t runs, int precision,
                       const char *path, int quiet);

/* The models[0..n_models) (1..5) of one dim, to paths[k]: the reference's
 * sweep runs them one after another over the same seeds (src/main.rs:109-127,
 * each through run_model_simulation).  With quiet = 0 that is what happens
 * here, message for message.  Quiet, the models whose files need the same
 * seeds (all of them in a fresh sweep) run together: one moment pass per
 * walk and one eigen solve per model (jne_eigenvalues_by_seed_models_device),
 * each file's records the same bytes as its own jne_run_simulation would
 * write.  Incompatible files are replaced as in jne_run_simulation. */
int jne_run_simulation_models(jne_ctx *ctx, const int *models, int n_models, int dim, uint32_t steps, uint64_t runs,
                              int precision, const char *const *paths, int quiet);

/* ------------------------------------------------------------------------
 * Query server.  No counterpart in the reference, which only runs batch
 * jobs; a query returns what its analysis step computes from a finished
 * file: the 0.90, 0.95 and 0.99 quantiles of the trace and max-eigenvalue
 * statistics (src/simulation_analyzers.rs:4-81) over `runs` walks.
 * Walk w (0-based) of a query is walk seed w+1 with Philox key word `seed`,
 * so a query's result depends only on (model, dim, steps, runs, seed,
 * precision): the same alone and in any mix of concurrent queries.
 * Queries are admitted in arrival order when their device footprint
 * (16 bytes per walk) fits the HBM budget; admitted queries share `lanes`
 * launch streams chunk by chunk (continuous batching), a free stream taking
 * the next chunk of the query with the least work left.  The quantiles are
 * selected on the device (bucket selection of the per-walk statistics, as
 * jne_percentiles_device); only the six values come back to the host.
 * ------------------------------------------------------------------------ */
#define JNE_ERR_TIMEOUT (-8)
typedef struct jne_server jne_server;
int jne_server_create(int device, uint64_t hbm_budget_bytes, int lanes, jne_server **out);
/* Stops the scheduler; queries not yet waited for are dropped. */
void jne_server_destroy(jne_server *server);
/* Enqueue a query (runs 1..2^31-1); returns at once with its id. */
int jne_server_submit(jne_server *server, int model, int dim, uint32_t steps, uint64_t runs, uint32_t seed,
                      int precision, uint64_t *id);
/* Wait for query id (timeout_ms < 0: no limit; JNE_ERR_TIMEOUT when it
 * runs out, the query stays pending).  quantiles[6] = trace p90, p95, p99,
 * max-eigenvalue p90, p95, p99; latency_ms = submit to completion.  An id
 * can be waited for once. */
int jne_server_wait(jne_server *server, uint64_t id, int timeout_ms, double *quantiles, double *latency_ms);
/* Walks completed and seconds with at least one launch in flight. */
int jne_server_counters(jne_server *server, uint64_t *walks, double *busy_seconds);

#ifdef __cplusplus
}
#endif
#endif /* JNE_H */
