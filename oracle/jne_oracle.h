/*
 * jne_oracle.h -- CPU oracle for the johansen-null-eigenspectra walk.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, and only as the checker or as the timed CPU baseline.
 *
 * It restates, in plain C, the reference's per-walk algorithm:
 *   src/rng_matrix.rs:57-141          Brownian path (scale by sqrt(1/T), cumsum)
 *   src/johansen_statistics.rs:59-85  walk driver (bm_previous, dbm = diff)
 *   src/johansen_statistics.rs:102-197 construct_f_matrix (models 0..4)
 *   src/matrix_utils.rs:67-85          sum_of_outer_products
 *   src/johansen_statistics.rs:24-47   generalized eigenproblem, |rho| sorted desc
 * and the build's own random stream (Philox4x32-10 + Box-Muller with
 * polynomial transcendentals), which replaces the reference's Xoshiro/Ziggurat
 * stream (src/rng_matrix.rs:11-37).
 *
 * The generalized eigen solve here is Cholesky + cyclic Jacobi (the problem is
 * symmetric-definite); the Python side of the oracle calls LAPACK dggev, the
 * routine the reference calls, and the tests cross-check the two.
 */
#ifndef JNE_ORACLE_H
#define JNE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11). */
void jo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Two Philox words -> two N(0,1) floats (the build's Box-Muller definition). */
void jo_normal_pair(uint32_t wa, uint32_t wb, float out[2]);

/* Normals of one walk: z[t*d + i], t < T, i < d (component fastest, like the
 * reference's column-major dim x T draw, src/rng_matrix.rs:11-37). */
void jo_walk_normals(uint32_t seed, uint32_t stream, int d, uint32_t T, float *z);

/* Number of eigenvalues per walk: dim+1 for models 1 and 3, else dim
 * (src/data_storage/thread_manager.rs:40-44). */
int jo_num_eigenvalues(int model, int d);

/* Brownian path from normals: B is d x (T+1), column-major, B[:,0] = 0.
 * (src/rng_matrix.rs:137-140 + src/matrix_utils.rs:51-63) */
void jo_brownian(int d, uint32_t T, const double *z, double *B);

/* F matrix (nF x T, column-major) from bm_previous (d x T, column-major). */
int jo_construct_f(int model, int d, uint32_t T, const double *bprev, double *F);

/* out (ra x rb, column-major) = sum_i a[:,i] b[:,i]^T, a: ra x n, b: rb x n. */
void jo_sum_outer(int ra, int rb, size_t n, const double *a, const double *b, double *out);

/* Eigenvalues of A v = rho B v (A sym PSD, B sym PD, n x n column-major),
 * returned as |rho| sorted descending.  Returns 0, or -1 if B is not PD. */
int jo_gen_eig_sym(int n, const double *A, const double *B, double *out);

/* One walk from given normals (double, z[t*d+i]) through the literal
 * restatement.  out has jo_num_eigenvalues(model, d) entries. */
int jo_walk_from_normals(int model, int d, uint32_t T, const double *z, double *out);

/* Intermediate matrices of one walk (for the Python oracle's LAPACK path):
 * S01raw (d x nF) and S11 (nF x nF), both column-major. */
int jo_walk_moments(int model, int d, uint32_t T, const double *z, double *s01raw, double *s11);

/* Many walks by seed on `threads` host threads (CPU baseline). */
int jo_walks_by_seed(int model, int d, uint32_t T, uint32_t stream, const uint32_t *seeds,
                     size_t n, int threads, double *out);

/* Percentile of sorted values with linear interpolation at rank p*(n-1)
 * (src/simulation_analyzers.rs:4-18). */
double jo_percentile_sorted(const double *sorted, size_t n, double p);

#ifdef __cplusplus
}
#endif
#endif
