/*
 * jne_oracle.c -- CPU oracle (test infrastructure only; see jne_oracle.h).
 *
 * Build with -ffp-contract=off: the normal generator must perform exactly the
 * IEEE operations written here (every fused multiply-add is an explicit
 * fmaf call) so that it is bit-identical to the device generator.
 */
#include "jne_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10.  Multipliers and Weyl key increments from the Random123     */
/* paper; one round maps (c0,c1,c2,c3) with key (k0,k1) to                    */
/*   (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0)).                */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void jo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Normal generator.  Coefficients from derive_rng_coeffs.py (Chebyshev       */
/* interpolants rounded to float32).                                          */
/* ------------------------------------------------------------------------ */
static const float kQ[10] = {-0x1p-1f,        0x1.55554ep-2f,  -0x1.000008p-2f, 0x1.99a13p-3f,
                             -0x1.5559e6p-3f, 0x1.237a56p-3f,  -0x1.fcc1aep-4f, 0x1.e3d33ap-4f,
                             -0x1.d962f2p-4f, 0x1.15425cp-4f};
static const float kS[5] = {0x1.1c5832p+0f, -0x1.d3ba5cp-2f, 0x1.cda0f6p-5f, -0x1.b1de6p-9f,
                            0x1.d53624p-14f};
static const float kC[6] = {0x1.6a09e6p-1f,  -0x1.bea5b6p-1f,  0x1.6f5a4ap-3f,
                            -0x1.e36a9ep-7f, 0x1.54c234p-11f, -0x1.26d016p-16f};
static const float kLn2Hi = 0x1.62e43p-1f;
static const float kLn2Lo = -0x1.05c61p-29f;

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ln(u) for u in [2^-33, 1] using only correctly rounded operations. */
static float ln_poly(float u) {
    uint32_t bits = f2u(u);
    int e = (int)(bits >> 23) - 127;
    uint32_t mbits = (bits & 0x007FFFFFu) | 0x3F800000u;
    if (mbits > 0x3FB504F3u) { /* m > sqrt(2): use m/2 */
        mbits -= 0x00800000u;
        e += 1;
    }
    float f = u2f(mbits) - 1.0f; /* exact */
    float q = kQ[9];
    for (int i = 8; i >= 0; --i) q = fmaf(q, f, kQ[i]);
    float f2 = f * f;
    float lnm = fmaf(f2, q, f);
    float ef = (float)e;
    return fmaf(ef, kLn2Hi, fmaf(ef, kLn2Lo, lnm));
}

void jo_normal_pair(uint32_t wa, uint32_t wb, float out[2]) {
    /* radius: u1 in (0, 1] */
    float u1 = fmaf((float)wa, 0x1p-32f, 0x1p-33f);
    float r = sqrtf(-2.0f * ln_poly(u1));
    /* angle: top two bits pick the quadrant, the low 30 bits the position g
     * in [-1/2, 1/2) around the quadrant's centre */
    uint32_t q = wb >> 30;
    int32_t gi = (int32_t)(wb & 0x3FFFFFFFu) - (int32_t)0x20000000;
    float g = (float)gi * 0x1p-30f;
    float x = g * g;
    float sp = kS[4];
    for (int i = 3; i >= 0; --i) sp = fmaf(sp, x, kS[i]);
    float cp = kC[5];
    for (int i = 4; i >= 0; --i) cp = fmaf(cp, x, kC[i]);
    float s = g * sp;           /* sin(pi g/2)/sqrt2 */
    float cq = cp - s;          /* cos(pi/4 + pi g/2) */
    float sq = cp + s;          /* sin(pi/4 + pi g/2) */
    float a = (q & 1u) ? sq : cq;
    float b = (q & 1u) ? cq : sq;
    uint32_t cs = ((q ^ (q >> 1)) & 1u) << 31; /* cos < 0 in quadrants 1, 2 */
    uint32_t ss = (q >> 1) << 31;              /* sin < 0 in quadrants 2, 3 */
    out[0] = u2f(f2u(r * a) ^ cs);
    out[1] = u2f(f2u(r * b) ^ ss);
}

/* Counter map: normal n = t*d + i of walk `seed` is output n%4 of the Philox
 * call with counter (n/4 low, n/4 high, seed, 0) and key (stream, JNE_KEY1). */
#define JNE_KEY1 0x4A4E4531u

void jo_walk_normals(uint32_t seed, uint32_t stream, int d, uint32_t T, float *z) {
    uint64_t total = (uint64_t)T * (uint64_t)d;
    uint64_t calls = (total + 3) / 4;
    const uint32_t key[2] = {stream, JNE_KEY1};
    for (uint64_t q = 0; q < calls; ++q) {
        uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), seed, 0u};
        uint32_t w[4];
        jo_philox4x32_10(ctr, key, w);
        float p0[2], p1[2];
        jo_normal_pair(w[0], w[1], p0);
        jo_normal_pair(w[2], w[3], p1);
        float v[4] = {p0[0], p0[1], p1[0], p1[1]};
        for (int j = 0; j < 4; ++j) {
            uint64_t n = 4 * q + (uint64_t)j;
            if (n < total) z[n] = v[j];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Literal restatement of the reference walk.                                  */
/* ------------------------------------------------------------------------ */
int jo_num_eigenvalues(int model, int d) { return (model == 1 || model == 3) ? d + 1 : d; }

void jo_brownian(int d, uint32_t T, const double *z, double *B) {
    /* src/rng_matrix.rs:135-140: delta_t = 1/steps, sqrt_dt = delta_t.sqrt(),
     * scaled = [start | z] * sqrt_dt, then cumulative sum along each row. */
    const double delta_t = 1.0 / (double)T;
    const double sqrt_dt = sqrt(delta_t);
    for (int i = 0; i < d; ++i) {
        double acc = 0.0;
        acc += 0.0 * sqrt_dt; /* start column (zeros) */
        B[i] = acc;
        for (uint32_t t = 0; t < T; ++t) {
            acc += z[(size_t)t * d + i] * sqrt_dt;
            B[(size_t)(t + 1) * d + i] = acc;
        }
    }
}

static double row_mean(const double *M, int rows, size_t cols, int r) {
    double s = 0.0;
    for (size_t c = 0; c < cols; ++c) s += M[c * rows + r];
    return s / (double)cols;
}

int jo_construct_f(int model, int d, uint32_t T, const double *bp, double *F) {
    const size_t cols = T;
    const double t = (double)cols;
    switch (model) {
    case 0: /* johansen_statistics.rs:106 */
        memcpy(F, bp, sizeof(double) * (size_t)d * cols);
        return 0;
    case 1: { /* :108-113 */
        const int nf = d + 1;
        for (size_t c = 0; c < cols; ++c) {
            for (int r = 0; r < d; ++r) F[c * nf + r] = bp[c * d + r];
            F[c * nf + d] = 1.0;
        }
        return 0;
    }
    case 2: { /* :115-138: rows 0..d-2 demeaned, last row (i+1)/T - 0.5 */
        const int nf = d;
        for (int r = 0; r < d - 1; ++r) {
            double mean = row_mean(bp, d, cols, r);
            for (size_t c = 0; c < cols; ++c) F[c * nf + r] = bp[c * d + r] - mean;
        }
        for (size_t c = 0; c < cols; ++c) F[c * nf + (d - 1)] = (double)(c + 1) / t - 0.5;
        return 0;
    }
    case 3: { /* :140-162: all rows demeaned, plus row (i+1)/T - 0.5 */
        const int nf = d + 1;
        for (int r = 0; r < d; ++r) {
            double mean = row_mean(bp, d, cols, r);
            for (size_t c = 0; c < cols; ++c) F[c * nf + r] = bp[c * d + r] - mean;
        }
        for (size_t c = 0; c < cols; ++c) F[c * nf + d] = (double)(c + 1) / t - 0.5;
        return 0;
    }
    case 4: { /* :164-195: X = [B rows 0..d-2; y^2], Z = [1; y], F = X - X Z'(ZZ')^-1 Z */
        const int nf = d;
        double *y = (double *)malloc(sizeof(double) * cols);
        if (!y) return -2;
        for (size_t c = 0; c < cols; ++c) y[c] = (double)(c + 1) / t;
        for (size_t c = 0; c < cols; ++c) {
            for (int r = 0; r < d - 1; ++r) F[c * nf + r] = bp[c * d + r];
            F[c * nf + (d - 1)] = y[c] * y[c]; /* powi(2) */
        }
        /* ZZ' (2x2) */
        double z00 = 0, z01 = 0, z11 = 0;
        for (size_t c = 0; c < cols; ++c) {
            z00 += 1.0 * 1.0;
            z01 += 1.0 * y[c];
            z11 += y[c] * y[c];
        }
        double z10 = z01;
        double det = z00 * z11 - z10 * z01;
        if (det == 0.0) { free(y); return -3; }
        double i00 = z11 / det, i01 = -z01 / det, i10 = -z10 / det, i11 = z00 / det;
        /* P = (X Z') inv Z, computed left to right; F = X - P */
        for (int r = 0; r < nf; ++r) {
            double xz0 = 0, xz1 = 0;
            for (size_t c = 0; c < cols; ++c) {
                xz0 += F[c * nf + r] * 1.0;
                xz1 += F[c * nf + r] * y[c];
            }
            double w0 = xz0 * i00 + xz1 * i10;
            double w1 = xz0 * i01 + xz1 * i11;
            for (size_t c = 0; c < cols; ++c) F[c * nf + r] -= w0 * 1.0 + w1 * y[c];
        }
        free(y);
        return 0;
    }
    default:
        return -1;
    }
}

void jo_sum_outer(int ra, int rb, size_t n, const double *a, const double *b, double *out) {
    for (int j = 0; j < rb; ++j)
        for (int i = 0; i < ra; ++i) out[(size_t)j * ra + i] = 0.0;
    for (size_t s = 0; s < n; ++s) {
        const double *ac = a + s * ra, *bc = b + s * rb;
        for (int j = 0; j < rb; ++j)
            for (int i = 0; i < ra; ++i) out[(size_t)j * ra + i] += ac[i] * bc[j];
    }
}

/* Cyclic Jacobi eigenvalues of a symmetric n x n matrix (column-major, in
 * place); eigenvalues to w. */
static void jacobi_eigvals(int n, double *a, double *w) {
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int q = 0; q < n; ++q) {
            dia += a[q * n + q] * a[q * n + q];
            for (int p = 0; p < q; ++p) off += a[q * n + p] * a[q * n + p];
        }
        if (off == 0.0 || off <= 1e-36 * dia) break;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double apq = a[q * n + p];
                if (apq == 0.0) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double tt = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) tt = -tt;
                double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
                for (int k = 0; k < n; ++k) { /* columns p, q */
                    double akp = a[p * n + k], akq = a[q * n + k];
                    a[p * n + k] = c * akp - s * akq;
                    a[q * n + k] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* rows p, q */
                    double apk = a[k * n + p], aqk = a[k * n + q];
                    a[k * n + p] = c * apk - s * aqk;
                    a[k * n + q] = s * apk + c * aqk;
                }
            }
        }
    }
    for (int i = 0; i < n; ++i) w[i] = a[i * n + i];
}

static int cholesky(int n, double *L /* in: B, out: lower factor, col-major */) {
    for (int j = 0; j < n; ++j) {
        double s = L[j * n + j];
        for (int k = 0; k < j; ++k) s -= L[k * n + j] * L[k * n + j];
        if (!(s > 0.0)) return -1;
        double ljj = sqrt(s);
        L[j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double v = L[j * n + i];
            for (int k = 0; k < j; ++k) v -= L[k * n + i] * L[k * n + j];
            L[j * n + i] = v / ljj;
        }
        for (int i = 0; i < j; ++i) L[j * n + i] = 0.0;
    }
    return 0;
}

static int cmp_desc(const void *x, const void *y) {
    double a = *(const double *)x, b = *(const double *)y;
    return (a < b) - (a > b);
}

int jo_gen_eig_sym(int n, const double *A, const double *B, double *out) {
    size_t nn = (size_t)n * n;
    double *L = (double *)malloc(sizeof(double) * nn * 2);
    if (!L) return -2;
    double *C = L + nn;
    memcpy(L, B, sizeof(double) * nn);
    if (cholesky(n, L) != 0) { free(L); return -1; }
    /* C = L^-1 A L^-T: first X = L^-1 A (columns), then C = L^-1 X^T */
    memcpy(C, A, sizeof(double) * nn);
    for (int pass = 0; pass < 2; ++pass) {
        for (int col = 0; col < n; ++col) {
            double *x = C + (size_t)col * n;
            for (int i = 0; i < n; ++i) {
                double v = x[i];
                for (int k = 0; k < i; ++k) v -= L[k * n + i] * x[k];
                x[i] = v / L[i * n + i];
            }
        }
        /* transpose */
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double tmp = C[j * n + i];
                C[j * n + i] = C[i * n + j];
                C[i * n + j] = tmp;
            }
    }
    /* symmetrize against rounding */
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double v = 0.5 * (C[j * n + i] + C[i * n + j]);
            C[j * n + i] = C[i * n + j] = v;
        }
    jacobi_eigvals(n, C, out);
    for (int i = 0; i < n; ++i) out[i] = fabs(out[i]);
    qsort(out, (size_t)n, sizeof(double), cmp_desc);
    free(L);
    return 0;
}

int jo_walk_moments(int model, int d, uint32_t T, const double *z, double *s01raw, double *s11) {
    if (model < 0 || model > 4 || d < 1 || T < 1) return -1;
    const int nf = jo_num_eigenvalues(model, d);
    const size_t cols = T;
    double *B = (double *)malloc(sizeof(double) * (size_t)d * (cols + 1));
    double *dB = (double *)malloc(sizeof(double) * (size_t)d * cols);
    double *F = (double *)malloc(sizeof(double) * (size_t)nf * cols);
    if (!B || !dB || !F) { free(B); free(dB); free(F); return -2; }
    jo_brownian(d, T, z, B);
    /* johansen_statistics.rs:80-82: bm_previous = cols 0..T-1,
     * dbm = cols 1..T - cols 0..T-1 */
    for (size_t c = 0; c < cols; ++c)
        for (int i = 0; i < d; ++i) dB[c * d + i] = B[(c + 1) * d + i] - B[c * d + i];
    int rc = jo_construct_f(model, d, T, B, F); /* B's first T columns are bm_previous */
    if (rc == 0) {
        /* johansen_statistics.rs:32-33 */
        jo_sum_outer(d, nf, cols, dB, F, s01raw);
        jo_sum_outer(nf, nf, cols, F, F, s11);
        const double delta_t = 1.0 / (double)T;
        for (int k = 0; k < nf * nf; ++k) s11[k] *= delta_t;
    }
    free(B); free(dB); free(F);
    return rc;
}

int jo_walk_from_normals(int model, int d, uint32_t T, const double *z, double *out) {
    if (model < 0 || model > 4 || d < 1 || T < 1) return -1;
    const int nf = jo_num_eigenvalues(model, d);
    double *s01 = (double *)malloc(sizeof(double) * (size_t)d * nf);
    double *s11 = (double *)malloc(sizeof(double) * (size_t)nf * nf);
    double *A = (double *)malloc(sizeof(double) * (size_t)nf * nf);
    if (!s01 || !s11 || !A) { free(s01); free(s11); free(A); return -2; }
    int rc = jo_walk_moments(model, d, T, z, s01, s11);
    if (rc == 0) {
        /* A = S01raw^T S01raw (johansen_statistics.rs:36) */
        for (int j = 0; j < nf; ++j)
            for (int i = 0; i < nf; ++i) {
                double v = 0.0;
                for (int k = 0; k < d; ++k) v += s01[(size_t)i * d + k] * s01[(size_t)j * d + k];
                A[(size_t)j * nf + i] = v;
            }
        rc = jo_gen_eig_sym(nf, A, s11, out);
    }
    free(s01); free(s11); free(A);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Multi-threaded by-seed driver (CPU baseline; one walk per task, like the   */
/* reference's rayon loop in parallel_compute.rs:14-41).                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int model, d;
    uint32_t T, stream;
    const uint32_t *seeds;
    size_t n;
    double *out;
    size_t next;
    pthread_mutex_t mu;
    int rc;
} by_seed_job;

static void *by_seed_worker(void *arg) {
    by_seed_job *job = (by_seed_job *)arg;
    const size_t total = (size_t)job->T * job->d;
    const int nf = jo_num_eigenvalues(job->model, job->d);
    float *zf = (float *)malloc(sizeof(float) * total);
    double *zd = (double *)malloc(sizeof(double) * total);
    if (!zf || !zd) {
        free(zf); free(zd);
        pthread_mutex_lock(&job->mu);
        job->rc = -2;
        pthread_mutex_unlock(&job->mu);
        return NULL;
    }
    for (;;) {
        pthread_mutex_lock(&job->mu);
        size_t w = job->next++;
        pthread_mutex_unlock(&job->mu);
        if (w >= job->n) break;
        jo_walk_normals(job->seeds[w], job->stream, job->d, job->T, zf);
        for (size_t k = 0; k < total; ++k) zd[k] = (double)zf[k];
        int rc = jo_walk_from_normals(job->model, job->d, job->T, zd, job->out + w * nf);
        if (rc != 0) {
            pthread_mutex_lock(&job->mu);
            job->rc = rc;
            pthread_mutex_unlock(&job->mu);
        }
    }
    free(zf); free(zd);
    return NULL;
}

int jo_walks_by_seed(int model, int d, uint32_t T, uint32_t stream, const uint32_t *seeds,
                     size_t n, int threads, double *out) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    by_seed_job job = {model, d, T, stream, seeds, n, out, 0, PTHREAD_MUTEX_INITIALIZER, 0};
    pthread_t tid[256];
    for (int i = 0; i < threads; ++i) pthread_create(&tid[i], NULL, by_seed_worker, &job);
    for (int i = 0; i < threads; ++i) pthread_join(tid[i], NULL);
    return job.rc;
}

double jo_percentile_sorted(const double *v, size_t n, double p) {
    if (n == 0) return NAN;
    double rank = p * (double)(n - 1);
    size_t lo = (size_t)floor(rank), hi = (size_t)ceil(rank);
    if (lo == hi) return v[lo];
    double w = rank - (double)lo;
    return v[lo] * (1.0 - w) + v[hi] * w;
}
