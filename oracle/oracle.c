/*
 * oracle.c -- CPU restatement of the reference DQMC Green's-function path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never loads it.
 * Column-major n x n doubles throughout (src/linalg.h:4).  Dense linear
 * algebra is written out here (unblocked Householder QR with LAPACK's
 * dlarfg sign convention, partial-pivot LU, unit-upper triangular inverse)
 * so the oracle has no LAPACK dependency; it agrees with the reference's
 * LAPACK calls to rounding, which is what the reference's own tests pin
 * (G to 1e-9..1e-12, tests/test_green_stack.c, tests/test_udv*.c).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, n) ((size_t)(i) + (size_t)(j) * (size_t)(n))

/* ------------------------------------------------------------------ RNG */

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* src/rng.c:16-21: four SplitMix64 outputs of the seed. */
void orc_rng_seed(uint64_t s[4], uint64_t seed)
{
    for (int i = 0; i < 4; i++) {
        uint64_t z = (seed += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
}

/* src/rng.c:23-37: xoshiro256** step, top 53 bits as a double in [0,1). */
double orc_rng_double(uint64_t s[4])
{
    const uint64_t out = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

/* One SplitMix64 output of x without state (src/replica.c:9-15). */
static uint64_t splitmix_value(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* src/replica.c:17-34 */
uint64_t orc_replica_seed(uint64_t base, int beta_index, int replica_id)
{
    const uint64_t legacy = base + 1000ULL * (uint64_t)beta_index;
    if (replica_id == 0) return legacy;
    uint64_t x = base;
    x ^= 0xD1B54A32D192ED03ULL * (uint64_t)(unsigned int)(beta_index + 1);
    x ^= 0xABC98388FB8FAC03ULL * (uint64_t)(unsigned int)(replica_id + 1);
    uint64_t m = splitmix_value(x);
    if (m == legacy) m = splitmix_value(m);
    return m;
}

/* -------------------------------------------------------------- lattice */

static void bond(int n, double *t, int i, int j, double v)
{
    t[IDX(i, j, n)] += v;
    t[IDX(j, i, n)] += v;
}

int orc_lattice(int type, int Lx, int Ly, int pbc, double t, double *thop,
                int *bipart)
{
    if (type == 0) { /* src/lattice.c:26-44 */
        const int n = Lx;
        memset(thop, 0, sizeof(double) * (size_t)n * (size_t)n);
        for (int x = 0; x < Lx; x++) {
            if (x + 1 < Lx) bond(n, thop, x, x + 1, t);
            else if (pbc && Lx > 1) bond(n, thop, x, 0, t);
            bipart[x] = (x % 2 == 0) ? 1 : -1;
        }
        return (pbc && Lx > 1 && (Lx % 2 != 0)) ? 0 : 1;
    }
    /* square, src/lattice.c:46-82 */
    const int n = Lx * Ly;
    memset(thop, 0, sizeof(double) * (size_t)n * (size_t)n);
    for (int y = 0; y < Ly; y++) {
        for (int x = 0; x < Lx; x++) {
            const int i = x + y * Lx;
            if (x + 1 < Lx) bond(n, thop, i, (x + 1) + y * Lx, t);
            else if (pbc && Lx > 1) bond(n, thop, i, 0 + y * Lx, t);
            if (y + 1 < Ly) bond(n, thop, i, x + (y + 1) * Lx, t);
            else if (pbc && Ly > 1) bond(n, thop, i, x, t);
            bipart[i] = ((x + y) % 2 == 0) ? 1 : -1;
        }
    }
    const int odd = (pbc && Lx > 1 && (Lx % 2 != 0)) ||
                    (pbc && Ly > 1 && (Ly % 2 != 0));
    return odd ? 0 : 1;
}

/* ---------------------------------------------------------------- model */

void orc_expm_sym(int n, const double *A, double scale, double *out)
{
    /* exp(X), X = scale*A: scale X by 2^-s until ||X||_1 <= 1/8, sum the
       Taylor series to 24 terms in long double, square s times. */
    const size_t nn = (size_t)n * (size_t)n;
    long double *X = malloc(sizeof(long double) * nn);
    long double *S = malloc(sizeof(long double) * nn);
    long double *P = malloc(sizeof(long double) * nn);
    long double *Q = malloc(sizeof(long double) * nn);
    long double norm = 0.0L;
    for (int j = 0; j < n; j++) {
        long double c = 0.0L;
        for (int i = 0; i < n; i++) c += fabsl((long double)scale * A[IDX(i, j, n)]);
        if (c > norm) norm = c;
    }
    int sq = 0;
    while (norm > 0.125L) { norm *= 0.5L; sq++; }
    const long double f = ldexpl((long double)scale, -sq);
    for (size_t k = 0; k < nn; k++) { X[k] = f * A[k]; S[k] = 0.0L; P[k] = 0.0L; }
    for (int i = 0; i < n; i++) { S[IDX(i, i, n)] = 1.0L; P[IDX(i, i, n)] = 1.0L; }
    for (int term = 1; term <= 24; term++) {
        /* P <- P * X / term */
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                long double acc = 0.0L;
                for (int k = 0; k < n; k++) acc += P[IDX(i, k, n)] * X[IDX(k, j, n)];
                Q[IDX(i, j, n)] = acc / (long double)term;
            }
        memcpy(P, Q, sizeof(long double) * nn);
        for (size_t k = 0; k < nn; k++) S[k] += P[k];
    }
    for (int r = 0; r < sq; r++) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                long double acc = 0.0L;
                for (int k = 0; k < n; k++) acc += S[IDX(i, k, n)] * S[IDX(k, j, n)];
                Q[IDX(i, j, n)] = acc;
            }
        memcpy(S, Q, sizeof(long double) * nn);
    }
    /* symmetrize (the exact result is symmetric) */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            out[IDX(i, j, n)] = (double)(0.5L * (S[IDX(i, j, n)] + S[IDX(j, i, n)]));
    free(X); free(S); free(P); free(Q);
}

void orc_model(int n, const double *thop, double U, double dtau,
               double *expK, double *expKinv, double *lambda)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *K = malloc(sizeof(double) * nn);
    memcpy(K, thop, sizeof(double) * nn);
    const double mu = U / 2.0; /* src/model.c:15, half = 1 */
    for (int i = 0; i < n; i++) K[IDX(i, i, n)] -= mu;
    orc_expm_sym(n, K, -dtau, expK);   /* src/model.c:29 */
    orc_expm_sym(n, K, dtau, expKinv); /* src/model.c:30 */
    free(K);
    *lambda = acosh(exp(dtau * U / 2.0)); /* src/field.c:9 */
}

/* ------------------------------------------------------ linear algebra */

void orc_matmul(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) C[IDX(i, j, n)] = 0.0;
        for (int k = 0; k < n; k++) {
            const double b = B[IDX(k, j, n)];
            for (int i = 0; i < n; i++) C[IDX(i, j, n)] += A[IDX(i, k, n)] * b;
        }
    }
}

/* Householder QR, LAPACK dgeqr2/dlarfg convention: H_j = I - tau v v^T,
   v(j) = 1, beta = -sign(alpha) * ||(alpha, x)||, R(j,j) = beta. */
void orc_geqrf(int n, double *A, double *tau)
{
    for (int j = 0; j < n; j++) {
        /* ||x|| scaled by max|x_i| so that entries near the double range
           neither overflow nor underflow when squared (dnrm2, dlapy2) */
        double scale = 0.0, ss = 0.0;
        for (int i = j + 1; i < n; i++) scale = fmax(scale, fabs(A[IDX(i, j, n)]));
        if (scale > 0.0)
            for (int i = j + 1; i < n; i++) {
                const double t = A[IDX(i, j, n)] / scale;
                ss += t * t;
            }
        const double xnorm = scale * sqrt(ss);
        const double alpha = A[IDX(j, j, n)];
        if (xnorm == 0.0) {
            tau[j] = 0.0;
        } else {
            const double beta = -copysign(hypot(alpha, xnorm), alpha);
            tau[j] = (beta - alpha) / beta;
            const double sc = 1.0 / (alpha - beta);
            for (int i = j + 1; i < n; i++) A[IDX(i, j, n)] *= sc;
            A[IDX(j, j, n)] = beta;
        }
        if (tau[j] != 0.0) {
            for (int k = j + 1; k < n; k++) {
                double w = A[IDX(j, k, n)];
                for (int i = j + 1; i < n; i++) w += A[IDX(i, j, n)] * A[IDX(i, k, n)];
                w *= tau[j];
                A[IDX(j, k, n)] -= w;
                for (int i = j + 1; i < n; i++) A[IDX(i, k, n)] -= A[IDX(i, j, n)] * w;
            }
        }
    }
}

/* Q = H_0 H_1 ... H_{n-1} from the reflectors below the diagonal. */
void orc_orgqr(int n, double *A, const double *tau)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *Q = calloc(nn, sizeof(double));
    for (int i = 0; i < n; i++) Q[IDX(i, i, n)] = 1.0;
    for (int j = n - 1; j >= 0; j--) {
        if (tau[j] == 0.0) continue;
        for (int k = j; k < n; k++) {
            double w = Q[IDX(j, k, n)];
            for (int i = j + 1; i < n; i++) w += A[IDX(i, j, n)] * Q[IDX(i, k, n)];
            w *= tau[j];
            Q[IDX(j, k, n)] -= w;
            for (int i = j + 1; i < n; i++) Q[IDX(i, k, n)] -= A[IDX(i, j, n)] * w;
        }
    }
    memcpy(A, Q, sizeof(double) * nn);
    free(Q);
}

/* Partial-pivot LU (dgetf2): pivot = first index of max |A(i,j)|, i >= j. */
int orc_getrf(int n, double *A, int *piv)
{
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        double best = fabs(A[IDX(j, j, n)]);
        for (int i = j + 1; i < n; i++) {
            const double v = fabs(A[IDX(i, j, n)]);
            if (v > best) { best = v; p = i; }
        }
        piv[j] = p;
        if (A[IDX(p, j, n)] == 0.0) {
            if (info == 0) info = j + 1;
            continue;
        }
        if (p != j) {
            for (int k = 0; k < n; k++) {
                const double t = A[IDX(j, k, n)];
                A[IDX(j, k, n)] = A[IDX(p, k, n)];
                A[IDX(p, k, n)] = t;
            }
        }
        const double inv = 1.0 / A[IDX(j, j, n)];
        for (int i = j + 1; i < n; i++) A[IDX(i, j, n)] *= inv;
        for (int k = j + 1; k < n; k++) {
            const double a = A[IDX(j, k, n)];
            for (int i = j + 1; i < n; i++) A[IDX(i, k, n)] -= A[IDX(i, j, n)] * a;
        }
    }
    return info;
}

int orc_inverse(int n, const double *A, double *Ainv)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *LU = malloc(sizeof(double) * nn);
    int *piv = malloc(sizeof(int) * (size_t)n);
    memcpy(LU, A, sizeof(double) * nn);
    const int info = orc_getrf(n, LU, piv);
    if (info != 0) { free(LU); free(piv); return info; }
    /* solve LU X = P I column by column */
    memset(Ainv, 0, sizeof(double) * nn);
    for (int i = 0; i < n; i++) Ainv[IDX(i, i, n)] = 1.0;
    for (int j = 0; j < n; j++) {
        if (piv[j] != j) {
            for (int k = 0; k < n; k++) {
                const double t = Ainv[IDX(j, k, n)];
                Ainv[IDX(j, k, n)] = Ainv[IDX(piv[j], k, n)];
                Ainv[IDX(piv[j], k, n)] = t;
            }
        }
    }
    for (int c = 0; c < n; c++) {
        double *x = &Ainv[IDX(0, c, n)];
        for (int i = 0; i < n; i++)
            for (int k = 0; k < i; k++) x[i] -= LU[IDX(i, k, n)] * x[k];
        for (int i = n - 1; i >= 0; i--) {
            for (int k = i + 1; k < n; k++) x[i] -= LU[IDX(i, k, n)] * x[k];
            x[i] /= LU[IDX(i, i, n)];
        }
    }
    free(LU);
    free(piv);
    return 0;
}

void orc_trtri_unit_upper(int n, double *T)
{
    /* column j of T^{-1}: back substitution against unit-upper T */
    const size_t nn = (size_t)n * (size_t)n;
    double *X = calloc(nn, sizeof(double));
    for (int j = 0; j < n; j++) {
        X[IDX(j, j, n)] = 1.0;
        for (int i = j - 1; i >= 0; i--) {
            double acc = 0.0;
            for (int k = i + 1; k <= j; k++) acc += T[IDX(i, k, n)] * X[IDX(k, j, n)];
            X[IDX(i, j, n)] = -acc;
        }
    }
    memcpy(T, X, sizeof(double) * nn);
    free(X);
}

static int all_finite(size_t m, const double *x)
{
    for (size_t i = 0; i < m; i++) if (!isfinite(x[i])) return 0;
    return 1;
}

/* sign of det(A) via LU (la_logdet, src/linalg.c:255-286); A destroyed. */
static int det_sign_lu(int n, double *A, int *ok)
{
    int *piv = malloc(sizeof(int) * (size_t)n);
    const int info = orc_getrf(n, A, piv);
    int s = 1;
    *ok = (info == 0);
    for (int i = 0; i < n && *ok; i++) {
        if (A[IDX(i, i, n)] < 0.0) s = -s;
        if (piv[i] != i) s = -s;
    }
    free(piv);
    return s;
}

/* QR of M, then U' = Q, D' = diag R, Rn = D'^-1 R (unit upper). */
static int qr_udv(int n, double *M, double *Q, double *D, double *Rn)
{
    double *tau = malloc(sizeof(double) * (size_t)n);
    orc_geqrf(n, M, tau);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) Rn[IDX(i, j, n)] = (i <= j) ? M[IDX(i, j, n)] : 0.0;
    orc_orgqr(n, M, tau);
    memcpy(Q, M, sizeof(double) * (size_t)n * (size_t)n);
    free(tau);
    for (int i = 0; i < n; i++) {
        D[i] = Rn[IDX(i, i, n)];
        if (!isfinite(D[i]) || D[i] == 0.0) return 1;
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) Rn[IDX(i, j, n)] /= D[i];
    return 0;
}

/* udv_lmul_impl, non-centered (src/linalg.c:515-642):
   M = (B U) diag(D); QR(M) -> U' = Q, D' = diag R, T' = (D'^-1 R) T. */
int orc_udv_lmul(int n, double *U, double *D, double *T, const double *B)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *M = malloc(sizeof(double) * nn), *R = malloc(sizeof(double) * nn);
    double *Tn = malloc(sizeof(double) * nn);
    int rc = 1;
    orc_matmul(n, B, U, M);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) M[IDX(i, j, n)] *= D[j];
    if (!all_finite(nn, M)) goto out;
    if (qr_udv(n, M, U, D, R)) goto out;
    orc_matmul(n, R, T, Tn);
    memcpy(T, Tn, sizeof(double) * nn);
    rc = all_finite(nn, U) && all_finite(nn, T) ? 0 : 1;
out:
    free(M); free(R); free(Tn);
    return rc;
}

/* udv_rmul_impl, non-centered (src/linalg.c:662-784):
   M = diag(D) (T B); QR(M) -> U' = U Q, D' = diag R, T' = D'^-1 R. */
int orc_udv_rmul(int n, double *U, double *D, double *T, const double *B)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *M = malloc(sizeof(double) * nn), *Q = malloc(sizeof(double) * nn);
    double *Un = malloc(sizeof(double) * nn);
    int rc = 1;
    orc_matmul(n, T, B, M);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) M[IDX(i, j, n)] *= D[i];
    if (!all_finite(nn, M)) goto out;
    if (qr_udv(n, M, Q, D, T)) goto out;
    orc_matmul(n, U, Q, Un);
    memcpy(U, Un, sizeof(double) * nn);
    rc = all_finite(nn, U) && all_finite(nn, T) ? 0 : 1;
out:
    free(M); free(Q); free(Un);
    return rc;
}

/* udv_combine (src/linalg.c:796-909): C = Dl (Tl Ur) Dr; QR(C);
   Uo = Ul Q, Do = diag R, To = (Do^-1 R) Tr. */
int orc_udv_combine(int n, const double *Ul, const double *Dl,
                    const double *Tl, const double *Ur, const double *Dr,
                    const double *Tr, double *Uo, double *Do, double *To)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *C = malloc(sizeof(double) * nn), *Q = malloc(sizeof(double) * nn);
    double *R = malloc(sizeof(double) * nn);
    int rc = 1;
    orc_matmul(n, Tl, Ur, C);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) C[IDX(i, j, n)] = Dl[i] * C[IDX(i, j, n)] * Dr[j];
    if (!all_finite(nn, C)) goto out;
    if (qr_udv(n, C, Q, Do, R)) goto out;
    orc_matmul(n, Ul, Q, Uo);
    orc_matmul(n, R, Tr, To);
    rc = all_finite(nn, Uo) && all_finite(nn, To) ? 0 : 1;
out:
    free(C); free(Q); free(R);
    return rc;
}

/* The big/small split of one D entry (src/linalg.c:966-986, 1190-1215):
   without an offset, Db = D if |D| > 1 else 1 and Ds = D / Db; with a log
   offset the true log-magnitude ell = log|D| + offset decides, and the
   INVERSE of the big part is kept (Dbi = sgn D e^-ell) so that nothing
   overflows.  Returns 0, or 3 (LINALG_FAILURE_EFFECTIVE_LOG_RANGE) when
   -|ell| is below log(DBL_TRUE_MIN) (:166-182).  *big as the sign rule
   (:1046-1052) reads it. */
static int split_entry(double d, double off, double *dbi, double *ds, int *big)
{
    if (off == 0.0) {
        *big = fabs(d) > 1.0;
        *dbi = *big ? 1.0 / d : 1.0;
        *ds = *big ? 1.0 : d;
        return 0;
    }
    const double ell = log(fabs(d)) + off;
    if (!isfinite(ell) || -fabs(ell) < log(4.9406564584124654e-324)) return 3;
    *big = ell > 0.0;
    *dbi = *big ? copysign(exp(-ell), d) : 1.0;
    *ds = *big ? 1.0 : copysign(exp(ell), d);
    return 0;
}

/* udv_inv_one_plus_work (src/linalg.c:911-1115), with the UDV's log
   offset (0 = not centered):
   G = (1 + U D T)^-1 = T^-1 (Db^-1 U^T T^-1 + Ds)^-1 Db^-1 U^T,
   sign = prod_{big} sgn D_i * sgn det U * sgn det M. */
int orc_udv_inv_one_plus_off(int n, const double *U, const double *D,
                             const double *T, double off, double *g,
                             int *sign)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *Dbi = malloc(sizeof(double) * (size_t)n);
    double *Ds = malloc(sizeof(double) * (size_t)n);
    int *big = malloc(sizeof(int) * (size_t)n);
    double *Ti = malloc(sizeof(double) * nn), *M = malloc(sizeof(double) * nn);
    double *Mi = malloc(sizeof(double) * nn), *W = malloc(sizeof(double) * nn);
    double *W2 = malloc(sizeof(double) * nn);
    int rc = 1;
    if (sign) *sign = 0;
    if (!all_finite(nn, U) || !all_finite((size_t)n, D) || !all_finite(nn, T) || !isfinite(off)) goto out;
    for (int i = 0; i < n; i++) {
        const int e = split_entry(D[i], off, &Dbi[i], &Ds[i], &big[i]);
        if (e) { rc = e; goto out; }
        if (!isfinite(Dbi[i]) || !isfinite(Ds[i]) || Dbi[i] == 0.0 || Ds[i] == 0.0) goto out;
    }
    memcpy(Ti, T, sizeof(double) * nn);
    orc_trtri_unit_upper(n, Ti);
    /* M = Db^-1 (U^T Ti) + diag(Ds) */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc += U[IDX(k, i, n)] * Ti[IDX(k, j, n)];
            M[IDX(i, j, n)] = off != 0.0 ? acc * Dbi[i] : (big[i] ? acc / D[i] : acc);
        }
    for (int i = 0; i < n; i++) M[IDX(i, i, n)] += Ds[i];
    if (!all_finite(nn, M)) goto out;
    if (orc_inverse(n, M, Mi) != 0 || !all_finite(nn, Mi)) goto out;
    {
        int s = 1, ok1, ok2;
        for (int i = 0; i < n; i++)
            if (big[i] && D[i] < 0.0) s = -s;
        memcpy(W, U, sizeof(double) * nn);
        const int su = det_sign_lu(n, W, &ok1);
        memcpy(W, M, sizeof(double) * nn);
        const int sm = det_sign_lu(n, W, &ok2);
        if (!ok1 || !ok2) goto out;
        if (sign) *sign = s * su * sm;
    }
    /* W2 = (Mi Db^-1) U^T ; g = Ti W2 */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) {
                const double mk = off != 0.0 ? Mi[IDX(i, k, n)] * Dbi[k]
                                             : (big[k] ? Mi[IDX(i, k, n)] / D[k] : Mi[IDX(i, k, n)]);
                acc += mk * U[IDX(j, k, n)];
            }
            W2[IDX(i, j, n)] = acc;
        }
    orc_matmul(n, Ti, W2, g);
    rc = all_finite(nn, g) ? 0 : 1;
out:
    if (rc && sign) *sign = 0;
    free(Dbi); free(Ds); free(big); free(Ti); free(M); free(Mi); free(W); free(W2);
    return rc;
}

int orc_udv_inv_one_plus(int n, const double *U, const double *D,
                         const double *T, double *g, int *sign)
{
    return orc_udv_inv_one_plus_off(n, U, D, T, 0.0, g, sign);
}

/* udv_inv_one_plus_two_sided_work (src/linalg.c:1117-1389):
   1 + L R with L = Ul Dl Tl, R = Ur Dr Tr (log offsets offl, offr; 0 = not
   centered) as Ul Dlb H Drb Tr,
   H = Dlb^-1 (Ul^T Tr^-1) Drb^-1 + Dls (Tl Ur) Drs,
   G = Tr^-1 Drb^-1 H^-1 Dlb^-1 Ul^T,
   sign = prod_{big} sgn Dl * prod_{big} sgn Dr * sgn det Ul * sgn det H. */
int orc_udv_inv_two_sided(int n, const double *Ul, const double *Dl,
                          const double *Tl, double offl, const double *Ur,
                          const double *Dr, const double *Tr, double offr,
                          double *g, int *sign)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *lbi = malloc(sizeof(double) * (size_t)n), *ls = malloc(sizeof(double) * (size_t)n);
    double *rbi = malloc(sizeof(double) * (size_t)n), *rs = malloc(sizeof(double) * (size_t)n);
    int *lb = malloc(sizeof(int) * (size_t)n), *rb = malloc(sizeof(int) * (size_t)n);
    double *TLUR = malloc(sizeof(double) * nn), *Ti = malloc(sizeof(double) * nn);
    double *H = malloc(sizeof(double) * nn), *Hi = malloc(sizeof(double) * nn);
    double *W = malloc(sizeof(double) * nn), *W2 = malloc(sizeof(double) * nn);
    int rc = 1;
    if (sign) *sign = 0;
    if (!all_finite(nn, Ul) || !all_finite((size_t)n, Dl) || !all_finite(nn, Tl) || !all_finite(nn, Ur) ||
        !all_finite((size_t)n, Dr) || !all_finite(nn, Tr) || !isfinite(offl) || !isfinite(offr))
        goto out;
    for (int i = 0; i < n; i++) {
        int e = split_entry(Dl[i], offl, &lbi[i], &ls[i], &lb[i]);
        if (!e) e = split_entry(Dr[i], offr, &rbi[i], &rs[i], &rb[i]);
        if (e) { rc = e; goto out; }
        if (lbi[i] == 0.0 || rbi[i] == 0.0 || ls[i] == 0.0 || rs[i] == 0.0) goto out;
    }
    orc_matmul(n, Tl, Ur, TLUR);
    memcpy(Ti, Tr, sizeof(double) * nn);
    orc_trtri_unit_upper(n, Ti);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double acc = 0.0; /* (Ul^T Tr^-1)(i, j) */
            for (int k = 0; k < n; k++) acc += Ul[IDX(k, i, n)] * Ti[IDX(k, j, n)];
            double large = offl != 0.0 ? acc * lbi[i] : (lb[i] ? acc / Dl[i] : acc);
            large = offr != 0.0 ? large * rbi[j] : (rb[j] ? large / Dr[j] : large);
            H[IDX(i, j, n)] = large + ls[i] * TLUR[IDX(i, j, n)] * rs[j];
        }
    if (!all_finite(nn, H)) goto out;
    if (orc_inverse(n, H, Hi) != 0 || !all_finite(nn, Hi)) goto out;
    {
        int s = 1, ok1, ok2;
        for (int i = 0; i < n; i++) {
            if (lb[i] && Dl[i] < 0.0) s = -s;
            if (rb[i] && Dr[i] < 0.0) s = -s;
        }
        memcpy(W, Ul, sizeof(double) * nn);
        const int su = det_sign_lu(n, W, &ok1);
        memcpy(W, H, sizeof(double) * nn);
        const int sh = det_sign_lu(n, W, &ok2);
        if (!ok1 || !ok2) goto out;
        if (sign) *sign = s * su * sh;
    }
    /* W2 = (Hi Dlb^-1) Ul^T, rows scaled by Drb^-1; g = Tr^-1 W2 */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) {
                const double hk = offl != 0.0 ? Hi[IDX(i, k, n)] * lbi[k]
                                              : (lb[k] ? Hi[IDX(i, k, n)] / Dl[k] : Hi[IDX(i, k, n)]);
                acc += hk * Ul[IDX(j, k, n)];
            }
            W2[IDX(i, j, n)] = offr != 0.0 ? acc * rbi[i] : (rb[i] ? acc / Dr[i] : acc);
        }
    orc_matmul(n, Ti, W2, g);
    rc = all_finite(nn, g) ? 0 : 1;
out:
    if (rc && sign) *sign = 0;
    free(lbi); free(ls); free(rbi); free(rs); free(lb); free(rb);
    free(TLUR); free(Ti); free(H); free(Hi); free(W); free(W2);
    return rc;
}

/* udv_recenter (src/linalg.c:453-513) with the hard margin 8 of the
   centered updates: D /= e^c, offset += c for the centre c of the log|D|
   range; 1 (nothing changed) when the radius leaves less than the margin
   below log(DBL_MAX) or a rescaled entry is not representable. */
int orc_udv_recenter(int n, double *D, double *off)
{
    double lo = 0.0, hi = 0.0;
    if (!isfinite(*off)) return 1;
    for (int i = 0; i < n; i++) {
        if (!isfinite(D[i]) || D[i] == 0.0) return 1;
        const double ld = log(fabs(D[i]));
        if (i == 0 || ld < lo) lo = ld;
        if (i == 0 || ld > hi) hi = ld;
    }
    const double c = lo + 0.5 * (hi - lo), r = 0.5 * (hi - lo);
    const double margin = log(1.7976931348623157e308) - r;
    if (!isfinite(c) || !isfinite(r) || !isfinite(margin) || margin < 8.0 || !isfinite(*off + c)) return 1;
    for (int i = 0; i < n; i++) {
        const double v = exp(log(fabs(D[i])) - c);
        if (!isfinite(v) || v == 0.0) return 1;
    }
    for (int i = 0; i < n; i++) D[i] = copysign(exp(log(fabs(D[i])) - c), D[i]);
    *off += c;
    return 0;
}

static double logaddexp2(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const double hi = a > b ? a : b, lo = a > b ? b : a;
    return hi + log1p(exp(lo - hi));
}

/* centered_update_has_margin (src/linalg.c:119-164): a bound on the
   largest QR input entry of the update, below log(DBL_MAX) - 8. */
static int update_has_margin(int n, const double *D, const double *T, const double *B, int left)
{
    const double limit = log(1.7976931348623157e308) - 8.0;
    double worst = -INFINITY;
    if (left) {
        double lb = -INFINITY;
        for (int i = 0; i < n; i++) {
            double row = -INFINITY;
            for (int k = 0; k < n; k++)
                if (B[IDX(i, k, n)] != 0.0) row = logaddexp2(row, log(fabs(B[IDX(i, k, n)])));
            if (row > lb) lb = row;
        }
        for (int j = 0; j < n; j++) {
            const double c = log(fabs(D[j])) + lb;
            if (c > worst) worst = c;
        }
    } else {
        for (int i = 0; i < n; i++) {
            double row = -INFINITY;
            for (int k = 0; k < n; k++) {
                double bmax = 0.0;
                for (int j = 0; j < n; j++) bmax = fmax(bmax, fabs(B[IDX(k, j, n)]));
                const double t = fabs(T[IDX(i, k, n)]);
                if (t != 0.0 && bmax != 0.0) row = logaddexp2(row, log(t) + log(bmax));
            }
            const double c = log(fabs(D[i])) + row;
            if (c > worst) worst = c;
        }
    }
    return isfinite(worst) && worst <= limit;
}

/* udv_lmul_centered_work / udv_rmul_centered (src/linalg.c:515-791,
   centered = 1): recentre, check the update's margin (inputs restored on
   failure), update, recentre.  0 ok, 1 LINALG_FAILURE_CENTERED_RADIUS,
   2 LINALG_FAILURE_CENTERED_UPDATE_MARGIN, 4 other failure. */
static int udv_mul_centered(int n, double *U, double *D, double *T, double *off, const double *B, int left)
{
    double *D0 = malloc(sizeof(double) * (size_t)n);
    const double off0 = *off;
    int rc = 0;
    memcpy(D0, D, sizeof(double) * (size_t)n);
    if (orc_udv_recenter(n, D, off)) { rc = 1; goto out; }
    if (!update_has_margin(n, D, T, B, left)) {
        memcpy(D, D0, sizeof(double) * (size_t)n);
        *off = off0;
        rc = 2;
        goto out;
    }
    if (left ? orc_udv_lmul(n, U, D, T, B) : orc_udv_rmul(n, U, D, T, B)) { rc = 4; goto out; }
    if (orc_udv_recenter(n, D, off)) rc = 1;
out:
    free(D0);
    return rc;
}

int orc_udv_lmul_centered(int n, double *U, double *D, double *T, double *off, const double *B)
{
    return udv_mul_centered(n, U, D, T, off, B, 1);
}

int orc_udv_rmul_centered(int n, double *U, double *D, double *T, double *off, const double *B)
{
    return udv_mul_centered(n, U, D, T, off, B, 0);
}

/* ------------------------------------------------------ B and the wrap */

void orc_expv(double lambda, double U, double dtau, double sigma,
              double expv[2])
{
    /* src/green.c:63-67 */
    const double c = -dtau * U / 2.0;
    expv[0] = exp(lambda * sigma * -1.0 + c);
    expv[1] = exp(lambda * sigma * 1.0 + c);
}

/* src/green.c:126-137: B = expK diag(d), d_j = expv[s_j > 0]. */
void orc_build_B(int n, const double *expK, const double expv[2],
                 const signed char *s_l, double *out)
{
    for (int j = 0; j < n; j++) {
        const double d = expv[s_l[j] > 0 ? 1 : 0];
        for (int i = 0; i < n; i++) out[IDX(i, j, n)] = expK[IDX(i, j, n)] * d;
    }
}

/* src/green.c:139-149: B^-1 = diag(1/d) expKinv. */
void orc_build_Binv(int n, const double *expKinv, const double expv[2],
                    const signed char *s_l, double *out)
{
    const double inv[2] = {1.0 / expv[0], 1.0 / expv[1]};
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            out[IDX(i, j, n)] = expKinv[IDX(i, j, n)] * inv[s_l[i] > 0 ? 1 : 0];
}

/* src/green.c:151-166: B_{b+len-1} ... B_b. */
void orc_build_Bblock(int n, const double *expK, const double expv[2],
                      const signed char *s, int l_begin, int len,
                      double *out)
{
    const size_t nn = (size_t)n * (size_t)n;
    if (len <= 0) {
        memset(out, 0, sizeof(double) * nn);
        for (int i = 0; i < n; i++) out[IDX(i, i, n)] = 1.0;
        return;
    }
    double *B = malloc(sizeof(double) * nn), *tmp = malloc(sizeof(double) * nn);
    orc_build_B(n, expK, expv, s + (size_t)l_begin * n, out);
    for (int p = 1; p < len; p++) {
        orc_build_B(n, expK, expv, s + (size_t)(l_begin + p) * n, B);
        orc_matmul(n, B, out, tmp);
        memcpy(out, tmp, sizeof(double) * nn);
    }
    free(B); free(tmp);
}

/* src/green.c:509-527: G <- B G B^-1. */
void orc_wrap(int n, const double *expK, const double *expKinv,
              const double expv[2], const signed char *s_l, double *g)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *B = malloc(sizeof(double) * nn), *Bi = malloc(sizeof(double) * nn);
    double *tmp = malloc(sizeof(double) * nn);
    orc_build_B(n, expK, expv, s_l, B);
    orc_build_Binv(n, expKinv, expv, s_l, Bi);
    orc_matmul(n, B, g, tmp);
    orc_matmul(n, tmp, Bi, g);
    free(B); free(Bi); free(tmp);
}

/* green_wrap_backward (src/green.c:529-546): G <- B_l^-1 G B_l. */
void orc_wrap_backward(int n, const double *expK, const double *expKinv,
                       const double expv[2], const signed char *s_l, double *g)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *B = malloc(sizeof(double) * nn), *Bi = malloc(sizeof(double) * nn);
    double *tmp = malloc(sizeof(double) * nn);
    orc_build_B(n, expK, expv, s_l, B);
    orc_build_Binv(n, expKinv, expv, s_l, Bi);
    orc_matmul(n, Bi, g, tmp);
    orc_matmul(n, tmp, B, g);
    free(B); free(Bi); free(tmp);
}

/* src/green.c:372-382: G_dn = delta - P_i P_j G_up(j,i). */
void orc_ph_down(int n, const int *bipart, const double *gu, double *gd)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            gd[IDX(i, j, n)] = ((i == j) ? 1.0 : 0.0) -
                               (double)(bipart[i] * bipart[j]) * gu[IDX(j, i, n)];
}

/* --------------------------------------------------------- measurement */

void orc_measure_sample(int n, const double *thop, double U,
                        const double *gu, const double *gd, double *out5)
{
    double ekin = 0.0, eint = 0.0, ntot = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double t = thop[IDX(i, j, n)];
            const double d = (i == j) ? 1.0 : 0.0;
            ekin += t * (d - gu[IDX(j, i, n)]);
            ekin += t * (d - gd[IDX(j, i, n)]);
        }
    for (int i = 0; i < n; i++) {
        const double nu = 1.0 - gu[IDX(i, i, n)], nd = 1.0 - gd[IDX(i, i, n)];
        eint += nu * nd;
        ntot += nu + nd;
    }
    out5[1] = ekin;
    out5[2] = U * eint;
    out5[3] = ntot;
    out5[4] = eint / (double)n;
    out5[0] = out5[1] + out5[2];
}

void orc_phase_cos(int Lx, int Ly, int nq, const int *mx, const int *my,
                   double *phase_cos)
{
    const int n = Lx * Ly;
    const double two_pi = 2.0 * acos(-1.0);
    for (int q = 0; q < nq; q++)
        for (int disp = 0; disp < n; disp++) {
            const int dx = disp % Lx, dy = disp / Lx;
            phase_cos[disp + (size_t)q * n] =
                cos(two_pi * ((double)mx[q] * dx / (double)Lx +
                              (double)my[q] * dy / (double)Ly));
        }
}

int orc_szz_sperp(int Lx, int Ly, int nq, const double *phase_cos,
                  const double *gu, const double *gd, double *szz,
                  double *sperp)
{
    const int n = Lx * Ly;
    double *cz = calloc((size_t)n, sizeof(double));
    double *cp = calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; i++) {
        const int xi = i % Lx, yi = i / Lx;
        for (int j = 0; j < n; j++) {
            const double d = (i == j) ? 1.0 : 0.0;
            const double si = gd[IDX(i, i, n)] - gu[IDX(i, i, n)];
            const double sj = gd[IDX(j, j, n)] - gu[IDX(j, j, n)];
            const double czz = 0.25 * (si * sj + (d - gu[IDX(j, i, n)]) * gu[IDX(i, j, n)] +
                                       (d - gd[IDX(j, i, n)]) * gd[IDX(i, j, n)]);
            const double cpp = 0.5 * ((d - gu[IDX(j, i, n)]) * gd[IDX(i, j, n)] +
                                      (d - gd[IDX(j, i, n)]) * gu[IDX(i, j, n)]);
            int dx = xi - j % Lx, dy = yi - j / Lx;
            if (dx < 0) dx += Lx;
            if (dy < 0) dy += Ly;
            cz[dx + dy * Lx] += czz;
            cp[dx + dy * Lx] += cpp;
        }
    }
    int bad = 0;
    for (int q = 0; q < nq; q++) {
        double a = 0.0, b = 0.0;
        for (int disp = 0; disp < n; disp++) {
            a += phase_cos[disp + (size_t)q * n] * cz[disp];
            b += phase_cos[disp + (size_t)q * n] * cp[disp];
        }
        if (szz) szz[q] = a / (double)n;
        if (sperp) sperp[q] = b / (double)n;
        if (!isfinite(a) || !isfinite(b)) bad = 1;
    }
    free(cz);
    free(cp);
    return bad;
}

/* src/measure.c:48-69 */
void orc_jackknife(const double *x, int N, double *mean, double *err)
{
    double sum = 0.0;
    for (int i = 0; i < N; i++) sum += x[i];
    const double mu = (N > 0) ? sum / (double)N : 0.0;
    *mean = mu;
    if (N < 2) { *err = 0.0; return; }
    double var = 0.0;
    for (int i = 0; i < N; i++) {
        const double d = (sum - x[i]) / (double)(N - 1) - mu;
        var += d * d;
    }
    *err = sqrt(var * (double)(N - 1) / (double)N);
}

/* ------------------------------------------------------- replica engine */

typedef struct {
    double *U, *D, *T;
    double off; /* log offset of a centered UDV (0 otherwise) */
} Udv;

struct orc_replica {
    int n, L, stab, M, cap, status, det_sign, drift_on;
    double U, dtau, lambda, sign;
    double *expK, *expKinv, *g, *gd, *C, *V;
    double ncache[2][2]; /* [isigma][is], src/field.c:10-17 */
    double expv[2];
    int delay;
    int *bipart, *b;
    signed char *s;
    uint64_t rng[4];
    unsigned long long acc, att, sweeps, drift_samples;
    double drift_max, drift_sum;
    Udv *S, left, comb;
    Udv *A;             /* prefix stack of the alternating mode (M + 1) */
    int alternating, dir, prefix_ok, suffix_ok;
    int rebuild;        /* 0 combine, 1 two_sided, 2 centered (src/green.h:10-14) */
    int reason;         /* LINALG_FAILURE_* of a failed replica */
    double *rlog;       /* ratio R_up R_dn of every proposal, [l * n + i] (NULL: off) */
};

static void udv_alloc(Udv *u, int n)
{
    const size_t nn = (size_t)n * (size_t)n;
    u->U = calloc(nn, sizeof(double));
    u->T = calloc(nn, sizeof(double));
    u->D = calloc((size_t)n, sizeof(double));
}

static void udv_eye(Udv *u, int n)
{
    memset(u->U, 0, sizeof(double) * (size_t)n * n);
    memset(u->T, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; i++) { u->U[IDX(i, i, n)] = 1.0; u->T[IDX(i, i, n)] = 1.0; u->D[i] = 1.0; }
    u->off = 0.0;
}

static void udv_cp(Udv *d, const Udv *s, int n)
{
    memcpy(d->U, s->U, sizeof(double) * (size_t)n * n);
    memcpy(d->T, s->T, sizeof(double) * (size_t)n * n);
    memcpy(d->D, s->D, sizeof(double) * (size_t)n);
    d->off = s->off;
}

/* green_udv_lmul / green_udv_rmul / dqmc_udv_lmul / dqmc_udv_rmul
   (src/green.c:22-38, src/dqmc.c:37-53): the centered variants in the
   centered rebuild mode.  0 ok; on failure the replica's reason is set. */
static int rep_lmul(orc_replica *R, Udv *u, const double *B)
{
    int rc;
    if (R->rebuild == 2) rc = orc_udv_lmul_centered(R->n, u->U, u->D, u->T, &u->off, B);
    else rc = orc_udv_lmul(R->n, u->U, u->D, u->T, B) ? 4 : 0;
    if (rc && rc < 4) R->reason = rc;
    return rc;
}

static int rep_rmul(orc_replica *R, Udv *u, const double *B)
{
    int rc;
    if (R->rebuild == 2) rc = orc_udv_rmul_centered(R->n, u->U, u->D, u->T, &u->off, B);
    else rc = orc_udv_rmul(R->n, u->U, u->D, u->T, B) ? 4 : 0;
    if (rc && rc < 4) R->reason = rc;
    return rc;
}

static void udv_release(Udv *u)
{
    free(u->U); free(u->T); free(u->D);
}

/* green_from_scratch (src/green.c:262-306) into out; returns 0 ok. */
static int from_scratch(orc_replica *R, int l0, double *out, int *sgn)
{
    const int n = R->n;
    Udv u;
    udv_alloc(&u, n);
    udv_eye(&u, n);
    double *B = malloc(sizeof(double) * (size_t)n * n);
    int rc = 0;
    for (int k = 0; k < R->L && rc == 0; k++) {
        const int l = (l0 + k) % R->L;
        orc_build_B(n, R->expK, R->expv, R->s + (size_t)l * n, B);
        rc = rep_lmul(R, &u, B);
    }
    if (rc == 0) {
        rc = orc_udv_inv_one_plus_off(n, u.U, u.D, u.T, u.off, out, sgn);
        if (rc == 3) R->reason = 3;
    }
    free(B);
    udv_release(&u);
    return rc;
}

/* G~_ii = g_ii - sum_p C_ip V_ip (src/green.c:384-392) */
static double delay_diag(const orc_replica *R, int i)
{
    const int n = R->n;
    double gii = R->g[IDX(i, i, n)];
    for (int p = 0; p < R->delay; p++) gii -= R->C[IDX(i, p, n)] * R->V[IDX(i, p, n)];
    return gii;
}

/* src/green.c:462-475: g -= C V^T */
static void delay_flush(orc_replica *R)
{
    const int n = R->n;
    for (int p = 0; p < R->delay; p++)
        for (int k = 0; k < n; k++) {
            const double v = R->V[IDX(k, p, n)];
            for (int j = 0; j < n; j++) R->g[IDX(j, k, n)] -= R->C[IDX(j, p, n)] * v;
        }
    R->delay = 0;
}

/* src/green.c:418-460 */
static void delay_accept(orc_replica *R, int i, double N)
{
    const int n = R->n;
    if (R->delay >= R->cap) delay_flush(R);
    const double denom = 1.0 + (1.0 - delay_diag(R, i)) * N;
    const double fac = N / denom;
    const int p = R->delay;
    double *c = &R->C[IDX(0, p, n)], *v = &R->V[IDX(0, p, n)];
    for (int j = 0; j < n; j++) {
        double gji = R->g[IDX(j, i, n)];
        for (int q = 0; q < p; q++) gji -= R->C[IDX(j, q, n)] * R->V[IDX(i, q, n)];
        c[j] = fac * (((i == j) ? 1.0 : 0.0) - gji);
    }
    for (int k = 0; k < n; k++) {
        double gik = R->g[IDX(i, k, n)];
        for (int q = 0; q < p; q++) gik -= R->C[IDX(i, q, n)] * R->V[IDX(k, q, n)];
        v[k] = gik;
    }
    R->delay++;
    if (R->delay == R->cap) delay_flush(R);
}

/* dqmc_record_stab_drift (src/dqmc.c:64-98), PH path. */
static int record_drift(orc_replica *R, int tau)
{
    if (!R->drift_on) return 0;
    const int n = R->n;
    double *ref = malloc(sizeof(double) * (size_t)n * n);
    int sg = 0;
    const int rc = from_scratch(R, tau, ref, &sg);
    if (rc == 0 && sg != 0) {
        double m = 0.0;
        for (size_t k = 0; k < (size_t)n * n; k++) {
            const double d = fabs(R->g[k] - ref[k]);
            if (d > m) m = d;
        }
        R->drift_samples++;
        R->drift_sum += m;
        if (m > R->drift_max || R->drift_samples == 1) R->drift_max = m;
    }
    free(ref);
    return (rc == 0 && sg != 0) ? 0 : 1;
}

orc_replica *orc_replica_new(int n, int L, int stab, const double *thop,
                             const int *bipart, double U, double dtau,
                             uint64_t seed, int delay_cap, int drift)
{
    orc_replica *R = calloc(1, sizeof(*R));
    const size_t nn = (size_t)n * (size_t)n;
    R->n = n; R->L = L; R->stab = stab > 0 ? stab : 8; R->U = U; R->dtau = dtau;
    R->cap = delay_cap > 0 ? delay_cap : 16; /* src/green.c:10 */
    R->M = (L + R->stab - 1) / R->stab;     /* src/green.c:172 */
    R->drift_on = drift;
    R->expK = malloc(sizeof(double) * nn);
    R->expKinv = malloc(sizeof(double) * nn);
    R->g = malloc(sizeof(double) * nn);
    R->gd = malloc(sizeof(double) * nn);
    R->C = malloc(sizeof(double) * nn);
    R->V = malloc(sizeof(double) * nn);
    R->bipart = malloc(sizeof(int) * (size_t)n);
    memcpy(R->bipart, bipart, sizeof(int) * (size_t)n);
    orc_model(n, thop, U, dtau, R->expK, R->expKinv, &R->lambda);
    for (int isg = 0; isg < 2; isg++)
        for (int is = 0; is < 2; is++)
            R->ncache[isg][is] = exp(-2.0 * R->lambda * (isg ? 1.0 : -1.0) * (is ? 1.0 : -1.0)) - 1.0;
    orc_expv(R->lambda, U, dtau, 1.0, R->expv);
    /* src/replica_run.c:112-116: seed, then L*n draws for the field */
    orc_rng_seed(R->rng, seed);
    R->s = malloc((size_t)L * n);
    for (int k = 0; k < L * n; k++) R->s[k] = (orc_rng_double(R->rng) < 0.5) ? -1 : 1;
    R->b = malloc(sizeof(int) * (size_t)(R->M + 1));
    R->S = calloc((size_t)R->M + 1, sizeof(Udv));
    for (int j = 0; j <= R->M; j++) {
        R->b[j] = j * R->stab > L ? L : j * R->stab; /* src/green.c:176-181 */
        udv_alloc(&R->S[j], n);
        udv_eye(&R->S[j], n);
    }
    udv_alloc(&R->left, n);
    udv_alloc(&R->comb, n);
    orc_replica_rebuild(R);
    return R;
}

/* dqmc_init_modes' initial rebuild (src/dqmc.c:464-481). */
int orc_replica_rebuild(orc_replica *R)
{
    int sg = 0;
    const int rc = from_scratch(R, 0, R->g, &sg);
    R->delay = 0;
    R->det_sign = sg;
    if (rc != 0 || sg == 0) { R->status = 1; R->sign = 0.0; return 1; }
    orc_ph_down(R->n, R->bipart, R->g, R->gd);
    R->sign = 1.0;
    return 0;
}

int orc_replica_status(const orc_replica *R) { return R->status; }

/* One slice of single-site Metropolis with delayed updates, then the flush
   (the loop bodies of src/dqmc.c:656-679 and :874-891; green_wrap and
   green_wrap_backward flush first, src/green.c:509-511, 529-531). */
static void metro_slice(orc_replica *R, int l)
{
    const int n = R->n;
    for (int i = 0; i < n; i++) {
        const signed char s = R->s[(size_t)l * n + i];
        const double Nu = R->ncache[1][s > 0];
        const double Nd = R->ncache[0][s > 0];
        const double gii = delay_diag(R, i);
        const double Ru = 1.0 + (1.0 - gii) * Nu;
        const double Rd = 1.0 + gii * Nd;
        const double Rt = Ru * Rd;
        if (R->rlog) R->rlog[(size_t)l * n + i] = Rt;
        R->att++;
        if (orc_rng_double(R->rng) < fabs(Rt)) {
            R->acc++;
            delay_accept(R, i, Nu);
            R->s[(size_t)l * n + i] = (signed char)-s;
            if (Rt < 0.0) R->sign = -R->sign;
        }
    }
    delay_flush(R);
}

/* green_from_boundary_factors (src/green.c:308-345): combine +
   inv_one_plus, or the two-sided inverse in the two_sided and centered
   modes; then the PH map and the sign reset (src/dqmc.c:731-738, 843-846). */
static int rebuild_from_factors(orc_replica *R, const Udv *prefix, const Udv *suffix)
{
    const int n = R->n;
    int sg = 0, rc;
    if (R->rebuild == 0)
        rc = orc_udv_combine(n, prefix->U, prefix->D, prefix->T, suffix->U, suffix->D, suffix->T, R->comb.U,
                             R->comb.D, R->comb.T) ||
             orc_udv_inv_one_plus(n, R->comb.U, R->comb.D, R->comb.T, R->g, &sg);
    else
        rc = orc_udv_inv_two_sided(n, prefix->U, prefix->D, prefix->T, prefix->off, suffix->U, suffix->D,
                                   suffix->T, suffix->off, R->g, &sg);
    if (rc == 3) R->reason = 3;
    if (rc || sg == 0) {
        R->status = 1;
        R->det_sign = 0;
        return 1;
    }
    R->det_sign = sg;
    orc_ph_down(n, R->bipart, R->g, R->gd);
    R->sign = 1.0;
    return 0;
}

/* green_from_left_udv (src/green.c:352-370) + PH map + sign reset. */
static int rebuild_from_left(orc_replica *R, const Udv *u)
{
    int sg = 0;
    const int rc = orc_udv_inv_one_plus_off(R->n, u->U, u->D, u->T, u->off, R->g, &sg);
    if (rc == 3) R->reason = 3;
    if (rc || sg == 0) {
        R->status = 1;
        R->det_sign = 0;
        return 1;
    }
    R->det_sign = sg;
    orc_ph_down(R->n, R->bipart, R->g, R->gd);
    R->sign = 1.0;
    return 0;
}

/* dqmc_sweep_forward, PH branch (src/dqmc.c:609-813).  carry != 0 is the
   alternating mode's forward leg: the suffix stack is rebuilt only when the
   previous backward sweep did not leave it, and the left UDV is stored into
   the prefix stack at every boundary and at tau = L. */
static int sweep_forward(orc_replica *R, int carry)
{
    const int n = R->n, L = R->L, stab = R->stab, M = R->M;
    double *Bb = malloc(sizeof(double) * (size_t)n * n);
    int tau_left = 0;

    if (!carry || !R->suffix_ok) {
        /* green_stack_build_suffix (src/green.c:203-226) */
        udv_eye(&R->S[M], n);
        udv_eye(&R->S[0], n);
        for (int j = M - 1; j >= 1; j--) {
            orc_build_Bblock(n, R->expK, R->expv, R->s, R->b[j], R->b[j + 1] - R->b[j], Bb);
            udv_cp(&R->S[j], &R->S[j + 1], n);
            if (rep_rmul(R, &R->S[j], Bb)) { R->status = 1; goto done; }
        }
    }
    udv_eye(&R->left, n);
    if (carry) udv_eye(&R->A[0], n);

    for (int l = 0; l < L; l++) {
        metro_slice(R, l);
        orc_wrap(n, R->expK, R->expKinv, R->expv, R->s + (size_t)l * n, R->g);
        if (((l + 1) % stab) == 0 && (l + 1) < L) {
            const int tau = l + 1, j = tau / stab, len = tau - tau_left;
            if (record_drift(R, tau)) { R->status = 1; goto done; }
            orc_build_Bblock(n, R->expK, R->expv, R->s, tau_left, len, Bb);
            if (rep_lmul(R, &R->left, Bb)) { R->status = 1; goto done; }
            tau_left = tau;
            if (carry) udv_cp(&R->A[j], &R->left, n);
            if (rebuild_from_factors(R, &R->left, &R->S[j])) goto done;
        }
    }
    {
        const int len = L - tau_left;
        if (record_drift(R, 0)) { R->status = 1; goto done; }
        if (len > 0) {
            orc_build_Bblock(n, R->expK, R->expv, R->s, tau_left, len, Bb);
            if (rep_lmul(R, &R->left, Bb)) { R->status = 1; goto done; }
        }
        if (carry) udv_cp(&R->A[M], &R->left, n);
        if (rebuild_from_left(R, &R->left)) goto done;
    }
    if (carry) {
        R->prefix_ok = 1;
        R->suffix_ok = 0;
        R->dir = 1;
    }
    R->sweeps++;
done:
    free(Bb);
    if (R->status) R->prefix_ok = R->suffix_ok = 0;
    return R->status;
}

/* dqmc_sweep_backward (src/dqmc.c:850-950) with dqmc_backward_boundary
   (:815-848): slices L-1 .. 0, G carried down by green_wrap_backward; at
   tau = j*stab the suffix UDV (B_{L-1} ... B_tau, accumulated by rmul) is
   stored as S[j] for the next forward sweep and G is rebuilt from the
   prefix A[j] of the previous forward sweep and that suffix. */
static int sweep_backward(orc_replica *R)
{
    const int n = R->n, L = R->L, stab = R->stab, M = R->M;
    double *Bb = malloc(sizeof(double) * (size_t)n * n);

    if (!R->prefix_ok) {
        /* green_stack_build_prefix (src/green.c:228-250) */
        udv_eye(&R->A[0], n);
        for (int j = 1; j <= M; j++) {
            orc_build_Bblock(n, R->expK, R->expv, R->s, R->b[j - 1], R->b[j] - R->b[j - 1], Bb);
            udv_cp(&R->A[j], &R->A[j - 1], n);
            if (rep_lmul(R, &R->A[j], Bb)) { R->status = 1; goto done; }
        }
    }
    udv_eye(&R->left, n);
    udv_eye(&R->S[M], n);
    orc_wrap_backward(n, R->expK, R->expKinv, R->expv, R->s + (size_t)(L - 1) * n, R->g);

    for (int l = L - 1; l >= 0; l--) {
        metro_slice(R, l);
        if (l == 0) break;
        if ((l % stab) == 0) {
            const int j = l / stab;
            if (record_drift(R, l)) { R->status = 1; goto done; }
            orc_build_Bblock(n, R->expK, R->expv, R->s, R->b[j], R->b[j + 1] - R->b[j], Bb);
            if (rep_rmul(R, &R->left, Bb)) { R->status = 1; goto done; }
            udv_cp(&R->S[j], &R->left, n);
            if (rebuild_from_factors(R, &R->A[j], &R->left)) goto done;
        }
        orc_wrap_backward(n, R->expK, R->expKinv, R->expv, R->s + (size_t)(l - 1) * n, R->g);
    }
    {
        const int len = R->b[1] - R->b[0];
        if (record_drift(R, 0)) { R->status = 1; goto done; }
        if (len > 0) {
            orc_build_Bblock(n, R->expK, R->expv, R->s, 0, len, Bb);
            if (rep_rmul(R, &R->left, Bb)) { R->status = 1; goto done; }
        }
        if (rebuild_from_left(R, &R->left)) goto done;
    }
    R->prefix_ok = 0;
    R->suffix_ok = 1;
    R->dir = 0;
    R->sweeps++;
done:
    free(Bb);
    if (R->status) R->prefix_ok = R->suffix_ok = 0;
    return R->status;
}

/* dqmc_sweep (src/dqmc.c:952-963) */
int orc_replica_sweep(orc_replica *R)
{
    if (R->status) return R->status;
    if (!R->alternating) return sweep_forward(R, 0);
    return R->dir ? sweep_backward(R) : sweep_forward(R, 1);
}

/* green_rebuild = combine (0), two_sided (1) or centered (2)
   (src/replica_run.c:37-45, src/green.c:116-119); the initial G is rebuilt
   in the new mode as dqmc_init_modes does (src/dqmc.c:464-466). */
int orc_replica_set_rebuild(orc_replica *R, int mode)
{
    R->rebuild = mode;
    R->status = 0;
    R->reason = 0;
    return orc_replica_rebuild(R);
}

int orc_replica_reason(const orc_replica *R) { return R->reason; }

/* sweep_order=alternating (src/dqmc.c:431-462): allocate the prefix stack;
   the next sweep is forward and rebuilds both stacks as it needs them. */
void orc_replica_set_alternating(orc_replica *R, int on)
{
    if (on && !R->A) {
        R->A = calloc((size_t)R->M + 1, sizeof(Udv));
        for (int j = 0; j <= R->M; j++) {
            udv_alloc(&R->A[j], R->n);
            udv_eye(&R->A[j], R->n);
        }
    }
    R->alternating = on ? 1 : 0;
    R->dir = 0;
    R->prefix_ok = R->suffix_ok = 0;
}

void orc_replica_green(const orc_replica *R, double *gu, double *gd)
{
    const size_t nn = (size_t)R->n * (size_t)R->n;
    if (gu) memcpy(gu, R->g, nn * sizeof(double));
    if (gd) memcpy(gd, R->gd, nn * sizeof(double));
}

void orc_replica_field(const orc_replica *R, signed char *s)
{
    memcpy(s, R->s, (size_t)R->L * (size_t)R->n);
}

void orc_replica_set_field(orc_replica *R, const signed char *s)
{
    memcpy(R->s, s, (size_t)R->L * (size_t)R->n);
    R->prefix_ok = R->suffix_ok = 0;
}

void orc_replica_stats(const orc_replica *R, unsigned long long *acc,
                       unsigned long long *att, double *sign, int *det_sign)
{
    *acc = R->acc;
    *att = R->att;
    *sign = R->sign;
    *det_sign = R->det_sign;
}

void orc_replica_rng_state(const orc_replica *R, uint64_t *s4)
{
    for (int i = 0; i < 4; i++) s4[i] = R->rng[i];
}

void orc_replica_model(const orc_replica *R, double *expK, double *expKinv,
                       double *lambda)
{
    const size_t nn = (size_t)R->n * (size_t)R->n;
    if (expK) memcpy(expK, R->expK, nn * sizeof(double));
    if (expKinv) memcpy(expKinv, R->expKinv, nn * sizeof(double));
    if (lambda) *lambda = R->lambda;
}

void orc_replica_drift(const orc_replica *R, double *max_inf,
                       double *mean_inf, unsigned long long *samples)
{
    *max_inf = R->drift_max;
    *samples = R->drift_samples;
    *mean_inf = R->drift_samples ? R->drift_sum / (double)R->drift_samples : 0.0;
}

/* Record the ratio of every proposal from now on (the last sweep's value
   per slice and site); out receives L * n values, slice-major. */
int orc_replica_log_ratios(orc_replica *R)
{
    if (!R->rlog) R->rlog = calloc((size_t)R->L * R->n, sizeof(double));
    return R->rlog ? 0 : 1;
}

void orc_replica_ratios(const orc_replica *R, double *out)
{
    if (R->rlog) memcpy(out, R->rlog, sizeof(double) * (size_t)R->L * R->n);
}

void orc_replica_free(orc_replica *R)
{
    if (!R) return;
    free(R->rlog);
    free(R->expK); free(R->expKinv); free(R->g); free(R->gd); free(R->C); free(R->V);
    free(R->bipart); free(R->b); free(R->s);
    for (int j = 0; j <= R->M; j++) udv_release(&R->S[j]);
    free(R->S);
    if (R->A) {
        for (int j = 0; j <= R->M; j++) udv_release(&R->A[j]);
        free(R->A);
    }
    udv_release(&R->left);
    udv_release(&R->comb);
    free(R);
}
