/*
 * exact.c -- the Green's function of a field configuration in extended
 * precision.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never loads it.
 *
 * Not a restatement of one of the reference's functions: this is the
 * quantity every FP64 path approximates,
 *
 *   G(tau) = (1 + B_{tau-1} ... B_0 B_{L-1} ... B_tau)^-1,
 *   B_l = expK diag(d_l),  d_{l,j} = expv[s_{l,j} > 0]
 *
 * (src/green.c:63-67, 126-137, 262-306), evaluated in the x87 80-bit long
 * double (64-bit significand, unit roundoff 5.4e-20 against FP64's 1.1e-16)
 * so that its own rounding is far below the FP64 paths' distances from it.
 * The chain is stabilised as the left UDV chain of green_from_scratch
 * (src/green.c:262-306): slices applied to U in groups of `blk`, then
 * M = X diag(D) factored by unblocked Householder QR, U = Q, D = diag R,
 * T = (D^-1 R) T; the inverse is G = N^-1 Db^-1 U^T with
 * N = Db^-1 U^T + Ds T (|Db| >= 1, |Ds| <= 1), N^-1 by Householder QR and
 * back substitution.  The FP64 inputs (expK, expv) are taken as exact.
 * How it is pinned: tests/test_oracle.py (mpmath at 50 digits, the U = 0
 * closed form, independence of the grouping).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef long double xr;
#define IX(i, j, n) ((size_t)(i) + (size_t)(j) * (size_t)(n))

/* C = A B (n x n, column-major) */
static void xmul(int n, const xr *A, const xr *B, xr *C)
{
    memset(C, 0, sizeof(xr) * (size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
            const xr b = B[IX(k, j, n)];
            if (b == 0) continue;
            const xr *a = A + IX(0, k, n);
            xr *c = C + IX(0, j, n);
            for (int i = 0; i < n; i++) c[i] += a[i] * b;
        }
}

/* Householder QR of A in place (R above, reflectors below, dlarfg signs);
   Q formed explicitly into Q. */
static void xqr(int n, xr *A, xr *Q)
{
    xr *tau = calloc((size_t)n, sizeof(xr));
    for (int j = 0; j < n; j++) {
        xr ss = 0;
        for (int i = j + 1; i < n; i++) ss += A[IX(i, j, n)] * A[IX(i, j, n)];
        const xr alpha = A[IX(j, j, n)];
        if (ss == 0) continue;
        const xr beta = -copysignl(sqrtl(alpha * alpha + ss), alpha);
        tau[j] = (beta - alpha) / beta;
        const xr sc = 1 / (alpha - beta);
        for (int i = j + 1; i < n; i++) A[IX(i, j, n)] *= sc;
        A[IX(j, j, n)] = beta;
        for (int k = j + 1; k < n; k++) {
            xr w = A[IX(j, k, n)];
            for (int i = j + 1; i < n; i++) w += A[IX(i, j, n)] * A[IX(i, k, n)];
            w *= tau[j];
            A[IX(j, k, n)] -= w;
            for (int i = j + 1; i < n; i++) A[IX(i, k, n)] -= A[IX(i, j, n)] * w;
        }
    }
    memset(Q, 0, sizeof(xr) * (size_t)n * n);
    for (int i = 0; i < n; i++) Q[IX(i, i, n)] = 1;
    for (int j = n - 1; j >= 0; j--) {
        if (tau[j] == 0) continue;
        for (int k = j; k < n; k++) {
            xr w = Q[IX(j, k, n)];
            for (int i = j + 1; i < n; i++) w += A[IX(i, j, n)] * Q[IX(i, k, n)];
            w *= tau[j];
            Q[IX(j, k, n)] -= w;
            for (int i = j + 1; i < n; i++) Q[IX(i, k, n)] -= A[IX(i, j, n)] * w;
        }
    }
    free(tau);
}

int orc_green_exact(int n, int L, int tau, int blk, const double *expK, const double expv[2],
                    const signed char *s, double *g)
{
    if (n <= 0 || L <= 0 || tau < 0 || tau >= L || blk <= 0) return 1;
    const size_t nn = (size_t)n * n;
    xr *K = malloc(sizeof(xr) * nn), *U = malloc(sizeof(xr) * nn), *T = malloc(sizeof(xr) * nn),
       *X = malloc(sizeof(xr) * nn), *Y = malloc(sizeof(xr) * nn), *Q = malloc(sizeof(xr) * nn),
       *D = malloc(sizeof(xr) * n);
    for (size_t k = 0; k < nn; k++) K[k] = expK[k];
    memset(U, 0, sizeof(xr) * nn);
    memset(T, 0, sizeof(xr) * nn);
    for (int i = 0; i < n; i++) {
        U[IX(i, i, n)] = 1;
        T[IX(i, i, n)] = 1;
        D[i] = 1;
    }
    int rc = 0;
    /* slices tau .. L-1, 0 .. tau-1, in groups of blk (green_from_scratch order) */
    for (int done = 0; done < L;) {
        const int take = blk < L - done ? blk : L - done;
        memcpy(X, U, sizeof(xr) * nn);
        for (int q = 0; q < take; q++) {
            const int l = (tau + done + q) % L;
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) X[IX(i, j, n)] *= (xr)expv[s[(size_t)l * n + i] > 0 ? 1 : 0];
            xmul(n, K, X, Y);
            memcpy(X, Y, sizeof(xr) * nn);
        }
        done += take;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) X[IX(i, j, n)] *= D[j];
        xqr(n, X, Q);
        memcpy(U, Q, sizeof(xr) * nn);
        for (int i = 0; i < n; i++) {
            D[i] = X[IX(i, i, n)];
            if (D[i] == 0 || !isfinite((double)D[i])) rc = 1;
        }
        /* T <- (D^-1 R) T */
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) Y[IX(i, j, n)] = i <= j ? X[IX(i, j, n)] / D[i] : 0;
        xmul(n, Y, T, X);
        memcpy(T, X, sizeof(xr) * nn);
    }
    if (!rc) {
        /* X = Db^-1 U^T, N = X + Ds T */
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const xr d = D[i];
                const int big = fabsl(d) > 1;
                const xr x = big ? U[IX(j, i, n)] / d : U[IX(j, i, n)];
                X[IX(i, j, n)] = x;
                Y[IX(i, j, n)] = x + (big ? 1 : d) * T[IX(i, j, n)];
            }
        xqr(n, Y, Q); /* Y = R_N (upper), Q = Q_N */
        /* W = Q^T X, then back substitution with R_N */
        xr *W = malloc(sizeof(xr) * nn);
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                xr acc = 0;
                for (int k = 0; k < n; k++) acc += Q[IX(k, i, n)] * X[IX(k, j, n)];
                W[IX(i, j, n)] = acc;
            }
        for (int j = 0; j < n; j++)
            for (int i = n - 1; i >= 0; i--) {
                xr acc = W[IX(i, j, n)];
                for (int k = i + 1; k < n; k++) acc -= Y[IX(i, k, n)] * W[IX(k, j, n)];
                W[IX(i, j, n)] = acc / Y[IX(i, i, n)];
            }
        for (size_t k = 0; k < nn; k++) {
            g[k] = (double)W[k];
            if (!isfinite(g[k])) rc = 1;
        }
        free(W);
    }
    free(K); free(U); free(T); free(X); free(Y); free(Q); free(D);
    return rc;
}
