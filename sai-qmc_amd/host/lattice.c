/*
 * lattice.c -- lattice, model exponential, replica seeds and jackknife for
 * the GPU driver (host side of the path; no device code here).
 */
#include "host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void bond(Lattice *L, int i, int j, double t)
{
    L->t[i + (size_t)j * L->n] += t;
    L->t[j + (size_t)i * L->n] += t;
}

/* lattice_chain / lattice_square (src/lattice.c:26-82): nearest-neighbour
   bonds to +x (and +y), wrapped when pbc and the extent exceeds 1;
   sublattice sign (-1)^(x+y); bipartite unless a periodic extent is odd. */
int lattice_build(Lattice *L, const Params *p)
{
    memset(L, 0, sizeof(*L));
    const int chain = strcmp(p->lattice, "chain") == 0;
    if (!chain && strcmp(p->lattice, "square") != 0) return 1;
    const int Lx = p->Lx, Ly = chain ? 1 : p->Ly;
    if (Lx <= 0 || Ly <= 0) return 1;
    L->n = Lx * Ly;
    L->Lx = Lx;
    L->Ly = Ly;
    L->is_chain = chain;
    L->t = calloc((size_t)L->n * L->n, sizeof(double));
    L->bipart = calloc((size_t)L->n, sizeof(int));
    if (!L->t || !L->bipart) return 1;
    for (int y = 0; y < Ly; y++)
        for (int x = 0; x < Lx; x++) {
            const int i = x + y * Lx;
            if (x + 1 < Lx) bond(L, i, (x + 1) + y * Lx, p->thop);
            else if (p->pbc && Lx > 1) bond(L, i, 0 + y * Lx, p->thop);
            if (!chain) {
                if (y + 1 < Ly) bond(L, i, x + (y + 1) * Lx, p->thop);
                else if (p->pbc && Ly > 1) bond(L, i, x, p->thop);
            }
            L->bipart[i] = ((x + y) % 2 == 0) ? 1 : -1;
        }
    const int odd = (p->pbc && Lx > 1 && Lx % 2) || (!chain && p->pbc && Ly > 1 && Ly % 2);
    L->is_bipartite = !odd;
    return 0;
}

void lattice_release(Lattice *L)
{
    free(L->t);
    free(L->bipart);
    memset(L, 0, sizeof(*L));
}

/* exp(scale*A) for symmetric A.  la_expm_sym (src/linalg.c:319-346) goes
   through dsyev; here it is the Taylor series with scaling and squaring
   (Moler & Van Loan, "Nineteen dubious ways to compute the exponential of
   a matrix", method 3): X = scale*A / 2^s with ||X||_1 <= 1/4, 18 terms
   (truncation below 1e-27), then s squarings, then symmetrised.  Agrees
   with the eigendecomposition route to a few ulps at the sizes used. */
static void matmul(int n, const double *a, const double *b, double *c)
{
    for (int j = 0; j < n; j++) {
        double *cj = c + (size_t)j * n;
        for (int i = 0; i < n; i++) cj[i] = 0.0;
        for (int k = 0; k < n; k++) {
            const double bkj = b[k + (size_t)j * n];
            const double *ak = a + (size_t)k * n;
            for (int i = 0; i < n; i++) cj[i] += ak[i] * bkj;
        }
    }
}

int expm_sym(int n, const double *A, double scale, double *out)
{
    const size_t nn = (size_t)n * n;
    double *x = malloc(sizeof(double) * nn), *term = malloc(sizeof(double) * nn),
           *tmp = malloc(sizeof(double) * nn);
    if (!x || !term || !tmp) {
        free(x); free(term); free(tmp);
        return 1;
    }
    double norm1 = 0.0;
    for (int j = 0; j < n; j++) {
        double c = 0.0;
        for (int i = 0; i < n; i++) c += fabs(scale * A[i + (size_t)j * n]);
        if (c > norm1) norm1 = c;
    }
    int sq = 0;
    while (norm1 > 0.25 && sq < 60) {
        norm1 *= 0.5;
        sq++;
    }
    const double f = scale / ldexp(1.0, sq);
    for (size_t i = 0; i < nn; i++) {
        x[i] = f * A[i];
        term[i] = (i % (size_t)(n + 1) == 0) ? 1.0 : 0.0;
        out[i] = term[i];
    }
    for (int k = 1; k <= 18; k++) { /* term <- term * X / k, out += term */
        matmul(n, term, x, tmp);
        for (size_t i = 0; i < nn; i++) {
            term[i] = tmp[i] / (double)k;
            out[i] += term[i];
        }
    }
    for (int k = 0; k < sq; k++) {
        matmul(n, out, out, tmp);
        memcpy(out, tmp, sizeof(double) * nn);
    }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) {
            const double m = 0.5 * (out[i + (size_t)j * n] + out[j + (size_t)i * n]);
            out[i + (size_t)j * n] = m;
            out[j + (size_t)i * n] = m;
        }
    int bad = 0;
    for (size_t i = 0; i < nn; i++) bad |= !isfinite(out[i]);
    free(x); free(term); free(tmp);
    return bad;
}

static unsigned long long splitmix64(unsigned long long x)
{
    unsigned long long z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* replica_seed (src/replica.c:17-34): replica 0 keeps the legacy
   base + 1000*beta_index; the others mix base, beta index and replica id
   through SplitMix64, avoiding a collision with the legacy seed. */
unsigned long long replica_seed(unsigned long long base, int beta_index, int replica_id)
{
    const unsigned long long legacy = base + 1000ULL * (unsigned long long)beta_index;
    if (replica_id == 0) return legacy;
    unsigned long long x = base;
    x ^= 0xD1B54A32D192ED03ULL * (unsigned long long)(unsigned int)(beta_index + 1);
    x ^= 0xABC98388FB8FAC03ULL * (unsigned long long)(unsigned int)(replica_id + 1);
    unsigned long long m = splitmix64(x);
    if (m == legacy) m = splitmix64(m);
    return m;
}

/* jackknife (src/measure.c:48-69): leave-one-out means, error
   sqrt((N-1)/N * sum (mean_i - mean)^2). */
void jackknife(const double *x, int N, double *mean, double *err)
{
    double sum = 0.0;
    for (int i = 0; i < N; i++) sum += x[i];
    const double mu = N > 0 ? sum / (double)N : 0.0;
    *mean = mu;
    if (N < 2) {
        *err = 0.0;
        return;
    }
    double var = 0.0;
    for (int i = 0; i < N; i++) {
        const double d = (sum - x[i]) / (double)(N - 1) - mu;
        var += d * d;
    }
    *err = sqrt(var * (double)(N - 1) / (double)N);
}
