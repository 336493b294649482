// Single-site Metropolis over one imaginary-time slice with delayed
// updates, one workgroup per replica (forward PH sweep,
// src/dqmc.c:656-679, src/green.c:384-475).
//
// Sites are visited in windows of W consecutive sites.  At a window start
// the window's W columns and W rows of G are staged in LDS; every accepted
// flip appends one column pair (c, v) to the pending rank-p update
// G~ = G - C V^T (the reference's delay_c / delay_v, src/green.c:418-460),
// built entirely from LDS (no global traffic on the sequential path).  At
// the window end the rank-p update is flushed into G in global memory
// (la_gemm_nt_rect, src/green.c:462-475).  The window size W plays the
// role of the delay depth k.
//
// The ratio of site i uses R = (1 + (1 - G~_ii) N_up)(1 + G~_ii N_dn)
// (src/green.c:401-416) and the test rng_double() < |R| (src/dqmc.c:668),
// always in FP64.  One RNG draw per proposal: the slice's n draws are taken
// up front from the replica's xoshiro256** state (draws never depend on
// acceptances).
#pragma once
#include "common.hpp"

template <typename T> struct MetroArgs {
    int n, L, W, l, batch;
    T *G;
    long long sG;
    signed char *field; // [L][n] per replica
    long long sF;
    T *dvec, *dinv;     // [L][n] per replica: expv[s>0], 1/expv[s>0]
    long long sV;
    T e[2], einv[2];
    double *dvec64, *dinv64; // FP64 copies kept in sync when T is float (may be null)
    double e64[2], einv64[2];
    uint64_t *rng;      // [4] per replica
    double nu[2], nd[2];
    unsigned long long *acc, *att;
    double *sign;
    int *status;
};

template <typename T>
static inline size_t metro_lds_bytes(int n, int W)
{
    return sizeof(T) * (size_t)4 * W * n + sizeof(double) * n + 2 * (size_t)n + sizeof(T) * (W + 4) + 64;
}

template <typename T>
__global__ __launch_bounds__(256) void k_metro_slice(MetroArgs<T> a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status[b]) return;
    const int n = a.n, W = a.W, tid = threadIdx.x, nt = blockDim.x;
    T *G = a.G + (long long)b * a.sG;
    signed char *fld = a.field + (long long)b * a.sF + (long long)a.l * n;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *Gc = (T *)lds_raw;       // W x n: Gc[w*n + j] = G(j, i0+w)
    T *Gr = Gc + (size_t)W * n;  // W x n: Gr[w*n + j] = G(i0+w, j)
    T *Cc = Gr + (size_t)W * n;  // W x n: pending columns c_q
    T *Vv = Cc + (size_t)W * n;  // W x n: pending rows v_q
    double *draws = (double *)(Vv + (size_t)W * n);
    signed char *fs = (signed char *)(draws + n);
    signed char *flip = fs + n;
    T *dW = (T *)(((uintptr_t)(flip + n) + 15) & ~(uintptr_t)15);

    for (int j = tid; j < n; j += nt) {
        fs[j] = fld[j];
        flip[j] = 0;
    }
    if (tid == 0) {
        uint64_t st[4];
        uint64_t *rs = a.rng + 4 * (long long)b;
        for (int q = 0; q < 4; q++) st[q] = rs[q];
        for (int i = 0; i < n; i++) draws[i] = xoshiro_double(st);
        for (int q = 0; q < 4; q++) rs[q] = st[q];
    }
    __syncthreads();

    int naccept = 0, negflips = 0;
    for (int i0 = 0; i0 < n; i0 += W) {
        const int wn = min(W, n - i0);
        for (int idx = tid; idx < wn * n; idx += nt) {
            const int w = idx / n, j = idx % n;
            Gc[w * n + j] = G[j + (long long)(i0 + w) * n];
        }
        for (int idx = tid; idx < wn * n; idx += nt) {
            const int w = idx % wn, j = idx / wn;
            Gr[w * n + j] = G[(i0 + w) + (long long)j * n];
        }
        __syncthreads();
        if (tid < wn) dW[tid] = Gc[tid * n + i0 + tid];
        __syncthreads();
        int p = 0;
        for (int w = 0; w < wn; w++) {
            const int i = i0 + w;
            const double gii = (double)dW[w];
            const int sp = fs[i] > 0;
            const double Nu = a.nu[sp], Nd = a.nd[sp];
            const double Ru = 1.0 + (1.0 - gii) * Nu;
            const double Rd = 1.0 + gii * Nd;
            const double R = Ru * Rd;
            if (draws[i] < fabs(R)) {
                const T fac = (T)(Nu / Ru);
                for (int j = tid; j < n; j += nt) {
                    T gji = Gc[w * n + j], gij = Gr[w * n + j];
                    for (int q = 0; q < p; q++) {
                        gji -= Cc[q * n + j] * Vv[q * n + i];
                        gij -= Cc[q * n + i] * Vv[q * n + j];
                    }
                    const T cj = fac * ((j == i ? T(1) : T(0)) - gji);
                    Cc[p * n + j] = cj;
                    Vv[p * n + j] = gij;
                    if (j > i && j < i0 + wn) dW[j - i0] -= cj * gij;
                }
                if (tid == 0) flip[i] = 1;
                naccept++;
                if (R < 0.0) negflips ^= 1;
                p++;
                __syncthreads();
            }
        }
        if (p > 0) {
            for (long long idx = tid; idx < (long long)n * n; idx += nt) {
                const int j = (int)(idx % n), m = (int)(idx / n);
                T g = G[idx];
                for (int q = 0; q < p; q++) g -= Cc[q * n + j] * Vv[q * n + m];
                G[idx] = g;
            }
            __syncthreads();
        }
    }
    T *dv = a.dvec + (long long)b * a.sV + (long long)a.l * n;
    T *di = a.dinv + (long long)b * a.sV + (long long)a.l * n;
    for (int j = tid; j < n; j += nt) {
        const signed char s = flip[j] ? (signed char)-fs[j] : fs[j];
        fld[j] = s;
        dv[j] = a.e[s > 0];
        di[j] = a.einv[s > 0];
        if (a.dvec64) {
            const long long o = (long long)b * a.sV + (long long)a.l * n + j;
            a.dvec64[o] = a.e64[s > 0];
            a.dinv64[o] = a.einv64[s > 0];
        }
    }
    if (tid == 0) {
        a.acc[b] += (unsigned long long)naccept;
        a.att[b] += (unsigned long long)n;
        if (negflips) a.sign[b] = -a.sign[b];
    }
}
