// Single-site Metropolis over one imaginary-time slice with delayed
// updates, one workgroup per replica and one thread per site (forward PH
// sweep, src/dqmc.c:656-679, src/green.c:384-475).
//
// Sites are visited in windows of W consecutive sites.  At a window start
// thread j holds in registers G~(j, i0+w) and G~(i0+w, j) for the window's W
// sites (the window's columns and rows of the current Green's function), and
// every thread holds the W diagonal entries G~(i0+w, i0+w).  An accepted flip
// at site i = i0+w forms the reference's delayed-update pair
// c = (N/R)(e_i - G~ e_i), v = G~^T e_i (green_delay_accept,
// src/green.c:418-460) straight from those registers, publishes it to LDS,
// and after one barrier every thread applies the rank-1 change to its
// remaining window registers: no global traffic and no O(k) loop on the
// sequential path.  At the window end the accumulated rank-p update
// G -= C V^T (green_delay_flush / la_gemm_nt_rect, src/green.c:462-475,
// src/linalg.c:209-219) is applied to G in global memory with MFMA 16x16x4
// tiles.  The window size W is the delay depth k.
//
// The ratio is R = (1 + (1 - G~_ii) N_up)(1 + G~_ii N_dn)
// (src/green.c:401-416), always evaluated in FP64, accepted iff
// rng_double() < |R| (src/dqmc.c:668).  Every proposal consumes one draw of
// the replica's xoshiro256** stream; the slice's n draws are taken up front.
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

#define METRO_THREADS 512
#define METRO_MAX_N 512

template <typename T>
static inline size_t metro_lds_bytes(int n, int W)
{
    return sizeof(T) * ((size_t)2 * W * n + 4 * W) + sizeof(double) * n + 2 * (size_t)n + 64;
}

// G[n x n] -= Cc[0:p]^T Vv[0:p] with Cc, Vv stored q-major (Cc[q*n + j]).
// Each MFMA tile holds a 16x16 block of G^T (output row = G column, output
// column = G row), so the 16 lanes of an accumulator row cover 16
// consecutive rows of one column of G: every load and store instruction
// touches four 128-byte segments.  B tiles per wave are kept in flight.
template <typename T>
__device__ void flush_rank_p(T *G, int n, const T *Cc, const T *Vv, int p)
{
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const int nt = (n + 15) >> 4;
    const int ntiles = nt * nt;
    constexpr int B = sizeof(T) == 8 ? 8 : 16;
    for (int t0 = wave * B; t0 < ntiles; t0 += nw * B) {
        acc_t acc[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int t = t0 + u;
            const int r0 = (t % nt) * 16, c0 = (t / nt) * 16;
            const int row = r0 + lc;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int col = c0 + Mfma<T>::row(lane, v);
                acc[u][v] = (t < ntiles && row < n && col < n) ? G[row + (long long)col * n] : T(0);
            }
        }
        for (int q0 = 0; q0 < p; q0 += 4) {
            const int q = q0 + lr;
#pragma unroll
            for (int u = 0; u < B; u++) {
                const int t = t0 + u;
                const int r0 = (t % nt) * 16, c0 = (t / nt) * 16;
                const T a = (q < p && c0 + lc < n) ? -Vv[q * n + c0 + lc] : T(0);
                const T b = (q < p && r0 + lc < n) ? Cc[q * n + r0 + lc] : T(0);
                acc[u] = Mfma<T>::mma(a, b, acc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int t = t0 + u;
            const int r0 = (t % nt) * 16, c0 = (t / nt) * 16;
            const int row = r0 + lc;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int col = c0 + Mfma<T>::row(lane, v);
                if (t < ntiles && row < n && col < n) G[row + (long long)col * n] = acc[u][v];
            }
        }
    }
}

template <typename T, int W>
__global__ __launch_bounds__(METRO_THREADS) void k_metro_slice(MetroArgs<T> a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status[b]) return;
    const int n = a.n, j = threadIdx.x;
    const bool act = j < n;
    T *G = a.G + (long long)b * a.sG;
    signed char *fld = a.field + (long long)b * a.sF + (long long)a.l * n;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *Cc = (T *)lds_raw;        // W x n pending columns c_q (also the row staging area)
    T *Vv = Cc + (size_t)W * n;  // W x n pending rows v_q
    // the accepted pair at the window's sites, double-buffered by the parity
    // of the accept count so a fast thread never overwrites values a slow one
    // is still reading
    T *cwin = Vv + (size_t)W * n;
    T *vwin = cwin + 2 * W;
    double *draws = (double *)(vwin + 2 * W);
    signed char *fs = (signed char *)(draws + n);
    signed char *flip = fs + n;

    if (act) {
        fs[j] = fld[j];
        flip[j] = 0;
    }
    if (j == 0) {
        uint64_t st[4];
        uint64_t *rs = a.rng + 4 * (long long)b;
        for (int q = 0; q < 4; q++) st[q] = rs[q];
        for (int i = 0; i < n; i++) draws[i] = xoshiro_double(st);
        for (int q = 0; q < 4; q++) rs[q] = st[q];
    }
    __syncthreads();

    int naccept = 0, negflips = 0;
    T gc[W], gr[W], dW[W];
    for (int i0 = 0; i0 < n; i0 += W) {
        const int wn = min(W, n - i0);
        // window columns (coalesced over j) and rows (staged through LDS)
#pragma unroll
        for (int w = 0; w < W; w++) gc[w] = (act && w < wn) ? G[j + (long long)(i0 + w) * n] : T(0);
        for (int base = threadIdx.x; base < wn * n; base += 8 * blockDim.x) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * blockDim.x;
                v[u] = idx < wn * n ? G[(i0 + idx % wn) + (long long)(idx / wn) * n] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * blockDim.x;
                if (idx < wn * n) Cc[(idx % wn) * n + idx / wn] = v[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < W; w++) {
            gr[w] = (act && w < wn) ? Cc[w * n + j] : T(0);
            dW[w] = (w < wn) ? Cc[w * n + i0 + w] : T(0);
        }
        __syncthreads();
        int p = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (w < wn) {
                const int i = i0 + w;
                const double gii = (double)dW[w];
                const int sp = fs[i] > 0;
                const double Nu = a.nu[sp], Nd = a.nd[sp];
                const double Ru = 1.0 + (1.0 - gii) * Nu;
                const double Rd = 1.0 + gii * Nd;
                const double R = Ru * Rd;
                if (draws[i] < fabs(R)) {
                    const T fac = (T)(Nu / Ru);
                    const T cj = fac * ((j == i ? T(1) : T(0)) - gc[w]);
                    const T vj = gr[w];
                    if (act) {
                        Cc[p * n + j] = cj;
                        Vv[p * n + j] = vj;
                        if (j >= i0 && j < i0 + wn) {
                            cwin[(p & 1) * W + j - i0] = cj;
                            vwin[(p & 1) * W + j - i0] = vj;
                        }
                    }
                    if (j == 0) flip[i] = 1;
                    naccept++;
                    if (R < 0.0) negflips ^= 1;
                    __syncthreads();
#pragma unroll
                    for (int w2 = w + 1; w2 < W; w2++) {
                        if (w2 < wn) {
                            const T cw = cwin[(p & 1) * W + w2];
                            const T vw = vwin[(p & 1) * W + w2];
                            gc[w2] -= cj * vw;
                            gr[w2] -= cw * vj;
                            dW[w2] -= cw * vw;
                        }
                    }
                    p++;
                }
            }
        }
        if (p > 0) {
            __syncthreads();
            flush_rank_p<T>(G, n, Cc, Vv, p);
        }
        __syncthreads();
    }
    T *dv = a.dvec + (long long)b * a.sV + (long long)a.l * n;
    T *di = a.dinv + (long long)b * a.sV + (long long)a.l * n;
    if (act) {
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
    if (j == 0) {
        a.acc[b] += (unsigned long long)naccept;
        a.att[b] += (unsigned long long)n;
        if (negflips) a.sign[b] = -a.sign[b];
    }
}
