// Single-site Metropolis over one segment of sites with the segment's
// submatrix of G held in registers, FP64 (forward PH sweep,
// src/dqmc.c:656-679, src/green.c:384-475).
//
// The ratio of site i needs only G~_ii, and the change of an accept at
// site a,
//   G~ <- G~ - c v^T,  c = f (e_a - G~ e_a),  v = G~^T e_a,  f = N_up / R_up
// (green_delay_accept, src/green.c:418-460), restricted to the rows and
// columns of the segment's sites S = [s0, s1) needs only G~[S, S].  So a
// segment is decided on its m x m submatrix alone (k_metro_sub): loaded
// once into registers, each accept applies its rank-1 change to it, and
// nothing of G outside S is read.  The change to all of G is then one
// rank-p update over the accepted sites A in accept order
// (Sherman-Morrison-Woodbury):
//   G' = G - (E_A - G[:, A]) K G[A, :],  K = Delta_A N^-1,
//   N = I + (I - G[A, A]) Delta_A,
// Delta_A the N_up of the accepted sites.  Gaussian elimination of N in
// accept order divides by exactly the R_up of the sequential updates (its
// leading minors are their products), so this is the sequential result
// reassociated.  k_metro_sub forms K by Gauss-Jordan elimination of N
// without pivoting, in registers; k_metro_pairs writes the pending pairs
//   C[q][:] = e_{a_q} - G[:, a_q],  V[q][:] = sum_r K[q][r] G[a_r, :]
// in the q-major layout of k_metro_slice, and the flush G -= C^T V that
// follows (k_flush64, the batched GEMM, or the wrap's fused column pass) is
// the one of the windowed path.
//
// Register layout of an m x m matrix (the submatrix, later N): entry (r, c)
// sits in lane r % 64, row slot r / 64, of wave c % 8, column slot c / 8.
// Row r of the columns of one wave is therefore in a single lane (read
// with readlane); a column is in a single wave and goes to the others
// through LDS.  Every wave keeps its own copy of the diagonal (lane r % 64
// holds G~_rr) and evaluates every proposal: the copies receive identical
// updates, so all waves take the same decisions, and an accept costs one
// workgroup barrier while a reject costs none.
#pragma once
#include "common.hpp"

#define MSUB_MAX 96       // largest segment (sites) of the submatrix path
#define MSUB_THREADS 512  // 8 waves
#define MP_COLS 64        // columns of G per k_metro_pairs workgroup

struct MetroSubArgs {
    int n, L, l, batch;
    int s0, s1;               // sites [s0, s1) of the slice, s1 - s0 <= MSUB_MAX
    const double *G;
    long long sG;
    signed char *field;       // [L][n] per replica
    long long sF;
    double *dvec, *dinv;      // [L][n] per replica
    long long sV;
    double e[2], einv[2];
    uint64_t *rng;            // [4] per replica
    double nu[2], nd[2];
    unsigned long long *acc, *att;
    double *sign;
    int *status;
    double *K;                // out: K, p x p row-major, leading dimension m
    long long sK;
    int *aq;                  // out: accepted sites, offsets from s0, in accept order
    long long sA;
    int *pcount;              // out: p
    double *rlog;             // null, or the ratio of every proposal, n per replica
};

// LDS: the submatrix at entry m x (m+1), column and row buffers 2 x 2 x
// MSUB_MAX, the segment's draws, N_up and site of each accept, spins and
// flips.
static inline size_t metro_sub_lds_bytes(int m)
{
    return sizeof(double) * ((size_t)m * (m + 1) + 6 * MSUB_MAX) + sizeof(int) * MSUB_MAX + 2 * MSUB_MAX + 64;
}

__device__ __forceinline__ double sub_bcast(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

template <int RS, int CS>
__global__ __launch_bounds__(MSUB_THREADS) void k_metro_sub(MetroSubArgs a)
{
    static_assert(RS * 64 >= 8 * CS && 8 * CS <= MSUB_MAX, "segment layout");
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status[b]) {
        if (threadIdx.x == 0) a.pcount[b] = 0;
        return;
    }
    const int n = a.n, s0 = a.s0, m = a.s1 - a.s0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double *G = a.G + (long long)b * a.sG;
    signed char *fld = a.field + (long long)b * a.sF + (long long)a.l * n;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int ldo = m + 1;
    double *Go = (double *)lds_raw;       // Go[c * ldo + r] = G[s0 + r][s0 + c] at entry
    double *colb = Go + (size_t)m * ldo;  // 2 x MSUB_MAX
    double *rowb = colb + 2 * MSUB_MAX;   // 2 x MSUB_MAX
    double *draws = rowb + 2 * MSUB_MAX;  // MSUB_MAX
    double *dlt = draws + MSUB_MAX;       // N_up of accept q
    int *acc_site = (int *)(dlt + MSUB_MAX);
    signed char *fs = (signed char *)(acc_site + MSUB_MAX);
    signed char *fl = fs + MSUB_MAX;

    // the submatrix and the diagonal, from clamped (always valid) entries;
    // entries past m are never stored and never read by a valid entry.  The
    // loads are issued first, so they arrive while thread 0 draws.
    double g[RS][CS], dg[RS];
#pragma unroll
    for (int s = 0; s < RS; s++) {
        const int r = min(lane + 64 * s, m - 1);
#pragma unroll
        for (int c = 0; c < CS; c++) g[s][c] = G[(s0 + r) + (long long)(s0 + min(wave + 8 * c, m - 1)) * n];
        dg[s] = G[(s0 + r) + (long long)(s0 + r) * n];
    }
    if (tid < m) {
        fs[tid] = fld[s0 + tid];
        fl[tid] = 0;
    }
    if (tid == 0) {
        uint64_t st[4];
        uint64_t *rs = a.rng + 4 * (long long)b;
        for (int q = 0; q < 4; q++) st[q] = rs[q];
        for (int i = 0; i < m; i++) draws[i] = xoshiro_double(st);
        for (int q = 0; q < 4; q++) rs[q] = st[q];
    }
#pragma unroll
    for (int s = 0; s < RS; s++)
#pragma unroll
        for (int c = 0; c < CS; c++) {
            const int r = lane + 64 * s, cc = wave + 8 * c;
            if (r < m && cc < m) Go[cc * ldo + r] = g[s][c];
        }
    __syncthreads();
    double dr[RS];
    int sp[RS];
#pragma unroll
    for (int s = 0; s < RS; s++) {
        const int r = min(lane + 64 * s, m - 1);
        dr[s] = draws[r];
        sp[s] = fs[r] > 0;
    }

    // 1. proposals in site order (sites 8c + w: column slot c of wave w)
    int p = 0, negflips = 0;
#pragma unroll
    for (int c = 0; c < CS; c++) {
        const int si = c / 8; // row slot of sites 8c .. 8c + 7
        const int wend = min(8, m - 8 * c); // sites of this slot (none past m)
        for (int w = 0; w < wend; w++) {
            const int i = 8 * c + w;
            const int li = i & 63;
            const double gii = sub_bcast(dg[si], li);
            const double dri = sub_bcast(dr[si], li);
            const int spi = __builtin_amdgcn_readlane(sp[si], li);
            const double Nu = spi ? a.nu[1] : a.nu[0];
            const double Nd = spi ? a.nd[1] : a.nd[0];
            const double Ru = 1.0 + (1.0 - gii) * Nu;
            const double Rd = 1.0 + gii * Nd;
            const double R = Ru * Rd;
            if (a.rlog && tid == 0) a.rlog[(long long)b * n + s0 + i] = R;
            if (dri < fabs(R)) {
                const double f = Nu / Ru;
                double *cb = colb + (p & 1) * MSUB_MAX, *rb = rowb + (p & 1) * MSUB_MAX;
                if (wave == w) {
#pragma unroll
                    for (int s = 0; s < RS; s++)
                        if (lane + 64 * s < m) cb[lane + 64 * s] = g[s][c];
                }
                if (lane == li) {
#pragma unroll
                    for (int c2 = 0; c2 < CS; c2++)
                        if (wave + 8 * c2 < m) rb[wave + 8 * c2] = g[si][c2];
                }
                if (tid == 0) {
                    acc_site[p] = i;
                    dlt[p] = Nu;
                    fl[i] = 1;
                }
                __syncthreads();
                double cs[RS], vd[RS], vr[CS];
#pragma unroll
                for (int s = 0; s < RS; s++) {
                    const int r = lane + 64 * s, rr = min(r, m - 1);
                    cs[s] = f * ((r == i ? 1.0 : 0.0) - cb[rr]);
                    vd[s] = rb[rr];
                }
#pragma unroll
                for (int c2 = 0; c2 < CS; c2++) vr[c2] = rb[min(wave + 8 * c2, m - 1)];
#pragma unroll
                for (int s = 0; s < RS; s++) {
#pragma unroll
                    for (int c2 = 0; c2 < CS; c2++) g[s][c2] -= cs[s] * vr[c2];
                    dg[s] -= cs[s] * vd[s];
                }
                p++;
                if (R < 0.0) negflips ^= 1;
            }
        }
    }
    __syncthreads();

    // 2. K = Delta_A N^-1 by Gauss-Jordan elimination of N in accept order:
    // N[q][r] = [q = r] + ([q = r] - G[a_q][a_r]) Delta_r
    if (p > 0) {
#pragma unroll
        for (int s = 0; s < RS; s++)
#pragma unroll
            for (int c = 0; c < CS; c++) {
                const int q = lane + 64 * s, r = wave + 8 * c;
                const int aqq = acc_site[min(q, p - 1)], arr = acc_site[min(r, p - 1)];
                const double dq = q == r ? 1.0 : 0.0;
                g[s][c] = dq + (dq - Go[arr * ldo + aqq]) * dlt[min(r, p - 1)];
            }
#pragma unroll
        for (int c = 0; c < CS; c++) {
            const int sk = c / 8;
            const int wend = min(8, p - 8 * c);
            for (int w = 0; w < wend; w++) {
                const int k = 8 * c + w;
                const int lk = k & 63;
                double *cb = colb + (k & 1) * MSUB_MAX;
                if (wave == w) {
#pragma unroll
                    for (int s = 0; s < RS; s++)
                        if (lane + 64 * s < p) cb[lane + 64 * s] = g[s][c];
                }
                __syncthreads();
                const double rp = 1.0 / cb[k];
                double t[RS], rk[CS];
#pragma unroll
                for (int s = 0; s < RS; s++) t[s] = cb[min(lane + 64 * s, p - 1)] * rp;
#pragma unroll
                for (int c2 = 0; c2 < CS; c2++) rk[c2] = sub_bcast(g[sk][c2], lk);
#pragma unroll
                for (int s = 0; s < RS; s++)
#pragma unroll
                    for (int c2 = 0; c2 < CS; c2++) {
                        const int q = lane + 64 * s, r = wave + 8 * c2;
                        double v = g[s][c2] - t[s] * rk[c2];
                        if (q == k) v = rk[c2] * rp;
                        if (r == k) v = q == k ? rp : -t[s];
                        g[s][c2] = v;
                    }
            }
        }
        double *K = a.K + (long long)b * a.sK;
#pragma unroll
        for (int s = 0; s < RS; s++) {
            const int q = lane + 64 * s;
            const double dq = dlt[min(q, p - 1)];
#pragma unroll
            for (int c = 0; c < CS; c++) {
                const int r = wave + 8 * c;
                if (q < p && r < p) K[(long long)q * m + r] = dq * g[s][c];
            }
        }
        if (tid < p) a.aq[(long long)b * a.sA + tid] = acc_site[tid];
    }

    double *dv = a.dvec + (long long)b * a.sV + (long long)a.l * n;
    double *di = a.dinv + (long long)b * a.sV + (long long)a.l * n;
    if (tid < m) {
        const signed char s = fl[tid] ? (signed char)-fs[tid] : fs[tid];
        fld[s0 + tid] = s;
        dv[s0 + tid] = a.e[s > 0];
        di[s0 + tid] = a.einv[s > 0];
    }
    if (tid == 0) {
        a.pcount[b] = p;
        a.acc[b] += (unsigned long long)p;
        a.att[b] += (unsigned long long)m;
        if (negflips) a.sign[b] = -a.sign[b];
    }
}

struct MetroPairsArgs {
    int n, batch, s0, m;
    const double *G;
    long long sG;
    const double *K;   // p x p row-major, leading dimension m
    long long sK;
    const int *aq;     // accepted sites (offsets from s0)
    long long sA;
    const int *pcount;
    const int *status;
    double *C, *V;     // out: q-major pending pairs, row q at q * n
    long long sP;
};

// C[q][j] = [j = s0 + a_q] - G[j][s0 + a_q] and V[q][j] = sum_r K[q][r]
// G[s0 + a_r][j] for the 64 columns (rows) j of one workgroup.  The
// segment's rows of those columns are staged in LDS with coalesced loads
// (the rows a_r lie in [s0, s0 + m)); V on FP64 MFMA 16x16x4, wave w taking
// columns 16w .. 16w + 15 and every 16-row tile of q, r four at a time in
// increasing order (A = K[q][r], B = G[s0 + a_r][j]).
__global__ __launch_bounds__(256) void k_metro_pairs(MetroPairsArgs a)
{
    const int nblk = (a.n + MP_COLS - 1) / MP_COLS;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / nblk, j0 = (logical % nblk) * MP_COLS;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int p = a.pcount[b];
    if (p <= 0) return;
    const int n = a.n, m = a.m, s0 = a.s0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const double *G = a.G + (long long)b * a.sG;
    const double *K = a.K + (long long)b * a.sK;
    double *C = a.C + (long long)b * a.sP, *V = a.V + (long long)b * a.sP;
    __shared__ int s_aq[MSUB_MAX];
    __shared__ double Ys[MSUB_MAX * (MP_COLS + 1)];
    if (tid < p) s_aq[tid] = a.aq[(long long)b * a.sA + tid];
    for (int i0 = tid; i0 < m * MP_COLS; i0 += 4 * 256) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int idx = min(i0 + 256 * u, m * MP_COLS - 1), i = idx % m, jj = idx / m;
            x[u] = G[(s0 + i) + (long long)min(j0 + jj, n - 1) * n];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int idx = i0 + 256 * u;
            if (idx < m * MP_COLS) Ys[(idx % m) * (MP_COLS + 1) + idx / m] = x[u];
        }
    }
    __syncthreads();

    // C: wave w takes q = w, w + 4, ..., four rows in flight
    {
        const int j = j0 + lane, jc = min(j, n - 1);
        for (int q0 = wave; q0 < p; q0 += 16) {
            double x[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int col = s0 + s_aq[min(q0 + 4 * u, p - 1)];
                x[u] = G[jc + (long long)col * n];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + 4 * u;
                if (q < p && j < n) {
                    const int col = s0 + s_aq[q];
                    C[(long long)q * n + j] = (j == col ? 1.0 : 0.0) - x[u];
                }
            }
        }
    }

    // V
    constexpr int QT = MSUB_MAX / 16;
    const int nqt = (p + 15) >> 4;
    f64x4 acc[QT];
#pragma unroll
    for (int qt = 0; qt < QT; qt++) acc[qt] = f64x4{0, 0, 0, 0};
    const int jl = 16 * wave + lc;
    for (int r0 = 0; r0 < p; r0 += 16) {
        double bv[4], av[4][QT];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = r0 + 4 * u + lr, rc = min(r, p - 1);
            bv[u] = Ys[s_aq[rc] * (MP_COLS + 1) + jl];
#pragma unroll
            for (int qt = 0; qt < QT; qt++)
                av[u][qt] = qt < nqt ? K[(long long)min(16 * qt + lc, p - 1) * m + rc] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double b_ = r0 + 4 * u + lr < p ? bv[u] : 0.0;
#pragma unroll
            for (int qt = 0; qt < QT; qt++)
                if (qt < nqt) acc[qt] = Mfma<double>::mma(av[u][qt], b_, acc[qt]);
        }
    }
    const int j = j0 + jl;
#pragma unroll
    for (int qt = 0; qt < QT; qt++)
        if (qt < nqt) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int q = 16 * qt + Mfma<double>::row(lane, v);
                if (q < p && j < n) V[(long long)q * n + j] = acc[qt][v];
            }
        }
}
