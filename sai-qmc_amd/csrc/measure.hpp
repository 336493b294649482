// Equal-time measurement at tau = 0, one workgroup per replica, FP64.
//
// G_dn is never formed: the particle-hole map G_dn(i,j) = delta_ij -
// P_i P_j G_up(j,i) (src/green.c:372-382) is applied on the fly.
// Scalars follow measure_sample (src/measure.c:6-39); the spin
// correlations follow spin_pair_correlations (src/structure_factor.c:
// 258-297), summed per lattice displacement as measure_szz_sperp_sample does
// (src/structure_factor.c:355-391) with one thread per displacement (no
// atomics), then cosine-transformed (src/structure_factor.c:311-326).  The
// sample is either written out raw or added with the replica's sign into a
// bin (src/replica.c:48-71, 240-269).
#pragma once
#include "common.hpp"

struct MeasureArgs {
    int n, Lx, Ly, nq, batch;
    const double *G;
    long long sG;
    const double *thop;
    const int *bipart;
    const double *phase; // phase[disp + q*n]
    double U, mu;
    // raw output (bin < 0)
    double *scal5;  // [batch][5]
    double *szz, *sperp; // [batch][nq]
    // binned output (bin >= 0)
    int bin, nbin;
    double *bins;   // [batch][nbin][8]
    int *counts;    // [batch][nbin]
    double *bszz, *bsperp; // [batch][nbin][nq]
    const double *sign;
    const unsigned long long *acc, *att;
    unsigned long long *acc_prev, *att_prev;
    int *status, *reason;
};

__global__ __launch_bounds__(256) void k_measure(MeasureArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int n = a.n, tid = threadIdx.x, nt = blockDim.x;
    const double *G = a.G + (long long)b * a.sG;
    extern __shared__ double lm[];
    double *cz = lm;      // [n]
    double *cp = lm + n;  // [n]
    double *gdiag = lm + 2 * n;
    double *red = lm + 3 * n; // 8 slots
    __shared__ int s_bad;
    if (tid == 0) s_bad = 0;
    for (int i = tid; i < n; i += nt) gdiag[i] = G[i + (long long)i * n];
    __syncthreads();

    // scalars: ekin = sum_ij t_ij [(delta - Gu_ji) + (delta - Gd_ji)],
    // (delta - Gd_ji) = P_i P_j Gu_ij (i != j), Gu_ii (i == j)
    double ek = 0.0, ei = 0.0, nt_ = 0.0;
    for (long long idx = tid; idx < (long long)n * n; idx += nt) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        const double t = a.thop[idx];
        if (t != 0.0) {
            const double gu_ji = G[j + (long long)i * n];
            const double d = (i == j) ? 1.0 : 0.0;
            const double pd = (double)(a.bipart[i] * a.bipart[j]);
            const double gd_ji = d - pd * G[i + (long long)j * n];
            ek += t * (d - gu_ji) + t * (d - gd_ji);
        }
    }
    for (int i = tid; i < n; i += nt) {
        const double nu = 1.0 - gdiag[i], nd = gdiag[i]; // 1 - Gd_ii = Gu_ii
        ei += nu * nd;
        nt_ += nu + nd;
    }
    ek = block_sum(ek, red);
    ei = block_sum(ei, red);
    nt_ = block_sum(nt_, red);

    // displacement sums
    const int Lx = a.Lx, Ly = a.Ly;
    for (int disp = tid; disp < n && a.nq > 0; disp += nt) {
        const int dx = disp % Lx, dy = disp / Lx;
        double sz = 0.0, sp = 0.0;
        for (int i = 0; i < n; i++) {
            const int xi = i % Lx, yi = i / Lx;
            int xj = xi - dx, yj = yi - dy;
            if (xj < 0) xj += Lx;
            if (yj < 0) yj += Ly;
            const int j = xj + yj * Lx;
            const double d = (i == j) ? 1.0 : 0.0;
            const double pij = (double)(a.bipart[i] * a.bipart[j]);
            const double gu_ii = gdiag[i], gu_jj = gdiag[j];
            const double gd_ii = 1.0 - gu_ii, gd_jj = 1.0 - gu_jj;
            const double gu_ij = G[i + (long long)j * n];
            const double gu_ji = G[j + (long long)i * n];
            const double gd_ij = d - pij * gu_ji;
            const double gd_ji = d - pij * gu_ij;
            const double si = gd_ii - gu_ii, sj = gd_jj - gu_jj;
            sz += 0.25 * (si * sj + (d - gu_ji) * gu_ij + (d - gd_ji) * gd_ij);
            sp += 0.5 * ((d - gu_ji) * gd_ij + (d - gd_ji) * gu_ij);
        }
        cz[disp] = sz;
        cp[disp] = sp;
        if (!isfinite(sz) || !isfinite(sp)) s_bad = 1;
    }
    __syncthreads();

    const double E = ek + a.U * ei;
    const double doublon = ei / (double)n;
    if (!isfinite(E) || !isfinite(nt_) || !isfinite(doublon)) s_bad = 1;
    __syncthreads();
    if (s_bad) {
        if (tid == 0 && a.status) {
            a.status[b] = 1;
            if (a.reason) a.reason[b] = 16;
        }
        return;
    }
    if (a.bin < 0) {
        if (tid == 0) {
            double *o = a.scal5 + 5 * (long long)b;
            o[0] = E; o[1] = ek; o[2] = a.U * ei; o[3] = nt_; o[4] = doublon;
        }
        for (int q = tid; q < a.nq; q += nt) {
            double z = 0.0, p = 0.0;
            for (int d = 0; d < n; d++) {
                const double c = a.phase[d + (long long)q * n];
                z += c * cz[d];
                p += c * cp[d];
            }
            a.szz[(long long)b * a.nq + q] = z / (double)n;
            a.sperp[(long long)b * a.nq + q] = p / (double)n;
        }
        return;
    }
    const double sg = a.sign[b];
    if (tid == 0) {
        double *o = a.bins + ((long long)b * a.nbin + a.bin) * 8;
        const double ehub = E;
        const double egc = E - a.mu * nt_;
        const double eph = E - 0.5 * a.U * nt_ + 0.25 * a.U * (double)n;
        o[0] += sg * ehub;
        o[1] += sg * egc;
        o[2] += sg * eph;
        o[3] += sg * nt_;
        o[4] += sg * doublon;
        o[5] += sg;
        o[6] += (double)(a.acc[b] - a.acc_prev[b]);
        o[7] += (double)(a.att[b] - a.att_prev[b]);
        a.acc_prev[b] = a.acc[b];
        a.att_prev[b] = a.att[b];
        a.counts[(long long)b * a.nbin + a.bin] += 1;
    }
    for (int q = tid; q < a.nq; q += nt) {
        double z = 0.0, p = 0.0;
        for (int d = 0; d < n; d++) {
            const double c = a.phase[d + (long long)q * n];
            z += c * cz[d];
            p += c * cp[d];
        }
        const long long o = ((long long)b * a.nbin + a.bin) * a.nq + q;
        a.bszz[o] += sg * (z / (double)n);
        a.bsperp[o] += sg * (p / (double)n);
    }
}

static inline size_t measure_lds_bytes(int n) { return sizeof(double) * (3 * (size_t)n + 8); }
