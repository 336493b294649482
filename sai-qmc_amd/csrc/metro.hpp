// Single-site Metropolis over one imaginary-time slice with delayed
// updates, one workgroup per replica (forward PH sweep, src/dqmc.c:656-679,
// src/green.c:384-475).
//
// G itself is not touched during the slice.  Accepted flips become pairs
// c = (N/R)(e_i - G~ e_i), v = G~^T e_i (green_delay_accept,
// src/green.c:418-460) appended to per-replica pending buffers C, V (n x P,
// q-major: C[q*n + j]); G~ = G - C V^T is never formed.  The sites are
// visited in windows of W:
//   1. the window's columns and rows of G~ are computed on MFMA from G and
//      the pending pairs (n x W x P products, staged through LDS);
//   2. wave 0 runs the window's W proposals alone on the W x W block of G~
//      in its registers: each accept forms the pair restricted to the window
//      and applies the rank-1 change to the block (no workgroup barrier per
//      proposal);
//   3. every thread j forms c_q(j), v_q(j) of the window's accepts from its
//      column/row registers and the window-restricted pairs (O(p^2) per
//      thread), and appends them to C, V.
// The rank-P flush G -= C V^T (green_delay_flush / la_gemm_nt_rect,
// src/green.c:462-475, src/linalg.c:209-219) is one batched GEMM after the
// kernel, with P per replica in `pcount`.  A launch covers the sites
// [s0, s1) of the slice: splitting a slice into segments, each followed by
// its flush, bounds P and with it the correction work of later windows.
//
// The ratio is R = (1 + (1 - G~_ii) N_up)(1 + G~_ii N_dn)
// (src/green.c:401-416), always evaluated in FP64, accepted iff
// rng_double() < |R| (src/dqmc.c:668).  Every proposal consumes one draw of
// the replica's xoshiro256** stream, in site order: a window's draws are
// taken while the window before it is decided.
#pragma once
#include "common.hpp"

#define METRO_THREADS 512

template <typename T> struct MetroArgs {
    int n, L, W, l, batch;
    int s0, s1;         // sites [s0, s1) of the slice in this launch
    const T *G;
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
    T *Cp, *Vp;         // pending pairs, n x n per replica (q-major)
    long long sP;
    int *pcount;        // pending pair count per replica (the flush GEMM's K)
    double *rlog;       // null, or the ratio of every proposal, n per replica
};

#define METRO_MAX_N 512

// LDS: corrected columns and rows (2 x W x (n+1)), the window slices of C
// and V (2 x n x W, P <= n), the window block and the window-restricted
// pairs (3 x W x W), N/R per site, the slice's draws, spins and flips.
template <typename T>
static inline size_t metro_lds_bytes(int n, int W)
{
    return sizeof(T) * (2 * (size_t)W * (n + 1) + 2 * (size_t)n * W + 3 * (size_t)W * W + W) + sizeof(double) * n +
           2 * (size_t)n + 64;
}

__device__ __forceinline__ float bcast_any(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double bcast_any(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Window corrections on MFMA, columns and rows in one pass:
//   Sc[w][j] = G[j][i0+w] - sum_q C[q][j] V[q][i0+w]
//   Sr[w][j] = G[i0+w][j] - sum_q C[q][i0+w] V[q][j]
// Sc, Sr are W x (n+1) in LDS; Vw[q*W + w] = V[q][i0+w] and Cw[q*W + w] =
// C[q][i0+w] are the window slices staged in LDS.  Each wave takes JT
// 16-row blocks of j for both outputs at once, streaming C and V from
// global memory UB k-steps (4*JT*UB loads per lane) at a time; the kernel
// runs JT = 2, UB = 4 on 8 waves (UB = 8 was 0.6 % slower per sweep once
// the k-steps past P were skipped: the loads past P are still issued).  Each accumulator row of a tile covers 16
// consecutive rows of one column of G, so the G loads are 128-byte segments.
template <typename T, int W, int JT, int UB>
__device__ void window_correct(const T *G, int n, int i0, int wn, const T *C, const T *V, int P, const T *Cw,
                               const T *Vw, T *Sc, T *Sr)
{
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int NTW = (W + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const int ntj = (n + 15) >> 4;
    const int ld = n + 1;
    for (int tb = wave; tb < ntj; tb += JT * nw) {
        int j0[JT];
        bool jok[JT];
#pragma unroll
        for (int t = 0; t < JT; t++) {
            j0[t] = (tb + t * nw) * 16;
            jok[t] = j0[t] + lc < n;
        }
        acc_t ac[JT][NTW], ar[JT][NTW];
#pragma unroll
        for (int t = 0; t < JT; t++)
#pragma unroll
            for (int tw = 0; tw < NTW; tw++) {
                const int w0 = tw * 16;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rr = Mfma<T>::row(lane, v);
                    // cols tile: rows w = w0 + rr, columns j = j0 + lc;
                    // rows tile: rows j = j0 + rr, columns w = w0 + lc
                    // (clamped addresses, zeroed below)
                    const int wc = i0 + min(w0 + rr, wn - 1), jc = min(j0[t] + lc, n - 1);
                    const int wr = i0 + min(w0 + lc, wn - 1), jr = min(j0[t] + rr, n - 1);
                    ac[t][tw][v] = G[jc + (long long)wc * n];
                    ar[t][tw][v] = G[wr + (long long)jr * n];
                }
            }
#pragma unroll
        for (int t = 0; t < JT; t++)
#pragma unroll
            for (int tw = 0; tw < NTW; tw++) {
                const int w0 = tw * 16;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rr = Mfma<T>::row(lane, v);
                    ac[t][tw][v] = (w0 + rr < wn && jok[t]) ? ac[t][tw][v] : T(0);
                    ar[t][tw][v] = (w0 + lc < wn && j0[t] + rr < n) ? ar[t][tw][v] : T(0);
                }
            }
        // big operands C[q][j0+lc] (cols) and V[q][j0+lc] (rows): UB k-steps
        // of the JT j blocks in flight per batch; another wave on the SIMD
        // computes while this one waits
        for (int q0 = 0; q0 < P; q0 += 4 * UB) {
            T cc[JT][UB], cv[JT][UB];
            // loads from clamped (always valid) addresses, zeroed afterwards:
            // no divergent branch between them
#pragma unroll
            for (int u = 0; u < UB; u++) {
                const int q = min(q0 + 4 * u + lr, P - 1);
#pragma unroll
                for (int t = 0; t < JT; t++) {
                    const long long o = (long long)q * n + min(j0[t] + lc, n - 1);
                    cc[t][u] = C[o];
                    cv[t][u] = V[o];
                }
            }
#pragma unroll
            for (int u = 0; u < UB; u++) {
                const bool qok = q0 + 4 * u + lr < P;
#pragma unroll
                for (int t = 0; t < JT; t++) {
                    cc[t][u] = (qok && jok[t]) ? cc[t][u] : T(0);
                    cv[t][u] = (qok && jok[t]) ? cv[t][u] : T(0);
                }
            }
            // k-steps wholly past P (the batch is UB steps whatever P is)
            // would only add products of zeros: skipped, wave-uniformly
#pragma unroll
            for (int u = 0; u < UB; u++) {
                if (q0 + 4 * u >= P) break;
                const int q = q0 + 4 * u + lr;
#pragma unroll
                for (int tw = 0; tw < NTW; tw++) {
                    const int w = tw * 16 + lc;
                    const int o = min(q, P - 1) * W + min(w, W - 1);
                    const bool ok = q < P && w < wn;
                    const T vw = ok ? Vw[o] : T(0);
                    const T cw = ok ? Cw[o] : T(0);
#pragma unroll
                    for (int t = 0; t < JT; t++) {
                        // cols: A[w][q] = -V[q][i0+w], B[q][j] = C[q][j]
                        ac[t][tw] = Mfma<T>::mma(-vw, cc[t][u], ac[t][tw]);
                        // rows: A[j][q] = -V[q][j], B[q][w] = C[q][i0+w]
                        ar[t][tw] = Mfma<T>::mma(-cv[t][u], cw, ar[t][tw]);
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < JT; t++)
#pragma unroll
            for (int tw = 0; tw < NTW; tw++) {
                const int w0 = tw * 16;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rr = Mfma<T>::row(lane, v);
                    if (w0 + rr < W && jok[t]) Sc[(w0 + rr) * ld + j0[t] + lc] = ac[t][tw][v];
                    if (w0 + lc < W && j0[t] + rr < n) Sr[(w0 + lc) * ld + j0[t] + rr] = ar[t][tw][v];
                }
            }
    }
}

template <typename T, int W>
__global__ __launch_bounds__(METRO_THREADS) void k_metro_slice(MetroArgs<T> a)
{
    static_assert(W == 8 || W == 16 || W == 32, "window of 8, 16 or 32 sites");
    constexpr int TH = METRO_THREADS;
    constexpr int RPL = W * W / 64; // rows of the window block per lane of wave 0
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status[b]) {
        if (threadIdx.x == 0) a.pcount[b] = 0;
        return;
    }
    const int n = a.n, j = threadIdx.x;
    const int lane = j & 63, wave = j >> 6;
    const bool act = j < n;
    const T *G = a.G + (long long)b * a.sG;
    T *Cp = a.Cp + (long long)b * a.sP;
    T *Vp = a.Vp + (long long)b * a.sP;
    signed char *fld = a.field + (long long)b * a.sF + (long long)a.l * n;

    extern __shared__ __align__(16) unsigned char lds_raw[];
    T *Sc = (T *)lds_raw;             // W x (n+1) corrected window columns
    T *Sr = Sc + (size_t)W * (n + 1); // W x (n+1) corrected window rows
    T *Cw = Sr + (size_t)W * (n + 1); // n x W window slice of C (q-major)
    T *Vw = Cw + (size_t)n * W;       // n x W window slice of V
    T *Bw = Vw + (size_t)n * W;       // W x W window block, Bw[r*W + c]
    T *cwin = Bw + W * W;            // cwin[u*W + r]: pair of the accept at window site u
    T *vwin = cwin + W * W;          // vwin[u*W + c]
    T *facw = vwin + W * W;          // N/R of the accept at window site u
    double *draws = (double *)(facw + W);
    signed char *fs = (signed char *)(draws + n);
    signed char *flip = fs + n;
    __shared__ unsigned s_mask;

    if (act) {
        fs[j] = fld[j];
        flip[j] = 0;
    }
    // The draws, one per proposal in site order: the first window's here,
    // each later window's by thread 64 (wave 1) while wave 0 decides the
    // window before it.
    const bool rng_thread = j == 64;
    uint64_t st[4] = {0, 0, 0, 0};
    uint64_t *rs = a.rng + 4 * (long long)b;
    if (rng_thread) {
        for (int q = 0; q < 4; q++) st[q] = rs[q];
        for (int i = a.s0; i < min(a.s1, a.s0 + W); i++) draws[i] = xoshiro_double(st);
    }
    __syncthreads();

    int naccept = 0, negflips = 0, P = 0;
    T gc[W];
    for (int i0 = a.s0; i0 < a.s1; i0 += W) {
        const int wn = min(W, a.s1 - i0);
        // 1. corrected window columns -> Sc, rows -> Sr, the window block -> Bw.
        // The window slices of C and V are staged four entries per thread
        // at a time, all loads issued before the first LDS store.
        for (int idx0 = threadIdx.x; idx0 < P * W; idx0 += 4 * TH) {
            T xc[4], xv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int idx = min(idx0 + k * TH, P * W - 1), q = idx / W, w = min(idx % W, wn - 1);
                xc[k] = Cp[(long long)q * n + i0 + w];
                xv[k] = Vp[(long long)q * n + i0 + w];
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int idx = idx0 + k * TH;
                if (idx < P * W) {
                    const bool ok = idx % W < wn;
                    Cw[idx] = ok ? xc[k] : T(0);
                    Vw[idx] = ok ? xv[k] : T(0);
                }
            }
        }
        __syncthreads();
        window_correct<T, W, 2, 4>(G, n, i0, wn, Cp, Vp, P, Cw, Vw, Sc, Sr);
        __syncthreads();
        for (int idx = threadIdx.x; idx < W * W; idx += blockDim.x) {
            const int r = idx / W, c = idx % W;
            Bw[idx] = (r < wn && c < wn) ? Sc[c * (n + 1) + i0 + r] : T(0);
        }
        __syncthreads();

        // 2. proposals on the window block, wave 0 alone
        if (wave == 0) {
            const int c = lane % W, h = lane / W;
            T bw[RPL];
#pragma unroll
            for (int k = 0; k < RPL; k++) {
                const int r = h * RPL + k;
                bw[k] = (r < wn && c < wn) ? Bw[r * W + c] : T(0);
            }
            // site data of the window in lanes 0..wn-1, broadcast per proposal
            double dr_l = 0.0, nu_l = 0.0, nd_l = 0.0;
            if (lane < wn) {
                const int sp = fs[i0 + lane] > 0;
                dr_l = draws[i0 + lane];
                nu_l = a.nu[sp];
                nd_l = a.nd[sp];
            }
            unsigned mask = 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
                if (w < wn) {
                    const int i = i0 + w;
                    const double gii = (double)bcast_any(bw[w % RPL], (w / RPL) * W + w);
                    const double Nu = bcast_any(nu_l, w), Nd = bcast_any(nd_l, w);
                    const double Ru = 1.0 + (1.0 - gii) * Nu;
                    const double Rd = 1.0 + gii * Nd;
                    const double R = Ru * Rd;
                    if (a.rlog && lane == 0) a.rlog[(long long)b * n + i] = R;
                    // N / R_up of an accept, started before the decision
                    const T fac = (T)(Nu / Ru);
                    if (bcast_any(dr_l, w) < fabs(R)) {
                        if constexpr (W == 16 && sizeof(T) == 8) {
                            // column w of the block to every lane of each
                            // 16-lane row (DPP row broadcast) gives c; row w,
                            // held by 16-lane row w / RPL, to all four rows
                            // (row swaps) gives v; G~ -= c v^T in registers.
                            // The pairs also go to LDS for the pair formation
                            // after the next barrier.
                            T cw[RPL];
#pragma unroll
                            for (int k = 0; k < RPL; k++)
                                cw[k] = fac * ((h * RPL + k == w ? T(1) : T(0)) - row_bcast(bw[k], w));
                            const T vc = row_to_all(bw[w % RPL], w / RPL);
                            if (c == w) {
#pragma unroll
                                for (int k = 0; k < RPL; k++) cwin[w * W + h * RPL + k] = cw[k];
                            }
                            if (h == w / RPL) vwin[w * W + c] = vc;
#pragma unroll
                            for (int k = 0; k < RPL; k++) bw[k] -= cw[k] * vc;
                        } else {
                            // column w of the block -> c, row w -> v (values before the update)
                            if (c == w) {
#pragma unroll
                                for (int k = 0; k < RPL; k++) {
                                    const int r = h * RPL + k;
                                    cwin[w * W + r] = fac * ((r == w ? T(1) : T(0)) - bw[k]);
                                }
                            }
                            if (h == w / RPL) vwin[w * W + c] = bw[w % RPL];
                            __builtin_amdgcn_s_waitcnt(0xc07f);
                            __builtin_amdgcn_wave_barrier();
                            const T vc = vwin[w * W + c];
#pragma unroll
                            for (int k = 0; k < RPL; k++) bw[k] -= cwin[w * W + h * RPL + k] * vc;
                            __builtin_amdgcn_s_waitcnt(0xc07f);
                            __builtin_amdgcn_wave_barrier();
                        }
                        if (lane == 0) {
                            facw[w] = fac;
                            flip[i] = 1;
                        }
                        mask |= 1u << w;
                        naccept++;
                        if (R < 0.0) negflips ^= 1;
                    }
                }
            }
            if (lane == 0) s_mask = mask;
        } else if (rng_thread) {
            for (int i = i0 + W; i < min(a.s1, i0 + 2 * W); i++) draws[i] = xoshiro_double(st);
        }
        __syncthreads();

        // 3. full pairs of the window's accepts, appended to C, V.  The c
        // and v halves are independent: with 2n threads or more, thread j <
        // n forms c_q(j) and thread n + j forms v_q(j); otherwise a thread
        // forms both, one after the other.  The mask is wave-uniform (scalar
        // branches).
        const unsigned mask = __builtin_amdgcn_readfirstlane(s_mask);
        const bool split = (int)blockDim.x >= 2 * n;
        const int jj = (split && j >= n) ? j - n : j;
        if (jj < n) {
            const int nparts = split ? 1 : 2;
            for (int part = 0; part < nparts; part++) {
                const bool vside = split ? (j >= n) : (part == 1);
                const T *Sx = vside ? Sr : Sc;
                const T *coef = vside ? cwin : vwin;
                T *dst = vside ? Vp : Cp;
#pragma unroll
                for (int w = 0; w < W; w++) {
                    const T x = Sx[min(w, wn - 1) * (n + 1) + jj];
                    gc[w] = (w < wn) ? x : T(0);
                }
                // right-looking: once an accept's entry is final it is
                // removed from every later site of the window at once (the
                // subtractions of each entry keep their order, and the
                // FMAs of different sites are independent)
#pragma unroll
                for (int u = 0; u < W; u++) {
                    if ((mask >> u) & 1u) {
                        gc[u] = vside ? gc[u] : facw[u] * ((jj == i0 + u ? T(1) : T(0)) - gc[u]);
                        const int q = P + __builtin_popcount(mask & ((1u << u) - 1u));
                        dst[(long long)q * n + jj] = gc[u];
#pragma unroll
                        for (int w = u + 1; w < W; w++) gc[w] -= gc[u] * coef[u * W + w];
                    }
                }
            }
        }
        P += __builtin_popcount(mask);
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
    if (rng_thread)
        for (int q = 0; q < 4; q++) rs[q] = st[q];
    if (j == 0) {
        a.pcount[b] = P;
        a.acc[b] += (unsigned long long)naccept;
        a.att[b] += (unsigned long long)(a.s1 - a.s0);
        if (negflips) a.sign[b] = -a.sign[b];
    }
}

// FP64 flush G -= C^T V of the pending pairs (green_delay_flush,
// src/green.c:462-475) for n a multiple of 64: C, V q-major (P x n per
// replica, P = pcount[b]).  One 64 x 64 block of G per workgroup (16 per
// replica at n = 256), 16 columns j per wave, all four 16-row i tiles:
// D^T[j][i] = sum_q V[q][j] C[q][i] on FP64 MFMA 16x16x4 (A = V^T, B = C,
// both read as 16 consecutive doubles per lane group, q in increasing order
// four at a time), then G = G_old - D.  The first sixteen q's are loaded
// first and the old G values of a lane only after the q loop: the loop's
// loads head the memory queue, and the wait for G_old meets the MFMAs of
// the other workgroups on the CU; no LDS, no barrier.  The q order and the
// final G_old - D match the rank-P GEMM this replaces.  Bounded to at
// least five waves per SIMD: the compiler then keeps it to 52 VGPRs, and
// more workgroups per CU hide the loads better than more loads in flight
// per workgroup (0.5 % per sweep at 16x16).
struct FlushArgs {
    int n, batch;
    double *G;
    long long sG;
    const double *C, *V;
    long long sP;
    const int *pcount;
    const int *status;
};

__global__ __launch_bounds__(256, 5) void k_flush64(FlushArgs a)
{
    const int n = a.n, nb = n >> 6, per = nb * nb;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / per, t = logical % per;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int P = a.pcount[b];
    if (P <= 0) return;
    const int i0 = (t % nb) * 64, j0 = (t / nb) * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    double *G = a.G + (long long)b * a.sG;
    const double *C = a.C + (long long)b * a.sP;
    const double *V = a.V + (long long)b * a.sP;
    const int jw = j0 + 16 * wave; // this wave's 16 columns
    double av[4], bv[4][4];
    // clamped (always valid) rows, zeroed past P after the loads
    auto load_q = [&](int q0) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long long qr = (long long)min(q0 + 4 * u + lr, P - 1) * n;
            av[u] = V[qr + jw + lc];
#pragma unroll
            for (int ti = 0; ti < 4; ti++) bv[u][ti] = C[qr + i0 + 16 * ti + lc];
        }
    };
    f64x4 acc[4];
#pragma unroll
    for (int ti = 0; ti < 4; ti++) acc[ti] = f64x4{0, 0, 0, 0};
    for (int q0 = 0; q0 < P; q0 += 16) {
        load_q(q0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (q0 + 4 * u >= P) break; // k-steps wholly past P: products of zeros
            const double a_ = q0 + 4 * u + lr < P ? av[u] : 0.0;
#pragma unroll
            for (int ti = 0; ti < 4; ti++) acc[ti] = Mfma<double>::mma(a_, bv[u][ti], acc[ti]);
        }
    }
    // accumulator entry v of lane (lr, lc) in i tile ti: D^T[jw + lr + 4v][i0 + 16 ti + lc]
    double old[4][4];
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) old[ti][v] = G[(i0 + 16 * ti + lc) + (long long)(jw + lr + 4 * v) * n];
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            G[(i0 + 16 * ti + lc) + (long long)(jw + lr + 4 * v) * n] = old[ti][v] - acc[ti][v];
}
