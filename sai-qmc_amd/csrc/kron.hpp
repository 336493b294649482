// B-matrix products on a 16 x 16 square lattice with separable hopping.
//
// With nearest-neighbour hopping along x and y, K = I_y (x) K_x + K_y (x) I_x
// and the two terms commute, so expK = exp(-dtau K) = E_y (x) E_x exactly
// (site i = x + 16 y): expK[(ix, iy), (jx, jy)] = E_y[iy][jy] E_x[ix][jx].
// Applied to a vector z of n = 256 entries, viewed as the 16 x 16 grid
// Z[ex][ey] = z[ex + 16 ey], (A_y (x) A_x) z is the grid A_x Z A_y^T: two
// 16 x 16 x 16 products (8 FP64 MFMAs) instead of a 256-long dot product
// per entry, 2 * 2 * 16 n flops per vector against 2 n^2.  The products
// expK diag(d) X of the wrap and of the block products
// (green_wrap_forward / green_wrap_backward / green_build_Bblock,
// src/green.c:126-166, 509-546) become one pass over the columns (left
// factor) or the rows (right factor, X M = (M^T X^T)^T) of X: HBM-bound,
// X read and written once.
//
// k_kron16: one workgroup per 16 vectors (columns, or rows with `rows`) of
// one replica.  The vectors are staged in LDS (pre-scaled), each wave
// transforms four of them in registers:
//   P^T = Z^T A_x^T   (A operand Z^T from LDS, B operand A_x^T),
//   out^T = A_y P^T   (A operand A_y, B operand P^T straight from the
//                      accumulator: FP64 16x16x4 keeps accumulator entry v
//                      of lane (lr, lc) at row lr + 4v, which is the B
//                      operand of k-step v with k = 4v + lr),
// and the results go back through LDS to coalesced stores (post-scaled).
// A column pass can first apply a pending rank-P delayed update to its 16
// columns (the last Metropolis segment's flush, green_delay_flush,
// src/green.c:462-475, when the wrap follows it directly), so that G is
// read and written once for both.
// Scales: `pre` / `post` along the transformed index before / after the
// transform, `oth` along the other index (it commutes with the transform).
#pragma once
#include "common.hpp"

#define KRON_L 16
#define KRON_N (KRON_L * KRON_L)
#define KRON_LD 273 // LDS doubles per vector: 16 grid rows of 17 (one pad), and one more so
                    // that the 16 vectors of a row-pass lane group fall on 16 banks

struct KronArgs {
    int batch, rows;     // rows = 0: transform the columns of X; 1: its rows (the launcher
                         // picks the kernel mode from it)
    const double *in;    // n x n per replica, column-major
    long long sIn;
    double *out;         // may not alias `in`
    long long sOut;
    const double *ax, *ay; // 16 x 16 factors, row-major: grid -> ax grid ay^T
    Vec<double> pre, post, oth;
    const int *status;
    // optional pending delayed update applied to the columns before the
    // transform (rows = 0 only): in <- in - C^T V with C, V q-major
    // (P x n per replica, stride sP), P = pcount[b] (fc == null: none)
    const double *fc, *fv;
    long long sP;
    const int *pcount;
};

__device__ __forceinline__ int kron_lds(int e) { return (e & 15) + 17 * (e >> 4); }

// MODE: KRON_COLS, KRON_ROWS, or KRON_COLS_PENDING (a column pass that
// applies the pending update first); one instantiation per mode, so the
// plain passes carry none of the pending update's registers.  At least four
// waves per SIMD (at most 128 VGPRs): with the LDS of four workgroups per
// CU, the pending pass then holds four workgroups instead of three.
#define KRON_COLS 0
#define KRON_ROWS 1
#define KRON_COLS_PENDING 2
template <int MODE>
__global__ __launch_bounds__(256, 4) void k_kron16(KronArgs a)
{
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / KRON_L, v0 = (logical % KRON_L) * KRON_L;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    __shared__ double Z[KRON_L * KRON_LD];
    const double *in = a.in + (long long)b * a.sIn;
    double *out = a.out + (long long)b * a.sOut;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lr = lane >> 4;

    // 1. stage: vector t = column (row) v0 + t, entry e along it
    double x[KRON_L];
    if (MODE != KRON_ROWS) {
        // thread = entry e = tid of 16 columns: 64 consecutive doubles per
        // wave.  The pending rank-P update of these 16 columns, D[e][t] =
        // sum_q C[q][e] V[q][v0 + t], is the longer dependence chain (its
        // loads, the MFMAs, the exchange through Z), so its first sixteen
        // q's are loaded ahead of the columns themselves; wave w takes rows
        // 64w .. 64w + 63 (four 16-row tiles), q four at a time in
        // increasing order, and D goes to Z in the staging layout
        const int P = (MODE == KRON_COLS_PENDING && a.fc) ? a.pcount[b] : 0;
        const double *C = a.fc + (long long)b * a.sP, *V = a.fv + (long long)b * a.sP;
        const int e0 = 64 * wave;
        double av[4][4], bv[4];
        // clamped (always valid) rows, zeroed past P after the loads
        auto load_q = [&](int q0) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const long long qr = (long long)min(q0 + 4 * u + lr, P - 1) * KRON_N;
                bv[u] = V[qr + v0 + lc];
#pragma unroll
                for (int ti = 0; ti < 4; ti++) av[u][ti] = C[qr + e0 + 16 * ti + lc];
            }
        };
        if (P > 0) load_q(0);
#pragma unroll
        for (int t = 0; t < KRON_L; t++) x[t] = in[tid + (long long)(v0 + t) * KRON_N];
        const double pe_raw = a.pre.raw(b, tid);
        if (P > 0) {
            f64x4 acc[4];
#pragma unroll
            for (int ti = 0; ti < 4; ti++) acc[ti] = f64x4{0, 0, 0, 0};
            for (int q0 = 0; q0 < P; q0 += 16) {
                if (q0 > 0) load_q(q0);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const double b_ = q0 + 4 * u + lr < P ? bv[u] : 0.0;
#pragma unroll
                    for (int ti = 0; ti < 4; ti++) acc[ti] = Mfma<double>::mma(av[u][ti], b_, acc[ti]);
                }
            }
#pragma unroll
            for (int ti = 0; ti < 4; ti++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    Z[lc * KRON_LD + kron_lds(e0 + 16 * ti + Mfma<double>::row(lane, v))] = acc[ti][v];
            __syncthreads();
#pragma unroll
            for (int t = 0; t < KRON_L; t++) x[t] -= Z[t * KRON_LD + kron_lds(tid)];
        }
        const double pe = a.pre.inv ? 1.0 / pe_raw : pe_raw;
#pragma unroll
        for (int t = 0; t < KRON_L; t++) Z[t * KRON_LD + kron_lds(tid)] = x[t] * pe * a.oth.at(b, v0 + t);
    } else {
        // thread = (row t, entries jq + 16k): 16 consecutive doubles per 16 lanes
        const int t = tid & 15, jq = tid >> 4;
#pragma unroll
        for (int k = 0; k < KRON_L; k++) x[k] = in[(v0 + t) + (long long)(jq + 16 * k) * KRON_N];
        const double oi = a.oth.at(b, v0 + t);
#pragma unroll
        for (int k = 0; k < KRON_L; k++) {
            const int e = jq + 16 * k;
            Z[t * KRON_LD + kron_lds(e)] = x[k] * a.pre.at(b, e) * oi;
        }
    }
    __syncthreads();

    // 2. transforms, four vectors per wave, in place in LDS (each wave reads
    // a vector completely before it writes it back)
    double bx[4], ay[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        bx[s] = a.ax[lc * KRON_L + 4 * s + lr]; // B[k = 4s + lr][n = lc] = ax[lc][4s + lr]
        ay[s] = a.ay[lc * KRON_L + 4 * s + lr]; // A[m = lc][k = 4s + lr]
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double *Zt = Z + (4 * wave + q) * KRON_LD;
        double za[4];
#pragma unroll
        for (int s = 0; s < 4; s++) za[s] = Zt[kron_lds((4 * s + lr) + 16 * lc)]; // Z[ex = 4s + lr][ey = lc]
        f64x4 p = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++) p = Mfma<double>::mma(za[s], bx[s], p);
        f64x4 o = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++) o = Mfma<double>::mma(ay[s], p[s], o);
        // o[v] = out^T[ey' = lr + 4v][ex' = lc]
#pragma unroll
        for (int v = 0; v < 4; v++) Zt[kron_lds(lc + 16 * (lr + 4 * v))] = o[v];
    }
    __syncthreads();

    // 3. store, post-scaled along the transformed index
    if (MODE != KRON_ROWS) {
        const double po = a.post.at(b, tid);
#pragma unroll
        for (int t = 0; t < KRON_L; t++) out[tid + (long long)(v0 + t) * KRON_N] = Z[t * KRON_LD + kron_lds(tid)] * po;
    } else {
        const int t = tid & 15, jq = tid >> 4;
#pragma unroll
        for (int k = 0; k < KRON_L; k++) {
            const int e = jq + 16 * k;
            out[(v0 + t) + (long long)e * KRON_N] = Z[t * KRON_LD + kron_lds(e)] * a.post.at(b, e);
        }
    }
}
