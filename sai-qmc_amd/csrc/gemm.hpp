// Batched GEMM on MFMA with fused diagonal scalings.
//
//   C_b = alpha * diag(cr) [ (diag(ar) op(A_b) diag(ac)) (diag(br) op(B_b) diag(bc)) ] diag(cc)
//         + diag(dg) + beta * C_b
//
// Every B-matrix product of the sweep is one of these: B_l = expK diag(d_l)
// and B_l^-1 = diag(1/d_l) expKinv (src/green.c:126-149) become operand
// scalings, so no B matrix is ever materialised; the UDV diagonal factors
// (src/linalg.c:556-561, 700-705, 840-846) and the Db/Ds split
// (src/linalg.c:1009-1022) are scalings and a diagonal add.
//
// Tile 64x64x16, 256 threads = 4 waves, each wave a 32x32 block of 2x2
// MFMA 16x16x4 tiles; operands staged through LDS k-major so fragment reads
// are conflict-free.  The grid is one-dimensional: tiles of a batch entry
// are consecutive logical ids and xcd_remap keeps them on one XCD.
#pragma once
#include "common.hpp"
#include <cstdlib>

template <typename T> struct GemmArgs {
    int M, N, K, batch;
    const T *A;
    long long lda, sA;
    int transA;
    const T *B;
    long long ldb, sB;
    int transB;
    T *C;
    long long ldc, sC;
    T alpha, beta;
    Vec<T> ar, ac, br, bc, cr, cc, dg;
    const int *status; // skip batch entries with status != 0 (may be null)
    const int *Kb;     // per-entry K <= K (may be null)
    int vec2;          // set by gemm_launch: element pairs may be loaded as one aligned vector
};

#define GEMM_BM 64
#define GEMM_BN 64
#define GEMM_BK 16

template <typename T>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs<T> g)
{
    const int tiles_m = (g.M + GEMM_BM - 1) / GEMM_BM;
    const int tiles_n = (g.N + GEMM_BN - 1) / GEMM_BN;
    const int per = tiles_m * tiles_n;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / per;
    const int t = logical % per;
    const int m0 = (t % tiles_m) * GEMM_BM;
    const int n0 = (t / tiles_m) * GEMM_BN;
    if (b >= g.batch) return;
    if (g.status && g.status[b]) return;
    const int K = g.Kb ? min(g.K, g.Kb[b]) : g.K;
    if (K <= 0 && g.beta == T(1) && !g.dg.p) return; // C unchanged

    __shared__ T As[GEMM_BK][GEMM_BM + 4];
    __shared__ T Bs[GEMM_BK][GEMM_BN + 4];

    const T *A = g.A + (long long)b * g.sA;
    const T *B = g.B + (long long)b * g.sB;
    T *C = g.C + (long long)b * g.sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = acc_t{0, 0, 0, 0};

    for (int k0 = 0; k0 < K; k0 += GEMM_BK) {
#pragma unroll
        for (int r = 0; r < (GEMM_BM * GEMM_BK) / 256; r++) {
            const int e = tid + 256 * r;
            int i, k;
            if (!g.transA) { i = e % GEMM_BM; k = e / GEMM_BM; }
            else { k = e % GEMM_BK; i = e / GEMM_BK; }
            const int gi = m0 + i, gk = k0 + k;
            T v = 0;
            if (gi < g.M && gk < K) {
                v = g.transA ? A[gk + (long long)gi * g.lda] : A[gi + (long long)gk * g.lda];
                v *= g.ar.at(b, gi) * g.ac.at(b, gk);
            }
            As[k][i] = v;
        }
#pragma unroll
        for (int r = 0; r < (GEMM_BN * GEMM_BK) / 256; r++) {
            const int e = tid + 256 * r;
            int k, j;
            if (!g.transB) { k = e % GEMM_BK; j = e / GEMM_BK; }
            else { j = e % GEMM_BN; k = e / GEMM_BN; }
            const int gk = k0 + k, gj = n0 + j;
            T v = 0;
            if (gk < K && gj < g.N) {
                v = g.transB ? B[gj + (long long)gk * g.ldb] : B[gk + (long long)gj * g.ldb];
                v *= g.br.at(b, gk) * g.bc.at(b, gj);
            }
            Bs[k][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GEMM_BK; kk += 4) {
            T a[2], bb[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++) a[mi] = As[kk + (lane >> 4)][wr * 32 + mi * 16 + (lane & 15)];
#pragma unroll
            for (int nj = 0; nj < 2; nj++) bb[nj] = Bs[kk + (lane >> 4)][wc * 32 + nj * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int nj = 0; nj < 2; nj++) acc[mi][nj] = Mfma<T>::mma(a[mi], bb[nj], acc[mi][nj]);
        }
        __syncthreads();
    }

    // epilogue: scales and the old C values are loaded as one batch from
    // clamped addresses; the stores are predicated
    T sr[2][4], sc[2], old[2][2][4];
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int v = 0; v < 4; v++) sr[mi][v] = g.cr.raw(b, min(m0 + wr * 32 + mi * 16 + Mfma<T>::row(lane, v), g.M - 1));
#pragma unroll
    for (int nj = 0; nj < 2; nj++) sc[nj] = g.cc.raw(b, min(n0 + wc * 32 + nj * 16 + (lane & 15), g.N - 1));
    if (g.beta != T(0)) {
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
#pragma unroll
            for (int nj = 0; nj < 2; nj++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = min(m0 + wr * 32 + mi * 16 + Mfma<T>::row(lane, v), g.M - 1);
                    const int j = min(n0 + wc * 32 + nj * 16 + (lane & 15), g.N - 1);
                    old[mi][nj][v] = C[i + (long long)j * g.ldc];
                }
    }
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = m0 + wr * 32 + mi * 16 + Mfma<T>::row(lane, v);
                const int j = n0 + wc * 32 + nj * 16 + (lane & 15);
                T r = g.alpha * acc[mi][nj][v] * g.cr.fix(sr[mi][v]) * g.cc.fix(sc[nj]);
                if (g.dg.p && i == j && i < g.M) r += g.dg.at(b, i);
                if (g.beta != T(0)) r += g.beta * old[mi][nj][v];
                if (i < g.M && j < g.N) C[i + (long long)j * g.ldc] = r;
            }
}

// Large-tile variant for M, N >= 128: 128x128x16 workgroup tile, each of
// the 4 waves a 64x64 block of 4x4 MFMA tiles.  Tiles are computed as C^T
// (the MFMA A operand is op(B), the B operand op(A)), so the 16 lanes of an
// accumulator row hold 16 consecutive rows of one column of C and the
// epilogue stores 128-byte segments.  The next k-tile is loaded into
// registers (two elements per load) while the current one is multiplied
// out of the other half of a double-buffered LDS; rows of the k-major LDS
// tiles are padded so a 4-row fragment read covers all 64 banks once.
#define GB_M 128
#ifndef GB_K
#define GB_K 16
#endif
#define GB_LD (GB_M + 16)
#define GB_PPT (GB_K / 4)       // element pairs per thread per operand and k-tile
#define GB_KH (GB_K / 2)        // pairs along k
#define GB_IS (256 / GB_KH)     // rows between a thread's pairs when pairs run along k
#define GB_LDK (GB_K + 4)       // j-major B tile (B not transposed): Bs[j*GB_LDK + k]
#define GB_BUF (GB_K * GB_LD + GB_M * GB_LDK) // one LDS buffer: A tile + B tile

template <typename T>
static inline size_t gemm_big_lds_bytes()
{
    return sizeof(T) * 2 * GB_BUF;
}

template <typename T> struct Two { T x, y; };

// Loads elements off and off + 1 of base (16-byte aligned pair) as one
// vector without a divergent branch: a pair outside the matrix is read from
// base[0] instead (always valid) and zeroed later.
template <typename T>
__device__ __forceinline__ Two<T> ld_pair(const T *base, long long off, bool ok)
{
    Two<T> r;
    const T *p = base + (ok ? off : 0);
    if constexpr (sizeof(T) == 8) {
        const double2 v = *reinterpret_cast<const double2 *>(p);
        r.x = v.x; r.y = v.y;
    } else {
        const float2 v = *reinterpret_cast<const float2 *>(p);
        r.x = v.x; r.y = v.y;
    }
    return r;
}

template <typename T, int TA, int TB>
__global__ __launch_bounds__(256, 2) void k_gemm_big(GemmArgs<T> g)
{
    const int tiles_m = (g.M + GB_M - 1) / GB_M;
    const int tiles_n = (g.N + GB_M - 1) / GB_M;
    const int per = tiles_m * tiles_n;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / per;
    const int t = logical % per;
    const int m0 = (t % tiles_m) * GB_M;
    const int n0 = (t / tiles_m) * GB_M;
    if (b >= g.batch) return;
    if (g.status && g.status[b]) return;
    const int K = g.Kb ? min(g.K, g.Kb[b]) : g.K;
    if (K <= 0 && g.beta == T(1) && !g.dg.p) return; // C unchanged

    extern __shared__ __align__(16) unsigned char gb_raw[];
    T *lds = reinterpret_cast<T *>(gb_raw);
    // buffer u: As = lds + u * GB_BUF (k-major), Bs = As + GB_K * GB_LD (k-major
    // when B is transposed; j-major otherwise, so that a thread's two
    // consecutive k of one column are one 16-byte store and the fragment
    // reads of 16 columns x 4 k spread over all banks)
    const T *A = g.A + (long long)b * g.sA;
    const T *B = g.B + (long long)b * g.sB;
    T *C = g.C + (long long)b * g.sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;

    // fixed per-thread load coordinates (GB_PPT pairs per operand per k-tile)
    //   A, not transposed: pair (2*i2, 2*i2+1) x k, i2 = tid % 64, k = tid / 64 + 4 r
    //   A, transposed:     pair i x (2*k2, 2*k2+1), k2 = tid % GB_KH, i = tid / GB_KH + GB_IS r
    //   B, not transposed: pair (2*k2, 2*k2+1) x j, k2 = tid % GB_KH, j = tid / GB_KH + GB_IS r
    //   B, transposed:     pair k x (2*j2, 2*j2+1), j2 = tid % 64, k = tid / 64 + 4 r
    // A k-tile is read in two phases: load_tile only issues the loads (from
    // clamped, always valid addresses, no divergent branches), so all of
    // them are in flight during the multiplication of the previous tile;
    // store_tile then zeroes what lies outside the matrix, applies the
    // diagonal scalings and writes LDS.  Pairs are aligned vectors: the
    // launcher sends a GEMM here only when every pair is (vec2).
    const int ia = 2 * (tid & 63);             // TA = 0: row pair; TB = 1: column pair
    const int kq = tid >> 6;                    // TA = 0 / TB = 1: k row of pair r is kq + 4 r
    const int kh = 2 * (tid % GB_KH);           // TA = 1 / TB = 0: k pair
    const int ih = tid / GB_KH;                 // TA = 1 / TB = 0: row / column of pair r is ih + GB_IS r
    Two<T> ra[GB_PPT], rb[GB_PPT];
    T ska[GB_PPT], skb[GB_PPT], skp[4];         // raw k-dependent scales of the tile in flight
    // k-independent scales, fixed per thread
    T sa_fixed[2] = {T(1), T(1)}, sb_fixed[2] = {T(1), T(1)};
    T sa_row[GB_PPT], sb_col[GB_PPT];
    if (!TA) {
        sa_fixed[0] = (m0 + ia < g.M) ? g.ar.at(b, m0 + ia) : T(0);
        sa_fixed[1] = (m0 + ia + 1 < g.M) ? g.ar.at(b, m0 + ia + 1) : T(0);
    }
    if (TB) {
        sb_fixed[0] = (n0 + ia < g.N) ? g.bc.at(b, n0 + ia) : T(0);
        sb_fixed[1] = (n0 + ia + 1 < g.N) ? g.bc.at(b, n0 + ia + 1) : T(0);
    }
#pragma unroll
    for (int r = 0; r < GB_PPT; r++) {
        const int i = m0 + ih + GB_IS * r, j = n0 + ih + GB_IS * r;
        sa_row[r] = (TA && i < g.M) ? g.ar.at(b, i) : T(1);
        sb_col[r] = (!TB && j < g.N) ? g.bc.at(b, j) : T(1);
    }
    const int Kc = K > 0 ? K - 1 : 0; // clamp for scale reads

    auto load_tile = [&](int k0) {
#pragma unroll
        for (int r = 0; r < GB_PPT; r++) {
            if (!TA) {
                const int i = m0 + ia, k = k0 + kq + 4 * r;
                const bool ok0 = k < K && i < g.M, ok1 = k < K && i + 1 < g.M;
                ra[r] = ld_pair(A, i + (long long)k * g.lda, ok0 || ok1);
                ska[r] = g.ac.raw(b, min(k, Kc));
            } else {
                const int i = m0 + ih + GB_IS * r, k = k0 + kh;
                const bool ok0 = i < g.M && k < K, ok1 = i < g.M && k + 1 < K;
                ra[r] = ld_pair(A, k + (long long)i * g.lda, ok0 || ok1);
            }
            if (!TB) {
                const int j = n0 + ih + GB_IS * r, k = k0 + kh;
                const bool ok0 = j < g.N && k < K, ok1 = j < g.N && k + 1 < K;
                rb[r] = ld_pair(B, k + (long long)j * g.ldb, ok0 || ok1);
            } else {
                const int j = n0 + ia, k = k0 + kq + 4 * r;
                const bool ok0 = k < K && j < g.N, ok1 = k < K && j + 1 < g.N;
                rb[r] = ld_pair(B, j + (long long)k * g.ldb, ok0 || ok1);
                skb[r] = g.br.raw(b, min(k, Kc));
            }
        }
        if (TA) {
            skp[0] = g.ac.raw(b, min(k0 + kh, Kc));
            skp[1] = g.ac.raw(b, min(k0 + kh + 1, Kc));
        }
        if (!TB) {
            skp[2] = g.br.raw(b, min(k0 + kh, Kc));
            skp[3] = g.br.raw(b, min(k0 + kh + 1, Kc));
        }
    };
    auto store_tile = [&](int u, int k0) {
        T *As = lds + u * GB_BUF;
        T *Bs = As + GB_K * GB_LD;
#pragma unroll
        for (int r = 0; r < GB_PPT; r++) {
            if (!TA) {
                const int i = m0 + ia, k = k0 + kq + 4 * r;
                const bool kok = k < K;
                const T s = g.ac.fix(ska[r]);
                T *d = As + (kq + 4 * r) * GB_LD + ia;
                d[0] = (kok && i < g.M) ? ra[r].x * (sa_fixed[0] * s) : T(0);
                d[1] = (kok && i + 1 < g.M) ? ra[r].y * (sa_fixed[1] * s) : T(0);
            } else {
                const int i = m0 + ih + GB_IS * r, k = k0 + kh;
                const bool iok = i < g.M;
                As[kh * GB_LD + ih + GB_IS * r] = (iok && k < K) ? ra[r].x * (sa_row[r] * g.ac.fix(skp[0])) : T(0);
                As[(kh + 1) * GB_LD + ih + GB_IS * r] =
                    (iok && k + 1 < K) ? ra[r].y * (sa_row[r] * g.ac.fix(skp[1])) : T(0);
            }
            if (!TB) {
                const int j = n0 + ih + GB_IS * r, k = k0 + kh;
                const bool jok = j < g.N;
                T *d = Bs + (ih + GB_IS * r) * GB_LDK + kh;
                d[0] = (jok && k < K) ? rb[r].x * (g.br.fix(skp[2]) * sb_col[r]) : T(0);
                d[1] = (jok && k + 1 < K) ? rb[r].y * (g.br.fix(skp[3]) * sb_col[r]) : T(0);
            } else {
                const int j = n0 + ia, k = k0 + kq + 4 * r;
                const bool kok = k < K;
                const T s = g.br.fix(skb[r]);
                T *d = Bs + (kq + 4 * r) * GB_LD + ia;
                d[0] = (kok && j < g.N) ? rb[r].x * (s * sb_fixed[0]) : T(0);
                d[1] = (kok && j + 1 < g.N) ? rb[r].y * (s * sb_fixed[1]) : T(0);
            }
        }
    };

    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[4][4];
#pragma unroll
    for (int it = 0; it < 4; it++)
#pragma unroll
        for (int jt = 0; jt < 4; jt++) acc[it][jt] = acc_t{0, 0, 0, 0};

    load_tile(0);
    store_tile(0, 0);
    __syncthreads();
    int u = 0;
    for (int k0 = 0; k0 < K; k0 += GB_K) {
        const bool more = k0 + GB_K < K;
        if (more) load_tile(k0 + GB_K);
        const T *As = lds + u * GB_BUF;
        const T *Bs = As + GB_K * GB_LD;
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            T fa[4], fb[4];
#pragma unroll
            for (int jt = 0; jt < 4; jt++)
                fa[jt] = TB ? Bs[(kk + lr) * GB_LD + wc * 64 + jt * 16 + lc]
                            : Bs[(wc * 64 + jt * 16 + lc) * GB_LDK + kk + lr];
#pragma unroll
            for (int it = 0; it < 4; it++) fb[it] = As[(kk + lr) * GB_LD + wr * 64 + it * 16 + lc];
#pragma unroll
            for (int it = 0; it < 4; it++)
#pragma unroll
                for (int jt = 0; jt < 4; jt++) acc[it][jt] = Mfma<T>::mma(fa[jt], fb[it], acc[it][jt]);
        }
        if (more) store_tile(u ^ 1, k0 + GB_K);
        __syncthreads();
        u ^= 1;
    }

    // epilogue: the 4 row and 16 column scales of a lane are loaded as one
    // batch, the old C values (beta != 0) one 16-row block at a time, all
    // from clamped addresses; the stores are predicated
    T sr[4], sc[4][4];
#pragma unroll
    for (int it = 0; it < 4; it++) sr[it] = g.cr.raw(b, min(m0 + wr * 64 + it * 16 + lc, g.M - 1));
#pragma unroll
    for (int jt = 0; jt < 4; jt++)
#pragma unroll
        for (int v = 0; v < 4; v++) sc[jt][v] = g.cc.raw(b, min(n0 + wc * 64 + jt * 16 + Mfma<T>::row(lane, v), g.N - 1));
#pragma unroll
    for (int it = 0; it < 4; it++) sr[it] = g.cr.fix(sr[it]);
#pragma unroll
    for (int jt = 0; jt < 4; jt++)
#pragma unroll
        for (int v = 0; v < 4; v++) sc[jt][v] = g.cc.fix(sc[jt][v]);
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const int i = m0 + wr * 64 + it * 16 + lc;
        const int ic = min(i, g.M - 1);
        T old[4][4];
        if (g.beta != T(0)) {
#pragma unroll
            for (int jt = 0; jt < 4; jt++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int j = min(n0 + wc * 64 + jt * 16 + Mfma<T>::row(lane, v), g.N - 1);
                    old[jt][v] = C[ic + (long long)j * g.ldc];
                }
        }
#pragma unroll
        for (int jt = 0; jt < 4; jt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = n0 + wc * 64 + jt * 16 + Mfma<T>::row(lane, v);
                T r = g.alpha * acc[it][jt][v] * (i < g.M ? sr[it] : T(0)) * sc[jt][v];
                if (g.dg.p && i == j && i < g.M) r += g.dg.at(b, i);
                if (g.beta != T(0)) r += g.beta * old[jt][v];
                if (i < g.M && j < g.N) C[i + (long long)j * g.ldc] = r;
            }
    }
}

// FP64 variant for op(A) = A or A^T (TA), op(B) = B or B^T (TB) with M, N
// multiples of 128 and 16-byte aligned columns; K a multiple of 16, or,
// with TA = 0 and TB = 1, any per-entry K (Kb: rows past it are read from
// row K-1 and zeroed by the k-scale).  The k-tiles go from memory
// straight into LDS (gfx950 global_load_lds_dwordx4, 16 bytes per lane,
// written at the wave's LDS base + 16 * lane), so no registers stage the
// next tile and no VALU work stores it; the tile of k + 16 is in flight
// while tile k is multiplied.  A tile: k-major rows of 128 doubles (one
// wave instruction each), padded to GD_LD; B tile: j-major, each column's 16
// k as 8 pairs whose order is XOR-swizzled by (j & 7), so the fragment
// reads of 16 columns spread over the banks.  Scalings: k-scales (ac, br)
// multiply the op(A) fragments, row / column scales (ar, cr / bc, cc) are
// applied in the epilogue.  With TB = 1 the B tile is k-major like A; with
// TA = 1 the A tile is i-major with the B tile's swizzled k pairs (a column
// of A is a row of op(A), 16 consecutive k per i).
#define GD_LD 144
#define GD_ABUF (GB_K * GD_LD)          // doubles of one A tile (and of a k-major B tile)
#define GD_BUF (2 * GD_ABUF)             // one buffer: A tile + B tile

static inline size_t gemm_dl_lds_bytes() { return sizeof(double) * 2 * GD_BUF; }
// DQMC_GPU_GEMM_DL=0 keeps every FP64 GEMM on k_gemm_big
static inline bool gemm_dl_enabled()
{
    static const int on = [] {
        const char *e = getenv("DQMC_GPU_GEMM_DL");
        return (e && e[0] == '0') ? 0 : 1;
    }();
    return on != 0;
}

typedef __attribute__((address_space(3))) void lds_void;

template <int TA, int TB>
__global__ __launch_bounds__(256, 2) void k_gemm_dl(GemmArgs<double> g)
{
    static_assert(GB_K == 16 && GB_M == 128, "16 x 128 tiles");
    const int tiles_m = g.M / GB_M, tiles_n = g.N / GB_M;
    const int per = tiles_m * tiles_n;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / per;
    const int t = logical % per;
    const int m0 = (t % tiles_m) * GB_M;
    const int n0 = (t / tiles_m) * GB_M;
    if (b >= g.batch) return;
    if (g.status && g.status[b]) return;
    const int K = g.Kb ? min(g.K, g.Kb[b]) : g.K;
    if (K <= 0 && g.beta == 1.0 && !g.dg.p) return; // C unchanged
    extern __shared__ __align__(16) unsigned char gd_raw[];
    double *lds = reinterpret_cast<double *>(gd_raw);
    const double *A = g.A + (long long)b * g.sA;
    const double *B = g.B + (long long)b * g.sB;
    double *C = g.C + (long long)b * g.sC;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lc = lane & 15, lr = lane >> 4;
    const int wr = wave >> 1, wc = wave & 1;
    // this lane's pieces of the tile loads: A rows k = 4 wave + q, i pair 2 lane;
    // B column groups of 8, column 8 grp + (lane >> 3), k pair (lane & 7) ^ (lane >> 3)
    const double *a_src = A + m0 + 2 * lane;
    const double *b_src = B + n0 + 2 * lane; // TB = 1: rows of op(B) are columns of B
    const int bj = lane >> 3, bkp = (lane & 7) ^ bj;
    auto issue = [&](int u, int k0) {
        double *As = lds + u * GD_BUF;
        double *Bs = As + GD_ABUF;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = 4 * wave + q;
            const int kr = min(k0 + k, K - 1); // rows past K: a valid row, zeroed by the k-scale
            if (TA) {
                const int grp = 4 * wave + q;
                const double *asrc = A + (k0 + 2 * bkp) + (long long)(m0 + 8 * grp + bj) * g.lda;
                __builtin_amdgcn_global_load_lds(asrc, (lds_void *)(As + grp * 8 * GB_K), 16, 0, 0);
            } else {
                __builtin_amdgcn_global_load_lds(a_src + (long long)kr * g.lda, (lds_void *)(As + k * GD_LD), 16, 0, 0);
            }
            if (TB) {
                __builtin_amdgcn_global_load_lds(b_src + (long long)kr * g.ldb, (lds_void *)(Bs + k * GD_LD), 16, 0, 0);
            } else {
                const int grp = 4 * wave + q;
                const double *bsrc = B + (k0 + 2 * bkp) + (long long)(n0 + 8 * grp + bj) * g.ldb;
                __builtin_amdgcn_global_load_lds(bsrc, (lds_void *)(Bs + grp * 8 * GB_K), 16, 0, 0);
            }
        }
    };
    // k-scales of this lane's four k rows of a tile (0 past K): the raw
    // entries are loaded with the tile, one k-step ahead, and combined
    // once the wait for the tile has covered them
    const bool ksc = g.ac.p || g.br.p || (K % GB_K) != 0;
    auto kraw = [&](int k0, double (&ra)[4], double (&rb)[4]) {
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int kc = min(k0 + 4 * kk + lr, K - 1);
            ra[kk] = g.ac.raw(b, kc);
            rb[kk] = g.br.raw(b, kc);
        }
    };

    f64x4 acc[4][4];
#pragma unroll
    for (int it = 0; it < 4; it++)
#pragma unroll
        for (int jt = 0; jt < 4; jt++) acc[it][jt] = f64x4{0, 0, 0, 0};

    double ra[4] = {1.0, 1.0, 1.0, 1.0}, rb[4] = {1.0, 1.0, 1.0, 1.0};
    if (K > 0) {
        if (ksc) kraw(0, ra, rb);
        issue(0, 0);
    }
    int u = 0;
    for (int k0 = 0; k0 < K; k0 += GB_K) {
        __builtin_amdgcn_s_waitcnt(0); // this wave's tile loads (and k-scales) have landed
        __syncthreads();               // ... and everyone's; the other buffer is free
        double sk[4] = {1.0, 1.0, 1.0, 1.0};
        if (ksc) {
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int k = k0 + 4 * kk + lr;
                sk[kk] = k < K ? g.ac.fix(ra[kk]) * g.br.fix(rb[kk]) : 0.0;
            }
        }
        if (k0 + GB_K < K) {
            if (ksc) kraw(k0 + GB_K, ra, rb);
            issue(u ^ 1, k0 + GB_K);
        }
        const double *As = lds + u * GD_BUF;
        const double *Bs = As + GD_ABUF;
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            const int k = kk + lr;
            double fa[4], fb[4];
#pragma unroll
            for (int jt = 0; jt < 4; jt++) {
                const int jl = wc * 64 + jt * 16 + lc;
                fa[jt] = TB ? Bs[k * GD_LD + jl] : Bs[jl * GB_K + 2 * ((k >> 1) ^ (jl & 7)) + (k & 1)];
            }
#pragma unroll
            for (int it = 0; it < 4; it++) {
                const int il = wr * 64 + it * 16 + lc;
                fb[it] = (TA ? As[il * GB_K + 2 * ((k >> 1) ^ (il & 7)) + (k & 1)] : As[k * GD_LD + il]) * sk[kk >> 2];
            }
#pragma unroll
            for (int it = 0; it < 4; it++)
#pragma unroll
                for (int jt = 0; jt < 4; jt++) acc[it][jt] = Mfma<double>::mma(fa[jt], fb[it], acc[it][jt]);
        }
        u ^= 1;
    }

    // epilogue as in k_gemm_big, with the row scales of op(A) (ar) and the
    // column scales of op(B) (bc) folded in
    double sr[4], sc[4][4];
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const int i = m0 + wr * 64 + it * 16 + lc;
        sr[it] = g.cr.fix(g.cr.raw(b, i)) * g.ar.fix(g.ar.raw(b, i));
    }
#pragma unroll
    for (int jt = 0; jt < 4; jt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int j = n0 + wc * 64 + jt * 16 + Mfma<double>::row(lane, v);
            sc[jt][v] = g.cc.fix(g.cc.raw(b, j)) * g.bc.fix(g.bc.raw(b, j));
        }
    auto store_block = [&](int it, int jt, const double (&old)[4], bool with_old) {
        const int i = m0 + wr * 64 + it * 16 + lc;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int j = n0 + wc * 64 + jt * 16 + Mfma<double>::row(lane, v);
            double r = g.alpha * acc[it][jt][v] * sr[it] * sc[jt][v];
            if (g.dg.p && i == j) r += g.dg.at(b, i);
            if (with_old) r += g.beta * old[v];
            C[i + (long long)j * g.ldc] = r;
        }
    };
    if (g.beta != 0.0) {
        // C += ...: the old values of the next 16 x 16 block are loaded
        // while this block is combined and stored
        auto load_block = [&](int it, int jt, double (&old)[4]) {
            const int i = m0 + wr * 64 + it * 16 + lc;
#pragma unroll
            for (int v = 0; v < 4; v++)
                old[v] = C[i + (long long)(n0 + wc * 64 + jt * 16 + Mfma<double>::row(lane, v)) * g.ldc];
        };
        double old[2][4];
        load_block(0, 0, old[0]);
#pragma unroll
        for (int blk = 0; blk < 16; blk++) {
            if (blk + 1 < 16) load_block((blk + 1) >> 2, (blk + 1) & 3, old[(blk + 1) & 1]);
            store_block(blk >> 2, blk & 3, old[blk & 1], true);
        }
    } else {
        const double none[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int it = 0; it < 4; it++)
#pragma unroll
            for (int jt = 0; jt < 4; jt++) store_block(it, jt, none, false);
    }
}

template <typename T>
static inline GemmArgs<T> gemm_args(int M, int N, int K, int batch, const T *A, long long lda,
                                    long long sA, int transA, const T *B, long long ldb,
                                    long long sB, int transB, T *C, long long ldc, long long sC)
{
    GemmArgs<T> g;
    g.M = M; g.N = N; g.K = K; g.batch = batch;
    g.A = A; g.lda = lda; g.sA = sA; g.transA = transA;
    g.B = B; g.ldb = ldb; g.sB = sB; g.transB = transB;
    g.C = C; g.ldc = ldc; g.sC = sC;
    g.alpha = T(1); g.beta = T(0);
    g.ar = g.ac = g.br = g.bc = g.cr = g.cc = g.dg = vnone<T>();
    g.status = nullptr;
    g.Kb = nullptr;
    g.vec2 = 0;
    return g;
}

template <typename T>
static inline void gemm_launch(const GemmArgs<T> &g0, hipStream_t s)
{
    GemmArgs<T> g = g0;
    {
        const uintptr_t al = 2 * sizeof(T) - 1;
        const bool even_k = (g.K % 2 == 0) && !g.Kb;
        g.vec2 = ((reinterpret_cast<uintptr_t>(g.A) & al) == 0) && ((reinterpret_cast<uintptr_t>(g.B) & al) == 0) &&
                 g.lda % 2 == 0 && g.ldb % 2 == 0 && g.sA % 2 == 0 && g.sB % 2 == 0 &&
                 (g.transA ? even_k : g.M % 2 == 0) && (g.transB ? g.N % 2 == 0 : even_k);
    }
    if constexpr (sizeof(T) == 8) {
        const bool aligned = (reinterpret_cast<uintptr_t>(g.A) & 15) == 0 && (reinterpret_cast<uintptr_t>(g.B) & 15) == 0 &&
                             g.lda % 2 == 0 && g.ldb % 2 == 0 && g.sA % 2 == 0 && g.sB % 2 == 0;
        // a k-major tile (op(A) = A, op(B) = B^T) takes any per-entry K; an
        // i- or j-major one needs K a multiple of 16
        const bool kfull = !g.Kb && g.K % GB_K == 0;
        if (aligned && g.M % GB_M == 0 && g.N % GB_M == 0 && g.K > 0 && gemm_dl_enabled() &&
            (!g.transA || kfull) && (g.transB || kfull)) {
            const int tb = (g.M / GB_M) * (g.N / GB_M) * g.batch;
            if (tb <= 0) return;
            const size_t lds = gemm_dl_lds_bytes();
            if (g.transA && g.transB) hipLaunchKernelGGL((k_gemm_dl<1, 1>), dim3(tb), dim3(256), lds, s, g);
            else if (g.transA) hipLaunchKernelGGL((k_gemm_dl<1, 0>), dim3(tb), dim3(256), lds, s, g);
            else if (g.transB) hipLaunchKernelGGL((k_gemm_dl<0, 1>), dim3(tb), dim3(256), lds, s, g);
            else hipLaunchKernelGGL((k_gemm_dl<0, 0>), dim3(tb), dim3(256), lds, s, g);
            return;
        }
    }
    if (g.M >= GB_M && g.N >= GB_M && g.vec2) {
        const int tb = ((g.M + GB_M - 1) / GB_M) * ((g.N + GB_M - 1) / GB_M) * g.batch;
        if (tb <= 0) return;
        const size_t lds = gemm_big_lds_bytes<T>();
        if (!g.transA && !g.transB) hipLaunchKernelGGL((k_gemm_big<T, 0, 0>), dim3(tb), dim3(256), lds, s, g);
        else if (g.transA && !g.transB) hipLaunchKernelGGL((k_gemm_big<T, 1, 0>), dim3(tb), dim3(256), lds, s, g);
        else if (!g.transA && g.transB) hipLaunchKernelGGL((k_gemm_big<T, 0, 1>), dim3(tb), dim3(256), lds, s, g);
        else hipLaunchKernelGGL((k_gemm_big<T, 1, 1>), dim3(tb), dim3(256), lds, s, g);
        return;
    }
    const int tiles = ((g.M + GEMM_BM - 1) / GEMM_BM) * ((g.N + GEMM_BN - 1) / GEMM_BN);
    const int total = tiles * g.batch;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_gemm<T>, dim3(total), dim3(256), 0, s, g);
}
