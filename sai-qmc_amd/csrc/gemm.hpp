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

    for (int k0 = 0; k0 < g.K; k0 += GEMM_BK) {
#pragma unroll
        for (int r = 0; r < (GEMM_BM * GEMM_BK) / 256; r++) {
            const int e = tid + 256 * r;
            int i, k;
            if (!g.transA) { i = e % GEMM_BM; k = e / GEMM_BM; }
            else { k = e % GEMM_BK; i = e / GEMM_BK; }
            const int gi = m0 + i, gk = k0 + k;
            T v = 0;
            if (gi < g.M && gk < g.K) {
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
            if (gk < g.K && gj < g.N) {
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

#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = m0 + wr * 32 + mi * 16 + Mfma<T>::row(lane, v);
                const int j = n0 + wc * 32 + nj * 16 + (lane & 15);
                if (i < g.M && j < g.N) {
                    T r = g.alpha * acc[mi][nj][v] * g.cr.at(b, i) * g.cc.at(b, j);
                    if (g.dg.p && i == j) r += g.dg.at(b, i);
                    T *cp = &C[i + (long long)j * g.ldc];
                    if (g.beta != T(0)) r += g.beta * *cp;
                    *cp = r;
                }
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
    return g;
}

template <typename T>
static inline void gemm_launch(const GemmArgs<T> &g, hipStream_t s)
{
    const int tiles = ((g.M + GEMM_BM - 1) / GEMM_BM) * ((g.N + GEMM_BN - 1) / GEMM_BN);
    const int total = tiles * g.batch;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_gemm<T>, dim3(total), dim3(256), 0, s, g);
}
