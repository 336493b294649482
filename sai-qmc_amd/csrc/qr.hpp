// FP64 factorisations for the UDV chain, one workgroup per batch entry.
//
// k_qr_udv: blocked right-looking Householder QR (panel width QR_NB, compact
// WY with the T factor of LAPACK dlarft, reflectors in dlarfg's convention:
// R(j,j) = -sign(alpha) ||x||, so D = diag R carries the reference's signs,
// src/linalg.c:575-612) followed by explicit formation of Q (dorgqr,
// src/linalg.c:596).  It emits the UDV pieces the reference assembles by
// hand: Q, D = diag R and Rn = D^-1 R (unit upper, zeros below)
// (src/linalg.c:614-627), plus sgn det Q = (-1)^(number of non-trivial
// reflectors), which replaces the LU determinant of U in the sign
// (src/linalg.c:1044-1083).
//
// k_trtri_unit: inverse of a unit upper-triangular matrix
// (dtrtri('U','U'), src/linalg.c:992-1004), block-column back substitution
// with the block column resident in LDS.
#pragma once
#include "common.hpp"

#define QR_NB 16
#define QR_THREADS 256

struct QrArgs {
    int n, batch;
    double *A;   // in: matrix, destroyed (R above, reflectors below)
    long long sA;
    double *Q;   // out: explicit Q (may be null)
    long long sQ;
    double *Rn;  // out: D^-1 R with zeros below the diagonal (may be null)
    long long sR;
    double *D;   // out: diag R
    long long sD;
    double *Tws; // workspace: ceil(n/QR_NB) * QR_NB * QR_NB per entry
    long long sT;
    int *qsign;  // out: sgn det Q (may be null)
    long long sq;
    int *status, *reason;
};

// reflector entry v(c, r) of the panel stored in P (column c, row r, leading
// dimension ld): 1 on the diagonal, stored values below, 0 above
__device__ __forceinline__ double panel_v(const double *P, int ld, int c, int r)
{
    return r > c ? P[c * ld + r] : (r == c ? 1.0 : 0.0);
}

// Apply (I - V op(T) V^T) to columns [j_begin, n) of the m-row block X
// (global, rows p0.., leading dimension n).  transT selects T^T (for Q^T A
// in the factorisation) or T (for forming Q).  One wave per column.
__device__ void apply_block_reflector(const double *P, int ld, const double *Tl, int nb, int m,
                                      double *X, int n, int p0, int j_begin, int transT)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    for (int j = j_begin + wave; j < n; j += nw) {
        double *col = X + (long long)j * n + p0;
        double w[QR_NB];
#pragma unroll
        for (int c = 0; c < QR_NB; c++) w[c] = 0.0;
        for (int r = lane; r < m; r += 64) {
            const double a = col[r];
#pragma unroll
            for (int c = 0; c < QR_NB; c++)
                if (c < nb) w[c] += panel_v(P, ld, c, r) * a;
        }
#pragma unroll
        for (int c = 0; c < QR_NB; c++) w[c] = wave_sum(w[c]);
        double wp[QR_NB];
#pragma unroll
        for (int c = 0; c < QR_NB; c++) {
            double s = 0.0;
            if (c < nb) {
                if (transT) {
                    for (int i = 0; i <= c; i++) s += Tl[i + c * QR_NB] * w[i];
                } else {
                    for (int k = c; k < nb; k++) s += Tl[c + k * QR_NB] * w[k];
                }
            }
            wp[c] = s;
        }
        for (int r = lane; r < m; r += 64) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < QR_NB; c++)
                if (c < nb) s += panel_v(P, ld, c, r) * wp[c];
            col[r] -= s;
        }
    }
}

__global__ __launch_bounds__(QR_THREADS) void k_qr_udv(QrArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int n = a.n, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    double *A = a.A + (long long)b * a.sA;
    double *Tw = a.Tws + (long long)b * a.sT;

    extern __shared__ double lds[];
    double *P = lds;                    // QR_NB x n panel, P[c*n + r]
    double *Tl = P + QR_NB * n;         // QR_NB x QR_NB, Tl[i + c*QR_NB]
    double *Y = Tl + QR_NB * QR_NB;     // QR_NB x QR_NB dots
    double *red = Y + QR_NB * QR_NB;    // reduction slots
    __shared__ double s_tau, s_sc;
    __shared__ int s_bad, s_ntau;
    if (tid == 0) { s_bad = 0; s_ntau = 0; }
    __syncthreads();

    const int npan = (n + QR_NB - 1) / QR_NB;
    for (int pi = 0; pi < npan; pi++) {
        const int p0 = pi * QR_NB;
        const int nb = min(QR_NB, n - p0);
        const int m = n - p0;
        for (int idx = tid; idx < nb * m; idx += blockDim.x) {
            const int c = idx / m, r = idx % m;
            const double v = A[(p0 + r) + (long long)(p0 + c) * n];
            if (!isfinite(v)) s_bad = 1;
            P[c * n + r] = v;
        }
        __syncthreads();
        for (int c = 0; c < nb; c++) {
            double part = 0.0;
            for (int r = c + 1 + tid; r < m; r += blockDim.x) part += P[c * n + r] * P[c * n + r];
            const double xnorm2 = block_sum(part, red);
            if (tid == 0) {
                const double alpha = P[c * n + c];
                double tau = 0.0, sc = 1.0;
                if (xnorm2 != 0.0) {
                    const double beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
                    tau = (beta - alpha) / beta;
                    sc = 1.0 / (alpha - beta);
                    P[c * n + c] = beta;
                    s_ntau++;
                }
                if (!isfinite(tau) || !isfinite(sc)) s_bad = 1;
                s_tau = tau;
                s_sc = sc;
                Tl[c + c * QR_NB] = tau;
            }
            __syncthreads();
            const double tau = s_tau, sc = s_sc;
            if (tau != 0.0)
                for (int r = c + 1 + tid; r < m; r += blockDim.x) P[c * n + r] *= sc;
            __syncthreads();
            if (tau != 0.0) {
                for (int c2 = c + 1 + wave; c2 < nb; c2 += nw) {
                    double part2 = 0.0;
                    for (int r = c + 1 + lane; r < m; r += 64) part2 += P[c * n + r] * P[c2 * n + r];
                    const double w = tau * (wave_sum(part2) + P[c2 * n + c]);
                    if (lane == 0) P[c2 * n + c] -= w;
                    for (int r = c + 1 + lane; r < m; r += 64) P[c2 * n + r] -= P[c * n + r] * w;
                }
            }
            __syncthreads();
        }
        // T factor (dlarft, forward, columnwise): Y(i,c) = v_i^T v_c, i < c
        for (int pr = wave; pr < nb * nb; pr += nw) {
            const int i = pr % nb, c = pr / nb;
            if (i >= c) continue;
            double part = 0.0;
            for (int r = c + lane; r < m; r += 64) part += panel_v(P, n, i, r) * panel_v(P, n, c, r);
            part = wave_sum(part);
            if (lane == 0) Y[i + c * QR_NB] = part;
        }
        __syncthreads();
        for (int c = 1; c < nb; c++) {
            // T(0:c, c) = -tau_c T(0:c, 0:c) Y(0:c, c)
            if (tid < c) {
                const double tc = Tl[c + c * QR_NB];
                double s = 0.0;
                for (int k = tid; k < c; k++) s += Tl[tid + k * QR_NB] * Y[k + c * QR_NB];
                Tl[tid + c * QR_NB] = -tc * s;
            }
            __syncthreads();
        }
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) {
            const int i = idx % QR_NB, c = idx / QR_NB;
            Tw[pi * QR_NB * QR_NB + idx] = (i < nb && c < nb && i <= c) ? Tl[idx] : 0.0;
        }
        for (int idx = tid; idx < nb * m; idx += blockDim.x) {
            const int c = idx / m, r = idx % m;
            A[(p0 + r) + (long long)(p0 + c) * n] = P[c * n + r];
        }
        __syncthreads();
        apply_block_reflector(P, n, Tl, nb, m, A, n, p0, p0 + nb, 1);
        __syncthreads();
    }

    // D = diag R, Rn = D^-1 R
    double *Dd = a.D + (long long)b * a.sD;
    for (int i = tid; i < n; i += blockDim.x) {
        const double d = A[i + (long long)i * n];
        Dd[i] = d;
        if (!isfinite(d) || d == 0.0) s_bad = (s_bad == 1) ? 1 : 2;
    }
    if (a.Rn) {
        double *Rn = a.Rn + (long long)b * a.sR;
        for (long long idx = tid; idx < (long long)n * n; idx += blockDim.x) {
            const int i = (int)(idx % n), j = (int)(idx / n);
            Rn[idx] = (i <= j) ? A[idx] / A[i + (long long)i * n] : 0.0;
        }
    }
    if (a.Q) {
        double *Q = a.Q + (long long)b * a.sQ;
        for (long long idx = tid; idx < (long long)n * n; idx += blockDim.x) {
            const int i = (int)(idx % n), j = (int)(idx / n);
            Q[idx] = (i == j) ? 1.0 : 0.0;
        }
        __syncthreads();
        for (int pi = npan - 1; pi >= 0; pi--) {
            const int p0 = pi * QR_NB;
            const int nb = min(QR_NB, n - p0);
            const int m = n - p0;
            for (int idx = tid; idx < nb * m; idx += blockDim.x) {
                const int c = idx / m, r = idx % m;
                P[c * n + r] = A[(p0 + r) + (long long)(p0 + c) * n];
            }
            for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x)
                Tl[idx] = Tw[pi * QR_NB * QR_NB + idx];
            __syncthreads();
            apply_block_reflector(P, n, Tl, nb, m, Q, n, p0, p0, 0);
            __syncthreads();
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (a.qsign) a.qsign[(long long)b * a.sq] = (s_ntau & 1) ? -1 : 1;
        if (s_bad && a.status) {
            a.status[b] = 1;
            if (a.reason) a.reason[b] = (s_bad == 1) ? 16 : 17;
        }
    }
}

static inline size_t qr_lds_bytes(int n)
{
    return sizeof(double) * ((size_t)QR_NB * n + 2 * QR_NB * QR_NB + 16);
}

// X = T^-1 for unit upper-triangular T (diagonal of T ignored).
struct TrtriArgs {
    int n, batch;
    const double *T;
    long long sT;
    double *X;
    long long sX;
    const int *status;
};

#define TR_NB 16

__global__ __launch_bounds__(256) void k_trtri_unit(TrtriArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int n = a.n, tid = threadIdx.x;
    const double *T = a.T + (long long)b * a.sT;
    double *X = a.X + (long long)b * a.sX;
    extern __shared__ double lds[];
    double *Xc = lds;              // n x TR_NB block column, Xc[r + jj*n]
    double *rhs = Xc + (long long)TR_NB * n; // TR_NB x TR_NB
    for (int J0 = 0; J0 < n; J0 += TR_NB) {
        const int jb = min(TR_NB, n - J0);
        for (int I0 = ((J0) / TR_NB) * TR_NB; I0 >= 0; I0 -= TR_NB) {
            const int ib = min(TR_NB, n - I0);
            // rhs(i, jj) = delta - sum_{k = I0+ib}^{J0+jb-1} T(I0+i, k) Xc(k, jj)
            {
                const int i = tid % TR_NB, jj = tid / TR_NB;
                if (i < ib && jj < jb) {
                    double s = (I0 + i == J0 + jj) ? 1.0 : 0.0;
                    for (int k = I0 + ib; k < J0 + jb; k++)
                        s -= T[(I0 + i) + (long long)k * n] * Xc[k + jj * n];
                    rhs[i + jj * TR_NB] = s;
                }
            }
            __syncthreads();
            // unit upper solve T(I0:I0+ib, I0:I0+ib) x = rhs, column jj per thread
            if (tid < jb) {
                const int jj = tid;
                for (int i = ib - 1; i >= 0; i--) {
                    double s = rhs[i + jj * TR_NB];
                    for (int k = i + 1; k < ib; k++) s -= T[(I0 + i) + (long long)(I0 + k) * n] * Xc[I0 + k + jj * n];
                    Xc[I0 + i + jj * n] = s;
                }
            }
            __syncthreads();
        }
        for (long long idx = tid; idx < (long long)n * jb; idx += blockDim.x) {
            const int r = (int)(idx % n), jj = (int)(idx / n);
            X[r + (long long)(J0 + jj) * n] = (r < J0 + jb) ? Xc[r + jj * n] : 0.0;
        }
        __syncthreads();
    }
}

static inline size_t trtri_lds_bytes(int n)
{
    return sizeof(double) * ((size_t)TR_NB * n + TR_NB * TR_NB);
}
