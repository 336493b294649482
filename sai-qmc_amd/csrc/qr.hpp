// FP64 factorisations for the UDV chain, one workgroup per batch entry.
//
// k_qr_udv: blocked right-looking Householder QR with reflectors in LAPACK
// dlarfg's convention (R(j,j) = -sign(alpha) ||x||, so D = diag R carries
// the reference's signs, src/linalg.c:575-612) and explicit formation of Q
// (dorgqr, src/linalg.c:596).  It emits the UDV pieces the reference
// assembles by hand: Q, D = diag R and Rn = D^-1 R (unit upper, zeros below)
// (src/linalg.c:614-627), plus sgn det Q = (-1)^(number of non-trivial
// reflectors), which replaces the LU determinant of U in the sign
// (src/linalg.c:1044-1083).
//
//   panel (QR_NB columns, resident in LDS): columns are owned round-robin by
//     the waves; step c needs one barrier: every wave applies reflector c to
//     the columns it owns using the unscaled column and the scalar
//     1/(alpha-beta), and the owner of column c+1 publishes its new norm;
//   T factor of the compact WY form (dlarft, forward, columnwise);
//   trailing update A <- (I - V T^T V^T) A and Q <- (I - V T V^T) Q on FP64
//     MFMA 16x16x4: each wave takes 16-column blocks, W = V^T A (2 row tiles,
//     K = panel height), W' = op(T) W, A -= V W'.
//
// k_trtri_unit: inverse of a unit upper-triangular matrix
// (dtrtri('U','U'), src/linalg.c:992-1004), one wave per 32-column block
// column, back substitution over 32-row blocks with MFMA products.
#pragma once
#include "common.hpp"

#define QR_NB 32
#define QR_THREADS 256
#define QR_WAVES (QR_THREADS / 64)

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

// reflector entry v(c, r) of the panel P (column c, row r, leading dim ld):
// 1 on the diagonal, stored values below, 0 above
__device__ __forceinline__ double panel_v(const double *P, int ld, int c, int r)
{
    return r > c ? P[c * ld + r] : (r == c ? 1.0 : 0.0);
}

// X[p0:p0+m, j0:j0+16] <- (I - V op(T) V^T) X[...] for one 16-column block,
// one wave.  V is the m x nb panel P (LDS), T is Tl (LDS, QR_NB x QR_NB,
// Tl[i + c*QR_NB], upper), Wl a wave-private LDS buffer of QR_NB x 16.
// Columns at or beyond jend are left untouched (loads masked, stores skipped).
__device__ void wy_apply_block(const double *P, int ld, const double *Tl, int nb, int m, double *X,
                               int n, int p0, int j0, int jend, int transT, double *Wl)
{
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15, lr = lane >> 4;
    const int j = j0 + lc;
    const bool jok = j < jend;
    double *col = X + (long long)(jok ? j : j0) * n + p0;
    // W = V^T X  (QR_NB x 16): row tile t covers reflectors 16t..16t+15.
    // Eight k-steps per batch so eight global loads are in flight per lane.
    f64x4 w0 = {0, 0, 0, 0}, w1 = {0, 0, 0, 0};
    for (int kb = 0; kb < m; kb += 32) {
        double bv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = kb + 4 * u + lr;
            bv[u] = (jok && k < m) ? col[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = kb + 4 * u + lr;
            const double a0 = (lc < nb && k < m) ? panel_v(P, ld, lc, k) : 0.0;
            const double a1 = (16 + lc < nb && k < m) ? panel_v(P, ld, 16 + lc, k) : 0.0;
            w0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv[u], w0, 0, 0, 0);
            w1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv[u], w1, 0, 0, 0);
        }
    }
    // spill W to LDS: W[i][jj] at Wl[i*16 + jj]
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int r = Mfma<double>::row(lane, v);
        Wl[r * 16 + lc] = w0[v];
        Wl[(16 + r) * 16 + lc] = w1[v];
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    // W' = op(T) W: op(T)[i][k] = T[k][i] (transT) or T[i][k]
    f64x4 u0 = {0, 0, 0, 0}, u1 = {0, 0, 0, 0};
    for (int k0 = 0; k0 < QR_NB; k0 += 4) {
        const int k = k0 + lr;
        const double b = Wl[k * 16 + lc];
        const double t0 = transT ? Tl[k + lc * QR_NB] : Tl[lc + k * QR_NB];
        const double t1 = transT ? Tl[k + (16 + lc) * QR_NB] : Tl[(16 + lc) + k * QR_NB];
        u0 = __builtin_amdgcn_mfma_f64_16x16x4f64(t0, b, u0, 0, 0, 0);
        u1 = __builtin_amdgcn_mfma_f64_16x16x4f64(t1, b, u1, 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int r = Mfma<double>::row(lane, v);
        Wl[r * 16 + lc] = u0[v];
        Wl[(16 + r) * 16 + lc] = u1[v];
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    // X -= V W' over 16-row tiles, four tiles (16 loads per lane) in flight
    for (int rb = 0; rb < m; rb += 64) {
        f64x4 c[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = rb + 16 * u + Mfma<double>::row(lane, v);
                c[u][v] = (jok && r < m) ? col[r] : 0.0;
            }
        for (int k0 = 0; k0 < QR_NB; k0 += 4) {
            const int k = k0 + lr;
            const double wv = Wl[k * 16 + lc];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = rb + 16 * u + lc;
                const double a = (k < nb && r < m) ? -panel_v(P, ld, k, r) : 0.0;
                c[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, wv, c[u], 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = rb + 16 * u + Mfma<double>::row(lane, v);
                if (jok && r < m) col[r] = c[u][v];
            }
    }
    __builtin_amdgcn_wave_barrier();
}

// Copy an m x nb block of global memory (leading dimension n) into the LDS
// panel P (leading dimension ld) or back, eight elements per thread in flight.
__device__ __forceinline__ void panel_load(double *P, int ld, const double *A, int n, int nb, int m, int *bad)
{
    const int total = nb * m;
    for (int base = threadIdx.x; base < total; base += 8 * blockDim.x) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = base + u * blockDim.x;
            v[u] = idx < total ? A[(idx % m) + (long long)(idx / m) * n] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int idx = base + u * blockDim.x;
            if (idx < total) {
                if (bad && !isfinite(v[u])) *bad = 1;
                P[(idx / m) * ld + (idx % m)] = v[u];
            }
        }
    }
}

__global__ __launch_bounds__(QR_THREADS) void k_qr_udv(QrArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int n = a.n, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    double *A = a.A + (long long)b * a.sA;
    double *Tw = a.Tws + (long long)b * a.sT;

    extern __shared__ double lds[];
    double *P = lds;                     // QR_NB x n panel, P[c*n + r]
    double *Tl = P + (size_t)QR_NB * n;  // QR_NB x QR_NB
    double *Wl = Tl + QR_NB * QR_NB;     // QR_WAVES x (QR_NB x 16)
    double *nrm = Wl + QR_WAVES * QR_NB * 16; // QR_NB column norms^2 (below diag)
    double *scl = nrm + QR_NB;           // QR_NB 1/(alpha - beta)
    __shared__ int s_bad, s_ntau;
    if (tid == 0) { s_bad = 0; s_ntau = 0; }

    const int npan = (n + QR_NB - 1) / QR_NB;
    for (int pi = 0; pi < npan; pi++) {
        const int p0 = pi * QR_NB;
        const int nb = min(QR_NB, n - p0);
        const int m = n - p0;
        __syncthreads();
        panel_load(P, n, A + p0 + (long long)p0 * n, n, nb, m, &s_bad);
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tl[idx] = 0.0;
        __syncthreads();
        if (wave == 0) {
            double s = 0.0;
            for (int r = 1 + lane; r < m; r += 64) s += P[r] * P[r];
            s = wave_sum(s);
            if (lane == 0) nrm[0] = s;
        }
        __syncthreads();
        // ---- panel: one barrier per column
        for (int c = 0; c < nb; c++) {
            const double alpha = P[c * n + c];
            const double xn2 = nrm[c];
            double tau = 0.0, sc = 0.0, beta = alpha;
            if (xn2 != 0.0) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tau = (beta - alpha) / beta;
                sc = 1.0 / (alpha - beta);
            }
            if (tid == 0) {
                if (xn2 != 0.0) s_ntau++;
                if (!isfinite(tau) || !isfinite(sc)) s_bad = 1;
                Tl[c + c * QR_NB] = tau;
                scl[c] = sc;
            }
            // reflector v = [1; sc * x]; columns c2 > c owned by this wave
            for (int c2 = c + 1; c2 < nb; c2++) {
                if ((c2 % QR_WAVES) != wave) continue;
                double d = 0.0;
                if (tau != 0.0) {
                    for (int r = c + 1 + lane; r < m; r += 64) d += P[c * n + r] * P[c2 * n + r];
                    d = wave_sum(d);
                    const double w = tau * (P[c2 * n + c] + sc * d);
                    const double ws = w * sc;
                    for (int r = c + 1 + lane; r < m; r += 64) P[c2 * n + r] -= ws * P[c * n + r];
                    if (lane == 0) P[c2 * n + c] -= w;
                }
                if (c2 == c + 1) {
                    double s = 0.0;
                    for (int r = c2 + 1 + lane; r < m; r += 64) s += P[c2 * n + r] * P[c2 * n + r];
                    s = wave_sum(s);
                    if (lane == 0) nrm[c2] = s;
                }
            }
            __syncthreads();
            if (tid == 0 && xn2 != 0.0) P[c * n + c] = beta;
        }
        __syncthreads();
        // scale the stored reflectors: v = sc * x below the diagonal
        for (int idx = tid; idx < nb * m; idx += blockDim.x) {
            const int c = idx / m, r = idx % m;
            if (r > c && scl[c] != 0.0) P[c * n + r] *= scl[c];
        }
        __syncthreads();
        // ---- T factor: Y(i,c) = v_i^T v_c for i < c into Wl scratch, then
        // T(0:c, c) = -tau_c T(0:c, 0:c) Y(0:c, c)
        double *Y = Wl;
        for (int pr = wave; pr < nb * nb; pr += QR_WAVES) {
            const int i = pr % nb, c = pr / nb;
            if (i >= c) continue;
            double s = 0.0;
            for (int r = c + lane; r < m; r += 64) s += panel_v(P, n, i, r) * panel_v(P, n, c, r);
            s = wave_sum(s);
            if (lane == 0) Y[i + c * QR_NB] = s;
        }
        __syncthreads();
        if (wave == 0) {
            for (int c = 1; c < nb; c++) {
                if (lane < c) {
                    double s = 0.0;
                    for (int k = lane; k < c; k++) s += Tl[lane + k * QR_NB] * Y[k + c * QR_NB];
                    Tl[lane + c * QR_NB] = -Tl[c + c * QR_NB] * s;
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tw[pi * QR_NB * QR_NB + idx] = Tl[idx];
        for (int idx = tid; idx < nb * m; idx += blockDim.x) {
            const int c = idx / m, r = idx % m;
            A[(p0 + r) + (long long)(p0 + c) * n] = P[c * n + r];
        }
        __syncthreads();
        // ---- trailing update A[p0:, p0+nb:] <- (I - V T^T V^T) A
        for (int j0 = p0 + nb + 16 * wave; j0 < n; j0 += 16 * QR_WAVES)
            wy_apply_block(P, n, Tl, nb, m, A, n, p0, j0, n, 1, Wl + wave * QR_NB * 16);
    }
    __syncthreads();

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
        for (int pi = npan - 1; pi >= 0; pi--) {
            const int p0 = pi * QR_NB;
            const int nb = min(QR_NB, n - p0);
            const int m = n - p0;
            __syncthreads();
            panel_load(P, n, A + p0 + (long long)p0 * n, n, nb, m, nullptr);
            for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tl[idx] = Tw[pi * QR_NB * QR_NB + idx];
            __syncthreads();
            for (int j0 = p0 + 16 * wave; j0 < n; j0 += 16 * QR_WAVES)
                wy_apply_block(P, n, Tl, nb, m, Q, n, p0, j0, n, 0, Wl + wave * QR_NB * 16);
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
    return sizeof(double) * ((size_t)QR_NB * n + QR_NB * QR_NB + QR_WAVES * QR_NB * 16 + 2 * QR_NB);
}

// X = T^-1 for unit upper-triangular T (the diagonal of T is not read).
struct TrtriArgs {
    int n, batch;
    const double *T;
    long long sT;
    double *X;
    long long sX;
    const int *status;
};

#define TR_NB 32

// One wave per 32-wide block column J of X: for block rows I = J down to 0,
// X_IJ = T_II^-1 (delta_IJ - sum_{K>I} T_IK X_KJ); the sum is a 32x32 MFMA
// product accumulated over 32-row blocks of K (read back from X in global
// memory, written earlier by the same wave), the 32x32 unit solve runs one
// column per lane.
__global__ __launch_bounds__(256) void k_trtri_unit(TrtriArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int n = a.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const double *T = a.T + (long long)b * a.sT;
    double *X = a.X + (long long)b * a.sX;
    extern __shared__ double lds[];
    double *S = lds + (size_t)wave * TR_NB * TR_NB; // per-wave 32x32 rhs, S[i + jj*32]
    const int nblk = (n + TR_NB - 1) / TR_NB;
    for (int J = wave; J < nblk; J += nw) {
        const int c0 = J * TR_NB;
        const int jb = min(TR_NB, n - c0);
        // zero the part of the block column below its diagonal block
        for (long long idx = lane; idx < (long long)(n - (c0 + jb)) * jb; idx += 64) {
            const int r = c0 + jb + (int)(idx % (n - c0 - jb)), jj = (int)(idx / (n - c0 - jb));
            X[r + (long long)(c0 + jj) * n] = 0.0;
        }
        for (int I = J; I >= 0; I--) {
            const int r0 = I * TR_NB;
            const int ib = min(TR_NB, n - r0);
            // S = delta - T[r0:r0+ib, r0+ib : c0+jb] X[r0+ib : c0+jb, c0:c0+jb]
            for (int ti = 0; ti < 2; ti++)
                for (int tj = 0; tj < 2; tj++) {
                    f64x4 acc = {0, 0, 0, 0};
                    const int i = r0 + ti * 16 + lc;
                    const int jj = tj * 16 + lc;
                    for (int kb = r0 + ib; kb < c0 + jb; kb += 32) {
                        double ta[8], xb[8];
#pragma unroll
                        for (int u = 0; u < 8; u++) {
                            const int k = kb + 4 * u + lr;
                            const bool kok = k < c0 + jb;
                            ta[u] = (kok && i < r0 + ib) ? T[i + (long long)k * n] : 0.0;
                            xb[u] = (kok && jj < jb) ? X[k + (long long)(c0 + jj) * n] : 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < 8; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ta[u], xb[u], acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int i = ti * 16 + Mfma<double>::row(lane, v);
                        const int jj = tj * 16 + lc;
                        const double d = (r0 + i == c0 + jj) ? 1.0 : 0.0;
                        S[i + jj * TR_NB] = d - acc[v];
                    }
                }
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_wave_barrier();
            // unit upper solve T_II x = S(:, jj), one column per lane
            if (lane < jb) {
                const int jj = lane;
                for (int i = ib - 1; i >= 0; i--) {
                    double s = S[i + jj * TR_NB];
                    for (int k = i + 1; k < ib; k++) s -= T[(r0 + i) + (long long)(r0 + k) * n] * S[k + jj * TR_NB];
                    S[i + jj * TR_NB] = s;
                    X[(r0 + i) + (long long)(c0 + jj) * n] = s;
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

static inline size_t trtri_lds_bytes(int n)
{
    (void)n;
    return sizeof(double) * (size_t)4 * TR_NB * TR_NB;
}
