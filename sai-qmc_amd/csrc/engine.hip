// MI355X DQMC Green's-function engine: device buffers, the forward sweep
// schedule and the C ABI declared in include/dqmc_gpu.h.
//
// One "model" groups the replicas that share (lattice, U, dtau, L, stab).
// Each model owns a HIP stream; every stage is one launch over all of the
// model's replicas (batched GEMMs over the whole chip, one workgroup per
// replica for the sequential stages).  A sweep's launch sequence is fixed,
// so it is captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dqmc_gpu.h"
#include "common.hpp"
#include "gemm.hpp"
#include "kron.hpp"
#include "measure.hpp"
#include "metro.hpp"
#include "metro_sub.hpp"
#include "persist.hpp"
#include "qr.hpp"

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static void set_err(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            set_err("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__,   \
                    __LINE__);                                                      \
            return 1;                                                               \
        }                                                                           \
    } while (0)

// ------------------------------------------------------------ small kernels

__global__ void k_udv_identity(double *U, long long sU, double *D, long long sD, double *T,
                               long long sT, int *q, long long sq, int n, int batch, double *off, long long so)
{
    const int b = blockIdx.y;
    if (b >= batch) return;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
         idx += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        const double v = (i == j) ? 1.0 : 0.0;
        U[b * sU + idx] = v;
        T[b * sT + idx] = v;
        if (j == 0) D[b * sD + i] = 1.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        q[b * sq] = 1;
        if (off) off[b * so] = 0.0; // udv_identity (src/linalg.c:432-441)
    }
}

__global__ void k_copy_d(const double *src, long long sS, double *dst, long long sD, long long len,
                         int batch)
{
    const int b = blockIdx.y;
    if (b >= batch) return;
    // four elements per thread, loads from clamped indices issued together
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long base = blockIdx.x * (long long)blockDim.x + threadIdx.x; base < len; base += 4 * stride) {
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = src[b * sS + min(base + k * stride, len - 1)];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (base + k * stride < len) dst[b * sD + base + k * stride] = v[k];
    }
}

// acceptance counters at the start of the next measured sweep: a bin adds
// only the accepts of its own sweeps (src/replica_run.c:293-295, 449-451)
__global__ void k_acc_mark(const unsigned long long *acc, const unsigned long long *att,
                           unsigned long long *acc_prev, unsigned long long *att_prev, int batch)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    acc_prev[b] = acc[b];
    att_prev[b] = att[b];
}

__global__ void k_d2f(const double *src, long long sS, float *dst, long long sD, long long len, int batch)
{
    const int b = blockIdx.y;
    if (b >= batch) return;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < len;
         idx += (long long)gridDim.x * blockDim.x)
        dst[b * sD + idx] = (float)src[b * sS + idx];
}

__global__ void k_f2d(const float *src, long long sS, double *dst, long long sD, long long len, int batch)
{
    const int b = blockIdx.y;
    if (b >= batch) return;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < len;
         idx += (long long)gridDim.x * blockDim.x)
        dst[b * sD + idx] = (double)src[b * sS + idx];
}

// out = expK diag(d) (the single-slice B matrix, src/green.c:126-137), FP32
// out = expK diag(d): the single-slice B matrix (src/green.c:126-137)
__global__ void k_scale_cols_d(const double *expK, const double *d, long long sdv, double *out, long long sO, int n,
                               int batch)
{
    const int b = blockIdx.y;
    if (b >= batch) return;
    const int nn = n * n, stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x + threadIdx.x; base < nn; base += 4 * stride) {
        double e[4], s[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int idx = min(base + k * stride, nn - 1);
            e[k] = expK[idx];
            s[k] = d[b * sdv + idx / n];
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (base + k * stride < nn) out[b * sO + base + k * stride] = e[k] * s[k];
    }
}

__global__ void k_scale_cols_f(const float *expK, const float *d, long long sdv, float *out,
                               long long sO, int n, int batch)
{
    const int b = blockIdx.y;
    if (b >= batch) return;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < (long long)n * n;
         idx += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(idx / n);
        out[b * sO + idx] = expK[idx] * d[b * sdv + j];
    }
}

// Rebuild operands of (1 + U D T)^-1 with D split as Db Ds (|Db| >= 1,
// |Ds| <= 1, src/linalg.c:963-984): since 1 + U D T = U Db N with
// N = Db^-1 U^T + Ds T, the Green's function is G = N^-1 X, X = Db^-1 U^T.
// Writes N, X and the sign of prod_{|D_i|>1} D_i.
// N = Db^-1 U^T + Ds T and X = Db^-1 U^T (Db = max(|D|, 1) with D's sign,
// Ds = D / Db) per 32 x 32 tile; U^T is read through an LDS tile so that
// both the U reads and the N, X, T accesses are coalesced.  Block (0, b)
// also records the sign of prod_{|D_i| > 1} D_i.
#define BN_T 32
// The big/small split of D_i (src/linalg.c:966-986, 1190-1215): without a
// log offset Db = D where |D| > 1 (else 1) and Ds = D / Db; with one, the
// true log-magnitude ell = log|D| + offset decides and the inverse of the
// big part is formed directly (sgn D e^-ell), so nothing overflows.  bi is
// the factor that divides by Db, s the small part; false when -|ell| is
// below log(DBL_TRUE_MIN) (src/linalg.c:166-182).
__device__ __forceinline__ bool d_split(double d, double off, double &bi, double &sm, bool &big)
{
    if (off == 0.0) {
        big = fabs(d) > 1.0;
        bi = big ? 1.0 / d : 1.0;
        sm = big ? 1.0 : d;
        return true;
    }
    const double ell = log(fabs(d)) + off;
    big = ell > 0.0;
    bi = big ? copysign(exp(-ell), d) : 1.0;
    sm = big ? 1.0 : copysign(exp(ell), d);
    return isfinite(ell) && -fabs(ell) >= -744.44007192138122; // log(4.9e-324)
}

__global__ __launch_bounds__(256) void k_build_N(const double *U, const double *D, const double *T, long long sU,
                                                 long long sD, const double *off, long long so, double *N, double *X,
                                                 int *bigsign, int n, int batch, int *status, int *reason)
{
    const int b = blockIdx.y;
    if (b >= batch || status[b]) return;
    const double ob = off ? off[b * so] : 0.0;
    const double *Ub = U + b * sU, *Tb = T + b * sU, *Db = D + b * sD;
    double *Nb = N + (long long)b * n * n, *Xb = X + (long long)b * n * n;
    const int nt = (n + BN_T - 1) / BN_T;
    const int i0 = (blockIdx.x % nt) * BN_T, j0 = (blockIdx.x / nt) * BN_T;
    const int tx = threadIdx.x % BN_T, ty = threadIdx.x / BN_T; // 32 x 8
    __shared__ double tl[BN_T][BN_T + 1];
    double u[BN_T / 8];
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) {
        const int i = min(i0 + ty + 8 * k, n - 1), j = min(j0 + tx, n - 1);
        u[k] = Ub[j + (long long)i * n]; // U(j, i) = U^T(i, j)
    }
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) tl[ty + 8 * k][tx] = u[k];
    __syncthreads();
    const int i = i0 + tx;
    const double d = Db[min(i, n - 1)];
    double dbi, dsm;
    bool big;
    (void)d_split(d, ob, dbi, dsm, big);
    double t[BN_T / 8];
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) t[k] = Tb[min(i, n - 1) + (long long)min(j0 + ty + 8 * k, n - 1) * n];
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) {
        const int j = j0 + ty + 8 * k;
        if (i < n && j < n) {
            const double uu = tl[tx][ty + 8 * k];
            const double x = big ? (ob == 0.0 ? uu / d : uu * dbi) : uu;
            Xb[i + (long long)j * n] = x;
            Nb[i + (long long)j * n] = x + dsm * t[k];
        }
    }
    if (blockIdx.x == 0) {
        __shared__ int neg, bad;
        if (threadIdx.x == 0) {
            neg = 0;
            bad = 0;
        }
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double bi, sm;
            bool bg;
            if (!d_split(Db[r], ob, bi, sm, bg)) bad = 1;
            if (bg && Db[r] < 0.0) atomicXor(&neg, 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            bigsign[b] = neg ? -1 : 1;
            if (bad) {
                status[b] = 1;
                reason[b] = DQMC_GPU_FAIL_EFFECTIVE_LOG_RANGE;
            }
        }
    }
}

// Db^-1 and Ds of the left and right D (Db = D where |D| > 1, else 1;
// Ds = D / Db) and the sign of prod_{|D_i| > 1} D_i over both, one
// workgroup per replica.
__global__ void k_split2(const double *Dl, long long sDl, const double *offl, long long sol, const double *Dr,
                         long long sDr, const double *offr, long long sor, int n, double *lbi, double *ls, double *rbi,
                         double *rs_, int *bigsign, int batch, int *status, int *reason)
{
    const int b = blockIdx.x;
    if (b >= batch || status[b]) return;
    const double ol = offl ? offl[b * sol] : 0.0, orr = offr ? offr[b * sor] : 0.0;
    __shared__ int neg, bad;
    if (threadIdx.x == 0) {
        neg = 0;
        bad = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double dl = Dl[b * sDl + i], dr = Dr[b * sDr + i];
        double li, lsv, ri, rsv;
        bool bl, br;
        const bool okl = d_split(dl, ol, li, lsv, bl), okr = d_split(dr, orr, ri, rsv, br);
        if (!okl || !okr) bad = 1;
        lbi[(long long)b * n + i] = li;
        ls[(long long)b * n + i] = lsv;
        rbi[(long long)b * n + i] = ri;
        rs_[(long long)b * n + i] = rsv;
        if ((bl && dl < 0.0) != (br && dr < 0.0)) atomicXor(&neg, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bigsign[b] = neg ? -1 : 1;
        if (bad) {
            status[b] = 1;
            reason[b] = DQMC_GPU_FAIL_EFFECTIVE_LOG_RANGE;
        }
    }
}

// udv_recenter (src/linalg.c:453-513) with the hard margin 8 of the
// centered updates, one workgroup per replica: c = centre of the log|D|
// range, Dout = D / e^c, offout = offin + c.  The output may be the input.
// A range whose half-width leaves less than the margin below log(DBL_MAX),
// or an entry that would not survive the rescaling, fails the replica
// (LINALG_FAILURE_CENTERED_RADIUS) and writes nothing.
__global__ __launch_bounds__(256) void k_udv_recenter(const double *Din, long long sDin, const double *oin,
                                                      long long soin, double *Dout, long long sDout, double *oout,
                                                      long long soout, int n, int batch, int *status, int *reason)
{
    const int b = blockIdx.x;
    if (b >= batch || status[b]) return;
    __shared__ double wlo[4], whi[4];
    __shared__ int bad;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) bad = 0;
    double d[2], ld[2];
    double lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int i = threadIdx.x + 256 * k;
        d[k] = i < n ? Din[b * sDin + i] : 1.0;
        ld[k] = log(fabs(d[k]));
        if (i < n) {
            lo = fmin(lo, ld[k]);
            hi = fmax(hi, ld[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (threadIdx.x + 256 * k < n && (!isfinite(d[k]) || d[k] == 0.0 || !isfinite(ld[k]))) bad = 1;
    lo = -wave_max(-lo);
    hi = wave_max(hi);
    if (lane == 0) {
        wlo[wave] = lo;
        whi[wave] = hi;
    }
    __syncthreads();
    lo = fmin(fmin(wlo[0], wlo[1]), fmin(wlo[2], wlo[3]));
    hi = fmax(fmax(whi[0], whi[1]), fmax(whi[2], whi[3]));
    const double off = oin ? oin[b * soin] : 0.0;
    const double c = lo + 0.5 * (hi - lo), r = 0.5 * (hi - lo);
    const double margin = 709.78271289338397 - r; // log(DBL_MAX) - r
    double v[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        v[k] = exp(ld[k] - c);
        if (threadIdx.x + 256 * k < n && (!isfinite(v[k]) || v[k] == 0.0)) bad = 1;
    }
    __syncthreads();
    const bool fail = bad || !isfinite(c) || !isfinite(r) || !isfinite(margin) || margin < 8.0 || !isfinite(off + c);
    if (fail) {
        if (threadIdx.x == 0) {
            status[b] = 1;
            reason[b] = DQMC_GPU_FAIL_CENTERED_RADIUS;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int i = threadIdx.x + 256 * k;
        if (i < n) Dout[b * sDout + i] = copysign(v[k], d[k]);
    }
    if (threadIdx.x == 0) oout[b * soout] = off + c;
}

// centered_update_has_margin (src/linalg.c:119-164), one workgroup per
// replica: a bound on the largest QR input entry of the update, which
// must stay below log(DBL_MAX) - 8; left, M = (op(B) U) D with U
// orthogonal: log|D_j| + max_i log sum_k |op(B)_ik|;
// right: max_i log|D_i| + log sum_k |T_ik| max_j |B_kj|.  Sums in the log
// domain shifted by their largest term.  A replica without the margin fails
// (LINALG_FAILURE_CENTERED_UPDATE_MARGIN).
__global__ __launch_bounds__(256) void k_update_margin(const double *B, long long sB, const double *T, long long sT,
                                                       const double *D, long long sD, int n, int left, int trans,
                                                       int batch, int *status, int *reason)
{
    const int b = blockIdx.x;
    if (b >= batch || status[b]) return;
    __shared__ double bmx[METRO_MAX_N];
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double *Bb = B + b * sB, *Tb = T ? T + b * sT : nullptr, *Db = D + b * sD;
    double worst = -INFINITY;
    if (left) {
        // row i of op(B): max_k |B_ik| then sum_k |B_ik| / max, rows over
        // threads (op(B) = B^T: column i of B)
        const long long si = trans ? n : 1, sk = trans ? 1 : n;
        double lb = -INFINITY;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            double mx = 0.0;
            for (int k = 0; k < n; k++) mx = fmax(mx, fabs(Bb[i * si + k * sk]));
            double sm = 0.0;
            if (mx > 0.0)
                for (int k = 0; k < n; k++) sm += fabs(Bb[i * si + k * sk]) / mx;
            if (mx > 0.0) lb = fmax(lb, log(mx) + log(sm));
        }
        lb = wave_max(lb);
        if (lane == 0) wmax[wave] = lb;
        __syncthreads();
        lb = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x) worst = fmax(worst, log(fabs(Db[j])) + lb);
    } else {
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            double mx = 0.0;
            for (int j = 0; j < n; j++) mx = fmax(mx, fabs(Bb[k + (long long)j * n]));
            bmx[k] = mx;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            double top = -INFINITY;
            for (int k = 0; k < n; k++) {
                const double t = fabs(Tb[i + (long long)k * n]);
                if (t != 0.0 && bmx[k] != 0.0) top = fmax(top, log(t) + log(bmx[k]));
            }
            double sm = 0.0;
            if (top > -INFINITY)
                for (int k = 0; k < n; k++) {
                    const double t = fabs(Tb[i + (long long)k * n]);
                    if (t != 0.0 && bmx[k] != 0.0) sm += exp(log(t) + log(bmx[k]) - top);
                }
            worst = fmax(worst, log(fabs(Db[i])) + top + log(sm));
        }
    }
    // NaN (a non-finite input) counts as no margin
    double w = isnan(worst) ? INFINITY : worst;
    w = wave_max(w);
    if (lane == 0) wmax[wave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double ww = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        if (!(isfinite(ww) && ww <= 709.78271289338397 - 8.0)) {
            status[b] = 1;
            reason[b] = DQMC_GPU_FAIL_CENTERED_UPDATE_MARGIN;
        }
    }
}

// M = Dlb^-1 Z^T Drb^-1 + Dls X Drs per 32 x 32 tile, Z read through an
// LDS tile (as in k_build_N)
__global__ __launch_bounds__(256) void k_build_M(const double *Z, const double *X, const double *lbi, const double *ls,
                                                 const double *rbi, const double *rs_, double *M, int n, int batch,
                                                 const int *status)
{
    const int b = blockIdx.y;
    if (b >= batch || status[b]) return;
    const long long nn = (long long)n * n;
    const double *Zb = Z + b * nn, *Xb = X + b * nn;
    double *Mb = M + b * nn;
    const int nt = (n + BN_T - 1) / BN_T;
    const int i0 = (blockIdx.x % nt) * BN_T, j0 = (blockIdx.x / nt) * BN_T;
    const int tx = threadIdx.x % BN_T, ty = threadIdx.x / BN_T;
    __shared__ double tl[BN_T][BN_T + 1];
    double z[BN_T / 8];
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) {
        const int i = min(i0 + ty + 8 * k, n - 1), j = min(j0 + tx, n - 1);
        z[k] = Zb[j + (long long)i * n]; // Z(j, i) = Z^T(i, j)
    }
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) tl[ty + 8 * k][tx] = z[k];
    __syncthreads();
    const int i = i0 + tx, ic = min(i, n - 1);
    const double li = lbi[(long long)b * n + ic], lsi = ls[(long long)b * n + ic];
    double x[BN_T / 8], rb[BN_T / 8], rsv[BN_T / 8];
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) {
        const int jc = min(j0 + ty + 8 * k, n - 1);
        x[k] = Xb[ic + (long long)jc * n];
        rb[k] = rbi[(long long)b * n + jc];
        rsv[k] = rs_[(long long)b * n + jc];
    }
#pragma unroll
    for (int k = 0; k < BN_T / 8; k++) {
        const int j = j0 + ty + 8 * k;
        if (i < n && j < n) Mb[i + (long long)j * n] = li * tl[tx][ty + 8 * k] * rb[k] + lsi * x[k] * rsv[k];
    }
}

// det sign of 1 + U D T = sgn(prod_big D) sgn det U sgn det M, with
// det M = det Q_M prod (D_M)_i (src/linalg.c:1039-1083); uq2, when given,
// is the sign of a second orthogonal factor (the two-sided form).
__global__ void k_sign_finish(const int *bigsign, const int *uq, long long suq, const int *uq2, long long suq2,
                              const int *mq, const double *DM, int n, int *det_sign, int *status, int *reason,
                              int batch)
{
    const int b = blockIdx.x;
    if (b >= batch || status[b]) return;
    __shared__ int neg, bad;
    if (threadIdx.x == 0) { neg = 0; bad = 0; }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = DM[(long long)b * n + i];
        if (d < 0.0) atomicXor(&neg, 1);
        if (!(d != 0.0) || !isfinite(d)) bad = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = bigsign[b] * uq[b * suq] * (uq2 ? uq2[b * suq2] : 1) * mq[b] * (neg ? -1 : 1);
        if (bad) s = 0;
        det_sign[b] = s;
        if (s == 0) {
            status[b] = 1;
            reason[b] = DQMC_GPU_FAIL_SIGN;
        }
    }
}

__global__ void k_mul_q(const int *a, long long sa, const int *b_, long long sb, int *out, long long so,
                        int batch, const int *status)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch || status[b]) return;
    out[b * so] = a[b * sa] * b_[b * sb];
}

// G <- Gnew after a rebuild (copy = 1), with the distance max|G - Gnew|
// (measure = 1; the wrapped G is the FP32 copy Gf when from_f32) and a
// non-finite check of Gnew and (check_old = 1) of the wrapped G; with
// copy = 0 only the distance is taken (Gnew then holds a from-scratch G,
// dqmc_record_stab_drift, src/dqmc.c:64-98).
// Many blocks per replica, results gathered with global atomics (the max
// of non-negative doubles is the max of their bit patterns) and finished
// by k_drift_finish.
__global__ void k_drift_copy(double *G, float *Gf, const double *Gnew, long long sG, int n, int batch, int measure,
                             int check_old, int from_f32, int copy, unsigned long long *tmax, int *tbad,
                             const int *status)
{
    const int b = blockIdx.y;
    if (b >= batch || status[b]) return;
    const long long nn = (long long)n * n, stride = (long long)gridDim.x * blockDim.x;
    double m = 0.0;
    int bad = 0;
    for (long long base = blockIdx.x * (long long)blockDim.x + threadIdx.x; base < nn; base += 4 * stride) {
        double v[4], o[4];
        // loads from clamped indices, issued together
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const long long idx = min(base + k * stride, nn - 1);
            v[k] = Gnew[b * sG + idx];
            o[k] = 0.0;
            if (measure || check_old) o[k] = from_f32 ? (double)Gf[b * sG + idx] : G[b * sG + idx];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const long long idx = base + k * stride;
            if (idx < nn) {
                if (!isfinite(v[k]) || (check_old && !isfinite(o[k]))) bad = 1;
                // a NaN on either side counts as an infinite distance
                if (measure) m = fmax(m, isnan(o[k] - v[k]) ? INFINITY : fabs(o[k] - v[k]));
                if (copy) {
                    G[b * sG + idx] = v[k];
                    if (Gf) Gf[b * sG + idx] = (float)v[k];
                }
            }
        }
    }
    if (measure) {
        // one atomic per workgroup: wave maxima meet in LDS first
        __shared__ double wm[16];
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, wm[w]);
            if (m > 0.0) atomicMax(&tmax[b], (unsigned long long)__double_as_longlong(m));
        }
    }
    if (bad) atomicOr(&tbad[b], 1);
}

// Per replica, after k_drift_copy: failures are recorded (a non-finite
// Gnew; with guard > 0 a distance above guard, DQMC_GPU_FAIL_DRIFT), the
// running sign is reset to the exact +1 when sign is given (a rebuild, PH
// path, src/dqmc.c:740-742, 799-801), the distance is accumulated as a
// drift sample when record is set, and the gather slots are cleared for the
// next call.  The slice tau of the sample and the model's completed-sweep
// counter are kept with each new maximum (max_tau, max_sweep of
// DqmcStabDrift, src/dqmc.c:90-95).
__global__ void k_drift_finish(int batch, int record, double guard, unsigned long long *tmax, int *tbad,
                               double *dmax, double *dsum, unsigned long long *dcnt, int *status, int *reason,
                               double *sign, int tau, const unsigned long long *sweeps, int *dtau,
                               unsigned long long *dsweep)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    if (!status[b]) {
        const double mm = __longlong_as_double((long long)tmax[b]);
        if (sign) sign[b] = 1.0;
        if (tbad[b]) {
            status[b] = 1;
            reason[b] = DQMC_GPU_FAIL_NONFINITE;
        } else if (guard > 0.0 && !(mm <= guard)) {
            status[b] = 1;
            reason[b] = DQMC_GPU_FAIL_DRIFT;
        }
        if (record && !tbad[b]) {
            if (dcnt[b] == 0 || mm > dmax[b]) {
                dmax[b] = mm;
                dtau[b] = tau;
                dsweep[b] = *sweeps;
            }
            dsum[b] += mm;
            dcnt[b] += 1;
        }
    }
    tmax[b] = 0;
    tbad[b] = 0;
}

// one completed sweep of a model (the counter behind max_sweep)
__global__ void k_sweep_tick(unsigned long long *sweeps)
{
    if (threadIdx.x == 0) *sweeps += 1;
}

// Column pre-pivoting of a UDV update's QR input M: perm = the columns of
// M in order of decreasing 2-norm (ties by index), so that the QR of
// M P = Q R factors the graded columns largest first.  Norms are compared
// as log(max|x|) + log ||x / max|x|||, so entries near the double range
// neither overflow nor underflow.  k_col_keys: one wave per column (lanes
// over rows, the column held in registers for both passes), many
// workgroups per replica; k_col_sort: one workgroup per replica, a bitonic
// sort of the (key, index) pairs in LDS.
#define CP_MAX 512
__global__ __launch_bounds__(256) void k_col_keys(const double *M, long long sM, int n, double *keys, int batch,
                                                  const int *status)
{
    const int b = blockIdx.y;
    if (b >= batch || status[b]) return;
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (j >= n) return; // wave-uniform
    const double *col = M + b * sM + (long long)j * n;
    double v[CP_MAX / 64];
#pragma unroll
    for (int k = 0; k < CP_MAX / 64; k++) {
        const int i = lane + 64 * k;
        v[k] = i < n ? col[i] : 0.0;
    }
    double mx = 0.0;
#pragma unroll
    for (int k = 0; k < CP_MAX / 64; k++) mx = fmax(mx, fabs(v[k]));
    mx = wave_max(mx);
    double ss = 0.0;
    if (mx > 0.0) {
        const double r = 1.0 / mx;
#pragma unroll
        for (int k = 0; k < CP_MAX / 64; k++) ss += (v[k] * r) * (v[k] * r);
    }
    ss = wave_sum(ss);
    if (lane == 0) {
        const double key = mx > 0.0 ? log(mx) + 0.5 * log(ss) : -INFINITY;
        keys[(long long)b * n + j] = isnan(key) ? -INFINITY : key;
    }
}

__global__ __launch_bounds__(256) void k_col_sort(const double *keys, int n, int *perm, int batch, const int *status)
{
    const int b = blockIdx.x;
    if (b >= batch || status[b]) return;
    __shared__ double key[CP_MAX];
    __shared__ int idx[CP_MAX];
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int j = threadIdx.x; j < n2; j += blockDim.x) {
        key[j] = j < n ? keys[(long long)b * n + j] : -INFINITY;
        idx[j] = j;
    }
    __syncthreads();
    // bitonic sort, descending key, ascending index among equal keys
    for (int k = 2; k <= n2; k <<= 1)
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int l = i ^ jj;
                if (l > i) {
                    const bool before = key[i] > key[l] || (key[i] == key[l] && idx[i] < idx[l]);
                    const bool up = (i & k) == 0; // this half ends up descending
                    if (up != before) {
                        const double tk = key[i];
                        key[i] = key[l];
                        key[l] = tk;
                        const int ti = idx[i];
                        idx[i] = idx[l];
                        idx[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    for (int k = threadIdx.x; k < n; k += blockDim.x) perm[(long long)b * n + k] = idx[k];
}

// out[:, k] = in[:, perm[k]] (rows = 0) or out[k, :] = in[perm[k], :]
// (rows = 1), n x n per batch entry (n * n < 2^31).  For even n a thread
// moves two consecutive entries of one column: one 16-byte store, and for
// a column gather one 16-byte load.
__global__ void k_gather(const double *in, long long sIn, double *out, long long sOut, const int *perm, int n,
                         int rows, int batch, const int *status)
{
    const int b = blockIdx.y;
    if (b >= batch || status[b]) return;
    const int nn = n * n;
    const int *pb = perm + (long long)b * n;
    const double *src = in + b * sIn;
    double *dst = out + b * sOut;
    const int step = gridDim.x * blockDim.x;
    if ((n & 1) == 0) {
        for (int e = 2 * (blockIdx.x * blockDim.x + threadIdx.x); e < nn; e += 2 * step) {
            const int k = e / n, i = e - k * n;
            double2 v;
            if (rows) {
                v.x = src[pb[i] + (long long)k * n];
                v.y = src[pb[i + 1] + (long long)k * n];
            } else {
                v = *reinterpret_cast<const double2 *>(src + i + (long long)pb[k] * n);
            }
            *reinterpret_cast<double2 *>(dst + e) = v;
        }
    } else {
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += step) {
            const int k = e / n, i = e - k * n;
            dst[e] = src[rows ? pb[i] + (long long)k * n : i + (long long)pb[k] * n];
        }
    }
}

// ---------------------------------------------------------------- buffers

template <typename T> static int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return 0;
    HIPCHK(hipMalloc((void **)p, count * sizeof(T)));
    return 0;
}

struct Udv {
    double *U, *D, *T;
    int *q;
    long long sU, sD, sq; // per-replica strides
    double *off = nullptr; // log offset of a centered UDV per replica (null: 0)
    long long so = 0;
};

struct Work {
    int n = 0, nb = 0;
    hipStream_t s = nullptr;
    double *W1 = nullptr, *W2 = nullptr, *W3 = nullptr, *W4 = nullptr;
    double *V1 = nullptr, *V2 = nullptr, *V3 = nullptr, *V4 = nullptr, *V5 = nullptr, *Tws = nullptr;
    int *I1 = nullptr, *I2 = nullptr, *status = nullptr, *reason = nullptr, *det_sign = nullptr;
    int *perm = nullptr; // column order of a pre-pivoted QR, n per entry
    bool pivot = false;  // UDV updates with column pre-pivoting (qr_update)
    float *F1 = nullptr, *F2 = nullptr;
    std::vector<void *> owned;

    int alloc(int n_, int nb_, int mixed)
    {
        n = n_;
        nb = nb_;
        const size_t nn = (size_t)n * n * nb;
        if (dalloc(&W1, nn) || dalloc(&W2, nn) || dalloc(&W3, nn) || dalloc(&W4, nn)) return 1;
        if (dalloc(&V1, (size_t)n * nb) || dalloc(&V2, (size_t)n * nb) || dalloc(&V3, (size_t)n * nb) ||
            dalloc(&V4, (size_t)n * nb) || dalloc(&V5, (size_t)n * nb))
            return 1;
        const int npan = (n + QR_NB - 1) / QR_NB;
        if (dalloc(&Tws, (size_t)npan * QR_NB * QR_NB * nb)) return 1;
        if (dalloc(&I1, nb) || dalloc(&I2, nb) || dalloc(&status, nb) || dalloc(&reason, nb) ||
            dalloc(&det_sign, nb) || dalloc(&perm, (size_t)n * nb))
            return 1;
        if (mixed && (dalloc(&F1, nn) || dalloc(&F2, nn))) return 1;
        void *ps[] = {W1, W2, W3, W4, V1, V2, V3, V4, V5, Tws, I1, I2, status, reason, det_sign, perm, F1, F2};
        for (void *p : ps)
            if (p) owned.push_back(p);
        HIPCHK(hipMemset(status, 0, sizeof(int) * nb));
        HIPCHK(hipMemset(reason, 0, sizeof(int) * nb));
        HIPCHK(hipMemset(det_sign, 0, sizeof(int) * nb));
        return 0;
    }
    void release()
    {
        for (void *p : owned) (void)hipFree(p);
        owned.clear();
    }
};

static dim3 grid2(long long len, int batch)
{
    long long bx = (len + 255) / 256;
    if (bx > 64) bx = 64;
    if (bx < 1) bx = 1;
    return dim3((unsigned)bx, (unsigned)batch);
}

// --------------------------------------------------------------- profiling
// With g_prof set (dqmc_gpu_profile_sweep), every stage is bracketed by HIP
// events on the stream it runs on; classes are DQMC_GPU_PROF_*.
struct ProfRec {
    int cls;
    hipEvent_t a, b;
};
static thread_local std::vector<ProfRec> *g_prof = nullptr;
struct ProfScope {
    ProfRec r;
    hipStream_t s = nullptr;
    bool on;
    ProfScope(int cls, hipStream_t st, bool enable = true) : on(enable && g_prof != nullptr)
    {
        if (!on) return;
        r.cls = cls;
        s = st;
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        (void)hipEventRecord(r.a, s);
    }
    ~ProfScope()
    {
        if (!on) return;
        (void)hipEventRecord(r.b, s);
        g_prof->push_back(r);
    }
};

// ------------------------------------------------------- linear algebra ops

static int launch_check(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_err("launch of %s failed: %s", what, hipGetErrorString(e));
        return 1;
    }
    return 0;
}

// QR of the n x n matrices A (stride n*n, destroyed): Q, Rn = D^-1 R, D and
// sgn det Q with the given per-replica strides.
static int qr(Work &w, double *A, double *Q, long long sQ, double *Rn, long long sR, double *D,
              long long sD, int *qsign, long long sq)
{
    ProfScope ps_(DQMC_GPU_PROF_QR, w.s);
    QrArgs a;
    a.n = w.n; a.batch = w.nb;
    a.A = A; a.sA = (long long)w.n * w.n;
    a.Q = Q; a.sQ = sQ;
    a.Rn = Rn; a.sR = sR;
    a.D = D; a.sD = sD;
    a.Tws = w.Tws; a.sT = (long long)((w.n + QR_NB - 1) / QR_NB) * QR_NB * QR_NB;
    a.qsign = qsign; a.sq = sq;
    a.status = w.status; a.reason = w.reason;
    if (w.n <= 64)
        hipLaunchKernelGGL(k_qr_udv<1>, dim3(w.nb), dim3(QR_THREADS), qr_lds_bytes(w.n), w.s, a);
    else if (w.n <= 128)
        hipLaunchKernelGGL(k_qr_udv<2>, dim3(w.nb), dim3(QR_THREADS), qr_lds_bytes(w.n), w.s, a);
    else if (w.n <= 256)
        hipLaunchKernelGGL(k_qr_udv<4>, dim3(w.nb), dim3(QR_THREADS), qr_lds_bytes(w.n), w.s, a);
    else
        hipLaunchKernelGGL(k_qr_udv<8>, dim3(w.nb), dim3(QR_THREADS), qr_lds_bytes(w.n), w.s, a);
    if (launch_check("k_qr_udv")) return 1;
    if (!Q) return 0;
    const dim3 fg(qr_formq_blocks(w.n, w.nb)), fb(FQ_WAVES * 64);
    const size_t fl = qr_formq_lds_bytes(w.n);
    if (w.n <= 64) hipLaunchKernelGGL(k_qr_formq_reg<16>, fg, fb, fl, w.s, a);
    else if (w.n <= 128) hipLaunchKernelGGL(k_qr_formq_reg<32>, fg, fb, fl, w.s, a);
    else if (w.n <= 256) hipLaunchKernelGGL(k_qr_formq_reg<64>, fg, fb, fl, w.s, a);
    else hipLaunchKernelGGL(k_qr_formq, fg, fb, fl, w.s, a);
    return launch_check("k_qr_formq");
}

// X = U^-1 B, U unit upper triangular (dtrsm, left, upper, unit diagonal)
static int trsm(Work &w, const double *U, long long sU, const double *B, long long sB, double *X, long long sX,
                int trans = 0, const double *rs = nullptr, long long srs = 0)
{
    ProfScope ps_(DQMC_GPU_PROF_TRSM, w.s);
    TrsmArgs a;
    a.n = w.n; a.batch = w.nb;
    a.U = U; a.sU = sU;
    a.B = B; a.sB = sB;
    a.X = X; a.sX = sX;
    a.status = w.status;
    a.trans = trans;
    a.rs = rs; a.srs = srs;
    hipLaunchKernelGGL(k_trsm_unit_upper, dim3(trsm_blocks(w.n, w.nb)), dim3(256), trsm_lds_bytes(w.n), w.s, a);
    return launch_check("k_trsm_unit_upper");
}

template <typename T> static int gemm(Work &w, GemmArgs<T> g)
{
    g.status = w.status;
    const bool full = !g.Kb && g.M == w.n && g.N == w.n && g.K == w.n;
    ProfScope ps_(g.Kb ? DQMC_GPU_PROF_FLUSH : DQMC_GPU_PROF_GEMM, w.s, g.Kb || full);
    gemm_launch(g, w.s);
    return launch_check("k_gemm");
}

static int copy_nn(Work &w, const double *src, long long sS, double *dst, long long sD, long long len)
{
    hipLaunchKernelGGL(k_copy_d, grid2(len, w.nb), dim3(256), 0, w.s, src, sS, dst, sD, len, w.nb);
    return launch_check("k_copy_d");
}

// Column pre-pivoting of the UDV updates (a model's DQMC_GPU_PIVOT=0 turns it off):
// the QR input M of an update is factored as M P = Q R with P ordering its
// columns by decreasing norm, so U' = Q, D' = diag R and T' = (D'^-1 R) P^T T.
// The normalised R then has entries of order one and the accumulated T
// stays bounded (at 16x16, U=4, beta=8 max|T| ~ 3e2 against 3e5 unpivoted),
// which is what keeps the boundary rebuild's T_l T_r^T product accurate.

// QR of an update's input M (n x n per entry, left unchanged when pivoted)
// into U' (Q), Rn (W3), D' and sgn det Q; with pivoting the columns are
// first ordered into W4 (w.perm holds the order) and factored from there.
// T' = Rn P^T T is left to update_T.
static int qr_update(Work &w, double *M, const Udv &out)
{
    const long long nn = (long long)w.n * w.n;
    double *qin = M;
    if (w.pivot) {
        // the keys go to V5 (n per entry), free during an update
        hipLaunchKernelGGL(k_col_keys, dim3((w.n + 3) / 4, w.nb), dim3(256), 0, w.s, M, nn, w.n, w.V5, w.nb,
                           w.status);
        if (launch_check("k_col_keys")) return 1;
        hipLaunchKernelGGL(k_col_sort, dim3(w.nb), dim3(256), 0, w.s, w.V5, w.n, w.perm, w.nb, w.status);
        if (launch_check("k_col_sort")) return 1;
        hipLaunchKernelGGL(k_gather, grid2(nn, w.nb), dim3(256), 0, w.s, M, nn, w.W4, nn, w.perm, w.n, 0, w.nb,
                           w.status);
        if (launch_check("k_gather")) return 1;
        qin = w.W4;
    }
    return qr(w, qin, out.U, out.sU, w.W3, nn, out.D, out.sD, out.q, out.sq);
}

// out.T = Rn P^T in.T after qr_update (Rn in W3); W2 and W4 are scratch
static int update_T(Work &w, const Udv &in, const Udv &out)
{
    const int n = w.n;
    const long long nn = (long long)n * n;
    if (w.pivot) {
        hipLaunchKernelGGL(k_gather, grid2(nn, w.nb), dim3(256), 0, w.s, in.T, in.sU, w.W2, nn, w.perm, n, 1, w.nb,
                           w.status);
        if (launch_check("k_gather")) return 1;
        return gemm(w, gemm_args<double>(n, n, n, w.nb, w.W3, n, nn, 0, w.W2, n, nn, 0, out.T, n, out.sU));
    }
    const bool inplace = in.T == out.T;
    GemmArgs<double> g = gemm_args<double>(n, n, n, w.nb, w.W3, n, nn, 0, in.T, n, in.sU, 0,
                                           inplace ? w.W4 : out.T, n, inplace ? nn : out.sU);
    if (gemm(w, g)) return 1;
    return inplace ? copy_nn(w, w.W4, nn, out.T, out.sU, nn) : 0;
}

// udv_combine (src/linalg.c:796-909): C = Dl (Tl Ur) Dr; QR(C) = Q R;
// out = (Ul Q, diag R, (D^-1 R) Tr).  rT: the right factor is stored
// transposed (U D T = R^T, the suffix stack), so R = Tr^T Dr Ur^T,
// C = Dl (Tl Tr^T) Dr and out T = (D^-1 R) Ur^T, which is no longer
// triangular; udv_inv_one_plus does not need it to be.
static int udv_combine(Work &w, const Udv &l, const Udv &r, const Udv &o, int rT = 0)
{
    ProfScope ps_(DQMC_GPU_PROF_COMBINE, w.s);
    const int n = w.n;
    const long long nn = (long long)n * n;
    GemmArgs<double> g = rT ? gemm_args<double>(n, n, n, w.nb, l.T, n, l.sU, 0, r.T, n, r.sU, 1, w.W1, n, nn)
                            : gemm_args<double>(n, n, n, w.nb, l.T, n, l.sU, 0, r.U, n, r.sU, 0, w.W1, n, nn);
    g.ar = vec(l.D, l.sD);
    g.bc = vec(r.D, r.sD);
    if (gemm(w, g)) return 1;
    if (qr(w, w.W1, w.W2, nn, w.W3, nn, o.D, o.sD, w.I1, 1)) return 1;
    g = gemm_args<double>(n, n, n, w.nb, l.U, n, l.sU, 0, w.W2, n, nn, 0, o.U, n, o.sU);
    if (gemm(w, g)) return 1;
    g = rT ? gemm_args<double>(n, n, n, w.nb, w.W3, n, nn, 0, r.U, n, r.sU, 1, o.T, n, o.sU)
           : gemm_args<double>(n, n, n, w.nb, w.W3, n, nn, 0, r.T, n, r.sU, 0, o.T, n, o.sU);
    if (gemm(w, g)) return 1;
    hipLaunchKernelGGL(k_mul_q, dim3((w.nb + 255) / 256), dim3(256), 0, w.s, l.q, l.sq, w.I1, 1LL, o.q,
                       o.sq, w.nb, w.status);
    return launch_check("k_mul_q");
}

// udv_inv_one_plus_work (src/linalg.c:911-1115) in the form
//   G = (Db^-1 U^T + Ds T)^-1 Db^-1 U^T = Rn_N^-1 D_N^-1 Q_N^T X,
// N = Q_N D_N Rn_N by Householder QR, so no T^-1 is ever formed and Rn_N^-1
// is applied by a triangular solve; the sign is
// sgn det(1 + U D T) = sgn det U * prod_{|D_i|>1} sgn D_i * sgn det N with
// det N = det Q_N prod (D_N)_i (src/linalg.c:1039-1083 computes the same
// quantity through the LU of the reference's M).
static int udv_inv_one_plus(Work &w, const Udv &u, double *Gout, long long sG)
{
    ProfScope ps_(DQMC_GPU_PROF_INV, w.s);
    const int n = w.n;
    const long long nn = (long long)n * n;
    const int bnt = (n + BN_T - 1) / BN_T;
    hipLaunchKernelGGL(k_build_N, dim3(bnt * bnt, w.nb), dim3(256), 0, w.s, u.U, u.D, u.T, u.sU, u.sD, u.off, u.so,
                       w.W1, w.W4, w.I1, n, w.nb, w.status, w.reason);
    if (launch_check("k_build_N")) return 1;                         // W1 = N, W4 = X
    if (qr(w, w.W1, w.W2, nn, w.W3, nn, w.V3, n, w.I2, 1)) return 1; // W2 = Q_N, V3 = D_N, W3 = Rn_N
    GemmArgs<double> g = gemm_args<double>(n, n, n, w.nb, w.W2, n, nn, 1, w.W4, n, nn, 0, w.W1, n, nn);
    g.cr = vec<double>(w.V3, n, 1);
    if (gemm(w, g)) return 1;                           // W1 = D_N^-1 Q_N^T X
    if (trsm(w, w.W3, nn, w.W1, nn, Gout, sG)) return 1; // G = Rn_N^-1 W1
    hipLaunchKernelGGL(k_sign_finish, dim3(w.nb), dim3(256), 0, w.s, w.I1, u.q, u.sq, (const int *)nullptr, 0LL,
                       w.I2, w.V3, n, w.det_sign, w.status, w.reason, w.nb);
    return launch_check("k_sign_finish");
}

// G = (1 + L R)^-1 at a block boundary with one QR, the fused form of
// udv_combine + udv_inv_one_plus (src/linalg.c:796-1115).  L = Ul Dl Tl
// (prefix) and R stored transposed (suffix stack: Ur Dr Tr = R^T, so
// R = Tr^T Dr Ur^T).  With the Db/Ds split of both diagonals
//   1 + L R = Ul Dlb M Drb Ur^T,
//   M = Dlb^-1 (Ul^T Ur) Drb^-1 + Dls (Tl Tr^T) Drs
// (both terms bounded: the big parts are divided out), so with
// M = Q_M D_M Rn_M
//   G = Ur Drb^-1 Rn_M^-1 D_M^-1 Q_M^T Dlb^-1 Ul^T:
// two GEMMs for the terms of M, one for Q_M^T Ul^T, one triangular solve
// and one GEMM for Ur.  sgn det(1 + L R) = sgn det Ul * sgn det Ur *
// prod_{|D|>1} sgn D (both sides) * sgn det Q_M * prod sgn (D_M)_i.
static int udv_inv_two_sided(Work &w, const Udv &l, const Udv &r, double *Gout, long long sG)
{
    ProfScope ps_(DQMC_GPU_PROF_INV, w.s);
    const int n = w.n;
    const long long nn = (long long)n * n;
    hipLaunchKernelGGL(k_split2, dim3(w.nb), dim3(256), 0, w.s, l.D, l.sD, l.off, l.so, r.D, r.sD, r.off, r.so, n,
                       w.V1, w.V2, w.V4, w.V5, w.I1, w.nb, w.status, w.reason);
    if (launch_check("k_split2")) return 1;
    // W1 = Ur^T Ul (k_build_M reads it transposed), W2 = Tl Tr^T
    GemmArgs<double> g = gemm_args<double>(n, n, n, w.nb, r.U, n, r.sU, 1, l.U, n, l.sU, 0, w.W1, n, nn);
    if (gemm(w, g)) return 1;
    g = gemm_args<double>(n, n, n, w.nb, l.T, n, l.sU, 0, r.T, n, r.sU, 1, w.W2, n, nn);
    if (gemm(w, g)) return 1;
    const int bnt = (n + BN_T - 1) / BN_T;
    hipLaunchKernelGGL(k_build_M, dim3(bnt * bnt, w.nb), dim3(256), 0, w.s, w.W1, w.W2, w.V1, w.V2, w.V4, w.V5, w.W4,
                       n, w.nb, w.status);
    if (launch_check("k_build_M")) return 1;
    if (qr(w, w.W4, w.W1, nn, w.W2, nn, w.V3, n, w.I2, 1)) return 1; // W1 = Q_M, W2 = Rn_M, V3 = D_M
    g = gemm_args<double>(n, n, n, w.nb, w.W1, n, nn, 1, l.U, n, l.sU, 1, w.W3, n, nn);
    g.br = vec<double>(w.V1, n);    // Dlb^-1 on the rows of Ul^T
    g.cr = vec<double>(w.V3, n, 1); // D_M^-1
    if (gemm(w, g)) return 1;                            // W3 = D_M^-1 Q_M^T Dlb^-1 Ul^T
    if (trsm(w, w.W2, nn, w.W3, nn, w.W4, nn)) return 1; // W4 = Rn_M^-1 W3
    g = gemm_args<double>(n, n, n, w.nb, r.U, n, r.sU, 0, w.W4, n, nn, 0, Gout, n, sG);
    g.ac = vec<double>(w.V4, n); // Drb^-1
    if (gemm(w, g)) return 1;    // G = Ur Drb^-1 W4
    hipLaunchKernelGGL(k_sign_finish, dim3(w.nb), dim3(256), 0, w.s, w.I1, l.q, l.sq, r.q, r.sq, w.I2, w.V3, n,
                       w.det_sign, w.status, w.reason, w.nb);
    return launch_check("k_sign_finish");
}

// ------------------------------------------------------------------ model

struct Model {
    int n = 0, L = 0, stab = 0, M = 0, W = 0, Wf = 0, nrep = 0, first = 0, cap = 0;
    int msplit = 1; // Metropolis launches (site segments, each with its flush) per slice
    bool tuned_separable = false; // B-matrix products on the separable path after tuning
    bool pivot = true; // column pre-pivoted UDV updates (DQMC_GPU_PIVOT=0: off)
    bool fuse = true;  // last segment's flush inside the wrap (DQMC_GPU_FUSE=0: off)
    int sub = 0;       // FP64 Metropolis segments on the submatrix path (0: windowed path)
    bool flush64 = true; // rank-P flush on k_flush64 where n % 64 == 0 (DQMC_GPU_FLUSH64=0: the GEMM)
    bool two_sided = true; // boundary rebuilds by udv_inv_two_sided (one QR) instead of combine + inv_one_plus
    bool persist_ok = true; // may run on the persistent sweep kernel (DQMC_GPU_PERSIST=0: never)
    bool in_pset = false;   // runs on it in the current launch set
    int Lx = 1, Ly = 1, nq = 0, nbin = 0;
    double U = 0, dtau = 0, mu = 0, lambda = 0;
    double expv[2], expvinv[2], nu[2], nd[2];
    std::vector<double> h_expK, h_expKinv, h_thop, h_phase;
    // expK = E_y (x) E_x and expKinv = F_y (x) F_x on a 16 x 16 lattice:
    // factor pairs (E_x, E_y), (E_x^T, E_y^T), (F_x, F_y), (F_x^T, F_y^T),
    // 16 x 16 row-major each; empty when expK does not split
    std::vector<double> h_kron;
    double *kronF = nullptr;
    std::vector<int> h_bipart;
    std::vector<unsigned long long> seeds;
    std::vector<signed char> h_fields;
    std::vector<uint64_t> h_rng;

    double *expK = nullptr, *expKinv = nullptr, *thop = nullptr, *phase = nullptr;
    float *expKf = nullptr, *expKinvf = nullptr;
    int *bipart = nullptr;
    double *G = nullptr, *Gnew = nullptr;
    float *Gf = nullptr;
    signed char *field = nullptr;
    double *dvec = nullptr, *dinv = nullptr;
    float *dvecf = nullptr, *dinvf = nullptr;
    uint64_t *rng = nullptr;
    double *SU = nullptr, *SD = nullptr, *ST = nullptr;
    int *Sq = nullptr;
    // log offsets of the centered rebuild mode (green_rebuild=centered):
    // suffix stack, prefix stack, left and combined UDVs, all 0 otherwise
    double *SO = nullptr, *AO = nullptr, *LO = nullptr, *CO = nullptr;
    bool centered = false;
    double *LU_ = nullptr, *LD = nullptr, *LT = nullptr;
    int *Lq = nullptr;
    double *CU = nullptr, *CD = nullptr, *CT = nullptr;
    int *Cq = nullptr;
    // from-scratch UDV of the drift samples (DQMC_GPU_DRIFT_SCRATCH only)
    double *XU = nullptr, *XD = nullptr, *XT = nullptr, *XO = nullptr;
    int *Xq = nullptr;
    unsigned long long *acc = nullptr, *att = nullptr, *acc_prev = nullptr, *att_prev = nullptr;
    int *pcount = nullptr; // pending delayed-update pairs per replica after a slice
    bool rlog_on = false;  // DQMC_GPU_RATIO_LOG=1: keep each proposal's ratio
    double *rlog = nullptr; // ratio R_up R_dn of every site of the latest slice, n per replica
    unsigned long long *dtmax = nullptr; // drift gather slots (k_drift_copy -> k_drift_finish)
    int *dtbad = nullptr;
    // block products Bblock_j = B_{(j+1)s-1} ... B_{js} (green_build_Bblock,
    // src/green.c:151-166), formed by the left chain after block j's
    // Metropolis and read by the next sweep's suffix stack: M x R x n^2
    double *BB = nullptr;
    bool bb_valid = false;
    double *sign = nullptr;
    double *dmax = nullptr, *dsum = nullptr;
    unsigned long long *dcnt = nullptr;
    int *dtaumax = nullptr;                 // tau of each replica's largest drift sample
    unsigned long long *dsweepmax = nullptr; // ... and the sweep it was taken in
    unsigned long long *sweeps = nullptr;    // completed sweeps of the model (device counter)
    double *bins = nullptr, *bszz = nullptr, *bsperp = nullptr;
    int *counts = nullptr;
    // sweep_order=alternating (src/dqmc.c:815-963): direction of the next
    // sweep, which carried stack is current, and the prefix stack
    // A[j] = B_{js-1} ... B_0 (UDV, j = 0..M) of the last forward sweep
    bool alternating = false;
    int dir = 0; // 0 forward, 1 backward
    bool prefix_ok = false, suffix_ok = false;
    double *AU = nullptr, *AD = nullptr, *AT = nullptr;
    int *Aq = nullptr;
    Work w;
    // the suffix stack of a forward sweep is built on w2's stream while
    // the main stream runs block 0 (fork at the sweep start, join before
    // the first boundary rebuild)
    Work w2;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool allocated = false;
    // one captured graph per sweep variant (SW_*)
    hipGraph_t graph[5] = {};
    hipGraphExec_t gexec[5] = {};
    std::vector<void *> owned;

    Udv left() const
    {
        return Udv{LU_, LD, LT, Lq, (long long)n * n, (long long)n, 1, LO, 1};
    }
    Udv xdrift() const
    {
        return Udv{XU, XD, XT, Xq, (long long)n * n, (long long)n, 1, XO, 1};
    }
    Udv comb() const
    {
        return Udv{CU, CD, CT, Cq, (long long)n * n, (long long)n, 1, CO, 1};
    }
    Udv stack(int j) const
    {
        const long long nn = (long long)n * n;
        return Udv{SU + j * nn, SD + (long long)j * n, ST + j * nn, Sq + j, (M + 1) * nn,
                   (long long)(M + 1) * n, (long long)(M + 1), SO ? SO + j : nullptr, (long long)(M + 1)};
    }
    Udv alt(int j) const
    {
        const long long nn = (long long)n * n;
        return Udv{AU + j * nn, AD + (long long)j * n, AT + j * nn, Aq + j, (M + 1) * nn,
                   (long long)(M + 1) * n, (long long)(M + 1), AO ? AO + j : nullptr, (long long)(M + 1)};
    }
};

struct dqmc_gpu_ctx {
    int device = 0;
    dqmc_gpu_opts opts;
    std::vector<Model *> models;
    // the models swept by k_sweep_persist, in launch order (most work
    // first), their descriptors on the device and the launch's workgroups
    std::vector<Model *> pset;
    PersistModel *d_pm = nullptr;
    size_t d_pm_cap = 0;
    int pblocks = 0;
    hipEvent_t pev = nullptr;
    unsigned long long *ptimers = nullptr; // phase timers of profiled persistent launches

    int nrep_total = 0;
    bool initialised = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

static int model_for_replica(dqmc_gpu_ctx *ctx, int replica, Model **m, int *local)
{
    for (Model *md : ctx->models) {
        if (replica >= md->first && replica < md->first + md->nrep) {
            *m = md;
            *local = replica - md->first;
            return 0;
        }
    }
    set_err("replica %d out of range", replica);
    return 1;
}

// Delay depth (window) for the Metropolis kernel: the largest of 32, 16, 8
// not above the request whose two W x n pending-update tiles fit in LDS.
// window of the Metropolis kernel: the largest of 32 / 16 / 8 sites not
// above the requested delay whose LDS fits (FP64 at most 16: the 32-site
// FP64 window does not fit the register file)
static int metro_window(int n, int want, size_t elem)
{
    const size_t budget = 150 * 1024;
    int W = want > 0 ? want : 32;
    if (elem == sizeof(double)) W = std::min(W, 16);
    const int opts[] = {32, 16, 8};
    for (int o : opts)
        if (o <= W && (elem == sizeof(double) ? metro_lds_bytes<double>(n, o) : metro_lds_bytes<float>(n, o)) <= budget)
            return o;
    return 8;
}

// prefix stack of the alternating mode; A[0] is the identity and never
// changes (green_stack_build_prefix and the forward leg store I there,
// src/green.c:231, src/dqmc.c:652-654)
static int alt_alloc(Model &m)
{
    if (m.AU) return 0;
    const size_t nn = (size_t)m.n * m.n, S = (size_t)(m.M + 1), R = (size_t)m.nrep;
    if (dalloc(&m.AU, nn * S * R) || dalloc(&m.AD, (size_t)m.n * S * R) || dalloc(&m.AT, nn * S * R) ||
        dalloc(&m.Aq, S * R) || dalloc(&m.AO, S * R))
        return 1;
    m.owned.push_back(m.AU); m.owned.push_back(m.AD); m.owned.push_back(m.AT); m.owned.push_back(m.Aq);
    m.owned.push_back(m.AO);
    HIPCHK(hipMemset(m.AO, 0, sizeof(double) * S * R));
    const Udv a0 = m.alt(0);
    hipLaunchKernelGGL(k_udv_identity, grid2((long long)nn, m.nrep), dim3(256), 0, m.w.s, a0.U, a0.sU, a0.D, a0.sD,
                       a0.T, a0.sU, a0.q, a0.sq, m.n, m.nrep, a0.off, a0.so);
    return launch_check("k_udv_identity");
}

static int model_alloc(dqmc_gpu_ctx *ctx, Model &m)
{
    const int n = m.n, R = m.nrep, L = m.L;
    const size_t nn = (size_t)n * n;
    const int mixed = ctx->opts.precision != DQMC_GPU_FP64;
    HIPCHK(hipStreamCreateWithFlags(&m.w.s, hipStreamNonBlocking));
    if (m.w.alloc(n, R, mixed)) return 1;
    m.w.pivot = m.pivot;
    const char *ov = getenv("DQMC_GPU_OVERLAP");
    if (!(ov && ov[0] == '0')) {
        HIPCHK(hipStreamCreateWithFlags(&m.w2.s, hipStreamNonBlocking));
        if (m.w2.alloc(n, R, 0)) return 1;
        m.w2.pivot = m.pivot;
        // a replica that fails on either stream is skipped on both
        m.w2.status = m.w.status;
        m.w2.reason = m.w.reason;
        HIPCHK(hipEventCreateWithFlags(&m.ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&m.ev_join, hipEventDisableTiming));
    }
    auto own = [&](void *p) {
        if (p) m.owned.push_back(p);
    };
    if (dalloc(&m.expK, nn) || dalloc(&m.expKinv, nn) || dalloc(&m.thop, nn) || dalloc(&m.bipart, n))
        return 1;
    own(m.expK); own(m.expKinv); own(m.thop); own(m.bipart);
    if (dalloc(&m.G, nn * R) || dalloc(&m.Gnew, nn * R) || dalloc(&m.field, (size_t)L * n * R) ||
        dalloc(&m.dvec, (size_t)L * n * R) || dalloc(&m.dinv, (size_t)L * n * R) ||
        dalloc(&m.rng, (size_t)4 * R))
        return 1;
    own(m.G); own(m.Gnew); own(m.field); own(m.dvec); own(m.dinv); own(m.rng);
    if (mixed) {
        if (dalloc(&m.expKf, nn) || dalloc(&m.expKinvf, nn) || dalloc(&m.Gf, nn * R) ||
            dalloc(&m.dvecf, (size_t)L * n * R) || dalloc(&m.dinvf, (size_t)L * n * R))
            return 1;
        own(m.expKf); own(m.expKinvf); own(m.Gf); own(m.dvecf); own(m.dinvf);
    }
    const size_t S = (size_t)(m.M + 1);
    if (dalloc(&m.SU, nn * S * R) || dalloc(&m.SD, (size_t)n * S * R) || dalloc(&m.ST, nn * S * R) ||
        dalloc(&m.Sq, S * R))
        return 1;
    own(m.SU); own(m.SD); own(m.ST); own(m.Sq);
    if (dalloc(&m.SO, S * R) || dalloc(&m.LO, (size_t)R) || dalloc(&m.CO, (size_t)R)) return 1;
    own(m.SO); own(m.LO); own(m.CO);
    HIPCHK(hipMemset(m.SO, 0, sizeof(double) * S * R));
    HIPCHK(hipMemset(m.LO, 0, sizeof(double) * R));
    HIPCHK(hipMemset(m.CO, 0, sizeof(double) * R));
    if (m.alternating && alt_alloc(m)) return 1;
    if (dalloc(&m.LU_, nn * R) || dalloc(&m.LD, (size_t)n * R) || dalloc(&m.LT, nn * R) ||
        dalloc(&m.Lq, R) || dalloc(&m.CU, nn * R) || dalloc(&m.CD, (size_t)n * R) ||
        dalloc(&m.CT, nn * R) || dalloc(&m.Cq, R))
        return 1;
    own(m.LU_); own(m.LD); own(m.LT); own(m.Lq); own(m.CU); own(m.CD); own(m.CT); own(m.Cq);
    if (ctx->opts.drift == DQMC_GPU_DRIFT_SCRATCH) {
        if (dalloc(&m.XU, nn * R) || dalloc(&m.XD, (size_t)n * R) || dalloc(&m.XT, nn * R) || dalloc(&m.Xq, R) ||
            dalloc(&m.XO, R))
            return 1;
        own(m.XU); own(m.XD); own(m.XT); own(m.Xq); own(m.XO);
        HIPCHK(hipMemset(m.XO, 0, sizeof(double) * R));
    }
    if (dalloc(&m.pcount, R) || dalloc(&m.dtmax, R) || dalloc(&m.dtbad, R)) return 1;
    own(m.pcount);
    if (m.rlog_on) {
        if (dalloc(&m.rlog, (size_t)n * R)) return 1;
        own(m.rlog);
        HIPCHK(hipMemset(m.rlog, 0, sizeof(double) * n * R));
    }
    own(m.dtmax);
    own(m.dtbad);
    HIPCHK(hipMemset(m.dtmax, 0, sizeof(unsigned long long) * R));
    HIPCHK(hipMemset(m.dtbad, 0, sizeof(int) * R));
    if (ctx->opts.precision != DQMC_GPU_MIXED_FP32_CHAIN) {
        if (dalloc(&m.BB, (size_t)m.M * R * m.n * m.n)) return 1;
        own(m.BB);
    }
    if (dalloc(&m.acc, R) || dalloc(&m.att, R) || dalloc(&m.acc_prev, R) || dalloc(&m.att_prev, R) ||
        dalloc(&m.sign, R) || dalloc(&m.dmax, R) || dalloc(&m.dsum, R) || dalloc(&m.dcnt, R))
        return 1;
    own(m.acc); own(m.att); own(m.acc_prev); own(m.att_prev); own(m.sign); own(m.dmax); own(m.dsum);
    own(m.dcnt);
    if (dalloc(&m.dtaumax, R) || dalloc(&m.dsweepmax, R) || dalloc(&m.sweeps, 1)) return 1;
    own(m.dtaumax); own(m.dsweepmax); own(m.sweeps);
    HIPCHK(hipMemset(m.dtaumax, 0, sizeof(int) * R));
    HIPCHK(hipMemset(m.dsweepmax, 0, sizeof(unsigned long long) * R));
    HIPCHK(hipMemset(m.sweeps, 0, sizeof(unsigned long long)));
    const int nbin = std::max(m.nbin, 1), nq = std::max(m.nq, 1);
    if (dalloc(&m.bins, (size_t)8 * nbin * R) || dalloc(&m.counts, (size_t)nbin * R) ||
        dalloc(&m.bszz, (size_t)nq * nbin * R) || dalloc(&m.bsperp, (size_t)nq * nbin * R))
        return 1;
    own(m.bins); own(m.counts); own(m.bszz); own(m.bsperp);
    if (m.nq > 0) {
        if (dalloc(&m.phase, (size_t)n * m.nq)) return 1;
        own(m.phase);
    }

    // uploads
    HIPCHK(hipMemcpy(m.expK, m.h_expK.data(), nn * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.expKinv, m.h_expKinv.data(), nn * 8, hipMemcpyHostToDevice));
    if (!m.h_kron.empty()) {
        if (dalloc(&m.kronF, m.h_kron.size())) return 1;
        m.owned.push_back(m.kronF);
        HIPCHK(hipMemcpy(m.kronF, m.h_kron.data(), m.h_kron.size() * 8, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(m.thop, m.h_thop.data(), nn * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.bipart, m.h_bipart.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.field, m.h_fields.data(), (size_t)L * n * R, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.rng, m.h_rng.data(), (size_t)4 * R * 8, hipMemcpyHostToDevice));
    {
        std::vector<double> dv((size_t)L * n * R), di((size_t)L * n * R);
        for (size_t k = 0; k < dv.size(); k++) {
            const int sp = m.h_fields[k] > 0;
            dv[k] = m.expv[sp];
            di[k] = m.expvinv[sp];
        }
        HIPCHK(hipMemcpy(m.dvec, dv.data(), dv.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m.dinv, di.data(), di.size() * 8, hipMemcpyHostToDevice));
        if (mixed) {
            std::vector<float> fv(dv.begin(), dv.end()), fi(di.begin(), di.end());
            HIPCHK(hipMemcpy(m.dvecf, fv.data(), fv.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(m.dinvf, fi.data(), fi.size() * 4, hipMemcpyHostToDevice));
            std::vector<float> k1(m.h_expK.begin(), m.h_expK.end()), k2(m.h_expKinv.begin(), m.h_expKinv.end());
            HIPCHK(hipMemcpy(m.expKf, k1.data(), nn * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(m.expKinvf, k2.data(), nn * 4, hipMemcpyHostToDevice));
        }
    }
    std::vector<double> ones(R, 1.0);
    HIPCHK(hipMemcpy(m.sign, ones.data(), R * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m.acc, 0, R * 8));
    HIPCHK(hipMemset(m.att, 0, R * 8));
    HIPCHK(hipMemset(m.acc_prev, 0, R * 8));
    HIPCHK(hipMemset(m.att_prev, 0, R * 8));
    HIPCHK(hipMemset(m.dmax, 0, R * 8));
    HIPCHK(hipMemset(m.dsum, 0, R * 8));
    HIPCHK(hipMemset(m.dcnt, 0, R * 8));
    HIPCHK(hipMemset(m.bins, 0, (size_t)8 * nbin * R * 8));
    HIPCHK(hipMemset(m.counts, 0, (size_t)nbin * R * sizeof(int)));
    HIPCHK(hipMemset(m.bszz, 0, (size_t)nq * nbin * R * 8));
    HIPCHK(hipMemset(m.bsperp, 0, (size_t)nq * nbin * R * 8));
    if (m.nq > 0) HIPCHK(hipMemcpy(m.phase, m.h_phase.data(), (size_t)n * m.nq * 8, hipMemcpyHostToDevice));
    // the last stack entry is the identity and never changes (src/green.c:205)
    const Udv sm = m.stack(m.M);
    hipLaunchKernelGGL(k_udv_identity, grid2(nn, R), dim3(256), 0, m.w.s, sm.U, sm.sU, sm.D, sm.sD, sm.T,
                       sm.sU, sm.q, sm.sq, n, R, sm.off, sm.so);
    if (launch_check("k_udv_identity")) return 1;
    HIPCHK(hipStreamSynchronize(m.w.s));
    m.allocated = true;
    return 0;
}

// ----------------------------------------------------------------- stages

static bool is_mixed(const dqmc_gpu_ctx *ctx) { return ctx->opts.precision != DQMC_GPU_FP64; }
// B-chains (green_build_Bblock) in FP32 only in DQMC_GPU_MIXED_FP32_CHAIN
static bool chain_f32(const dqmc_gpu_ctx *ctx) { return ctx->opts.precision == DQMC_GPU_MIXED_FP32_CHAIN; }

template <typename T, int W> static void launch_metro(const MetroArgs<T> &a, hipStream_t s)
{
    // at least 512 threads: sites beyond n idle in the proposals but add
    // loads in flight to the rank-p flushes
    const int threads = std::min(METRO_THREADS, std::max(512, ((a.n + 63) / 64) * 64));
    hipLaunchKernelGGL((k_metro_slice<T, W>), dim3(a.batch), dim3(threads), metro_lds_bytes<T>(a.n, W), s, a);
}

template <typename T> static void launch_metro_w(const MetroArgs<T> &a, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (a.W >= 32) {
            launch_metro<T, 32>(a, s);
            return;
        }
    }
    if (a.W >= 16) launch_metro<T, 16>(a, s);
    else launch_metro<T, 8>(a, s);
}

// Metropolis over slice l (src/dqmc.c:656-679): the proposal kernel leaves
// the slice's accepted pairs in C, V (workspace), then one batched GEMM
// with per-replica K applies G -= C V^T (green_delay_flush,
// src/green.c:462-475).
// the rank-P update G -= C^T V of the pending pairs in W3 / W4
// (green_delay_flush, src/green.c:462-475)
static int flush_pending(Model &m)
{
    const int n = m.n;
    const long long nn = (long long)n * n;
    if (n % 64 == 0 && m.flush64) {
        // rank-P update on its own kernel (same sums as the GEMM)
        ProfScope psf_(DQMC_GPU_PROF_FLUSH, m.w.s);
        FlushArgs f;
        f.n = n; f.batch = m.nrep; f.G = m.G; f.sG = nn; f.C = m.w.W3; f.V = m.w.W4; f.sP = nn;
        f.pcount = m.pcount; f.status = m.w.status;
        hipLaunchKernelGGL(k_flush64, dim3((n / 64) * (n / 64) * m.nrep), dim3(256), 0, m.w.s, f);
        return launch_check("k_flush64");
    }
    GemmArgs<double> g = gemm_args<double>(n, n, n, m.nrep, m.w.W3, n, nn, 0, m.w.W4, n, nn, 1, m.G, n, nn);
    g.alpha = -1.0;
    g.beta = 1.0;
    g.Kb = m.pcount;
    return gemm(m.w, g);
}

// FP64 Metropolis of one slice on the submatrix path: m.sub segments of at
// most MSUB_MAX sites, each decided by k_metro_sub, its pending pairs
// written by k_metro_pairs (K in W2, the accepted sites in perm, the pairs
// in W3 / W4) and flushed; with defer the last segment's pairs stay pending
// for the wrap.
static int stage_metro_sub(Model &m, int l, bool defer)
{
    const int n = m.n;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    const int per = (n + m.sub - 1) / m.sub;
    if (per > MSUB_MAX) {
        set_err("submatrix Metropolis segment of %d sites exceeds %d", per, MSUB_MAX);
        return 1;
    }
    for (int s0 = 0; s0 < n; s0 += per) {
        const int s1 = std::min(n, s0 + per), sm = s1 - s0;
        MetroSubArgs a;
        a.n = n; a.L = m.L; a.l = l; a.batch = m.nrep; a.s0 = s0; a.s1 = s1;
        a.G = m.G; a.sG = nn; a.field = m.field; a.sF = sV;
        a.dvec = m.dvec; a.dinv = m.dinv; a.sV = sV;
        a.e[0] = m.expv[0]; a.e[1] = m.expv[1];
        a.einv[0] = m.expvinv[0]; a.einv[1] = m.expvinv[1];
        a.rng = m.rng;
        a.nu[0] = m.nu[0]; a.nu[1] = m.nu[1]; a.nd[0] = m.nd[0]; a.nd[1] = m.nd[1];
        a.acc = m.acc; a.att = m.att; a.sign = m.sign; a.status = m.w.status;
        a.K = m.w.W2; a.sK = nn; a.aq = m.w.perm; a.sA = n; a.pcount = m.pcount; a.rlog = m.rlog;
        {
            ProfScope psp_(DQMC_GPU_PROF_PROPOSE, m.w.s);
            if (sm <= 64)
                hipLaunchKernelGGL((k_metro_sub<1, 8>), dim3(m.nrep), dim3(MSUB_THREADS), metro_sub_lds_bytes(sm),
                                   m.w.s, a);
            else
                hipLaunchKernelGGL((k_metro_sub<2, 12>), dim3(m.nrep), dim3(MSUB_THREADS), metro_sub_lds_bytes(sm),
                                   m.w.s, a);
            if (launch_check("k_metro_sub")) return 1;
            MetroPairsArgs pa;
            pa.n = n; pa.batch = m.nrep; pa.s0 = s0; pa.m = sm;
            pa.G = m.G; pa.sG = nn; pa.K = m.w.W2; pa.sK = nn; pa.aq = m.w.perm; pa.sA = n;
            pa.pcount = m.pcount; pa.status = m.w.status; pa.C = m.w.W3; pa.V = m.w.W4; pa.sP = nn;
            hipLaunchKernelGGL(k_metro_pairs, dim3(((n + MP_COLS - 1) / MP_COLS) * m.nrep), dim3(256), 0, m.w.s, pa);
            if (launch_check("k_metro_pairs")) return 1;
        }
        if (defer && s1 == n) continue;
        if (flush_pending(m)) return 1;
    }
    return 0;
}

static int stage_metro(dqmc_gpu_ctx *ctx, Model &m, int l, bool defer = false)
{
    ProfScope ps_(DQMC_GPU_PROF_METRO, m.w.s);
    if (!is_mixed(ctx) && m.sub > 0) return stage_metro_sub(m, l, defer);
    const int n = m.n;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    const int W = is_mixed(ctx) ? m.Wf : m.W;
    // segment boundaries on window multiples
    const int nwin = (n + W - 1) / W, per = (nwin + m.msplit - 1) / m.msplit;
    for (int s0 = 0; s0 < n; s0 += per * W) {
        const int s1 = std::min(n, s0 + per * W);
        if (is_mixed(ctx)) {
            MetroArgs<float> a;
            a.n = n; a.L = m.L; a.W = m.Wf; a.l = l; a.batch = m.nrep; a.s0 = s0; a.s1 = s1;
            a.G = m.Gf; a.sG = nn; a.field = m.field; a.sF = sV;
            a.dvec = m.dvecf; a.dinv = m.dinvf; a.sV = sV;
            a.e[0] = (float)m.expv[0]; a.e[1] = (float)m.expv[1];
            a.einv[0] = (float)m.expvinv[0]; a.einv[1] = (float)m.expvinv[1];
            a.dvec64 = m.dvec; a.dinv64 = m.dinv;
            a.e64[0] = m.expv[0]; a.e64[1] = m.expv[1];
            a.einv64[0] = m.expvinv[0]; a.einv64[1] = m.expvinv[1];
            a.rng = m.rng;
            a.nu[0] = m.nu[0]; a.nu[1] = m.nu[1]; a.nd[0] = m.nd[0]; a.nd[1] = m.nd[1];
            a.acc = m.acc; a.att = m.att; a.sign = m.sign; a.status = m.w.status;
            a.Cp = m.w.F1; a.Vp = m.w.F2; a.sP = nn; a.pcount = m.pcount; a.rlog = m.rlog;
            {
                ProfScope psp_(DQMC_GPU_PROF_PROPOSE, m.w.s);
                launch_metro_w<float>(a, m.w.s);
            }
            if (launch_check("k_metro_slice")) return 1;
            GemmArgs<float> g =
                gemm_args<float>(n, n, n, m.nrep, m.w.F1, n, nn, 0, m.w.F2, n, nn, 1, m.Gf, n, nn);
            g.alpha = -1.0f;
            g.beta = 1.0f;
            g.Kb = m.pcount;
            if (gemm(m.w, g)) return 1;
        } else {
            MetroArgs<double> a;
            a.n = n; a.L = m.L; a.W = m.W; a.l = l; a.batch = m.nrep; a.s0 = s0; a.s1 = s1;
            a.G = m.G; a.sG = nn; a.field = m.field; a.sF = sV;
            a.dvec = m.dvec; a.dinv = m.dinv; a.sV = sV;
            a.e[0] = m.expv[0]; a.e[1] = m.expv[1];
            a.einv[0] = m.expvinv[0]; a.einv[1] = m.expvinv[1];
            a.dvec64 = nullptr; a.dinv64 = nullptr;
            a.rng = m.rng;
            a.nu[0] = m.nu[0]; a.nu[1] = m.nu[1]; a.nd[0] = m.nd[0]; a.nd[1] = m.nd[1];
            a.acc = m.acc; a.att = m.att; a.sign = m.sign; a.status = m.w.status;
            a.Cp = m.w.W3; a.Vp = m.w.W4; a.sP = nn; a.pcount = m.pcount; a.rlog = m.rlog;
            {
                ProfScope psp_(DQMC_GPU_PROF_PROPOSE, m.w.s);
                launch_metro_w<double>(a, m.w.s);
            }
            if (launch_check("k_metro_slice")) return 1;
            // defer: the last segment's pairs stay pending for the wrap
            if (defer && s1 == n) continue;
            if (flush_pending(m)) return 1;
        }
    }
    return 0;
}

// out = (A_y (x) A_x) applied to the columns (rows = 0) or the rows
// (rows = 1) of in, with the scales of k_kron16; f selects the factor pair:
// 0 expK from the left, 1 expK from the right, 2 expKinv from the left,
// 3 expKinv from the right
static int kron(Model &m, Work &w, int rows, int f, const double *in, double *out, Vec<double> pre,
                Vec<double> post, Vec<double> oth, bool pending = false)
{
    KronArgs a;
    // the last Metropolis segment's rank-P update, left pending in W3 / W4
    // (stage_metro with defer), applied by this column pass
    a.fc = pending ? m.w.W3 : nullptr;
    a.fv = pending ? m.w.W4 : nullptr;
    a.sP = (long long)m.n * m.n;
    a.pcount = m.pcount;
    a.batch = m.nrep;
    a.rows = rows;
    a.in = in;
    a.sIn = (long long)m.n * m.n;
    a.out = out;
    a.sOut = a.sIn;
    a.ax = m.kronF + (size_t)f * 2 * KRON_N;
    a.ay = a.ax + KRON_N;
    a.pre = pre;
    a.post = post;
    a.oth = oth;
    a.status = w.status;
    if (rows)
        hipLaunchKernelGGL(k_kron16<KRON_ROWS>, dim3(KRON_L * m.nrep), dim3(256), 0, w.s, a);
    else if (pending)
        hipLaunchKernelGGL(k_kron16<KRON_COLS_PENDING>, dim3(KRON_L * m.nrep), dim3(256), 0, w.s, a);
    else
        hipLaunchKernelGGL(k_kron16<KRON_COLS>, dim3(KRON_L * m.nrep), dim3(256), 0, w.s, a);
    return launch_check("k_kron16");
}

// whether the last Metropolis segment of a slice leaves its rank-P update to
// the wrap that follows it (separable FP64 products; DQMC_GPU_FUSE=0 off)
static bool fuse_flush(dqmc_gpu_ctx *ctx, const Model &m)
{
    return m.fuse && m.kronF && !is_mixed(ctx);
}

// G <- B_l G B_l^-1 = expK diag(d_l) G diag(1/d_l) expKinv (src/green.c:509-527);
// pending: first apply the rank-P update the Metropolis stage left pending
static int stage_wrap(dqmc_gpu_ctx *ctx, Model &m, int l, bool pending = false)
{
    ProfScope ps_(DQMC_GPU_PROF_WRAP, m.w.s);
    const int n = m.n, R = m.nrep;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    if (is_mixed(ctx)) {
        GemmArgs<float> g = gemm_args<float>(n, n, n, R, m.expKf, n, 0, 0, m.Gf, n, nn, 0, m.w.F1, n, nn);
        g.br = vec<float>(m.dvecf + (long long)l * n, sV);
        g.bc = vec<float>(m.dinvf + (long long)l * n, sV);
        if (gemm(m.w, g)) return 1;
        g = gemm_args<float>(n, n, n, R, m.w.F1, n, nn, 0, m.expKinvf, n, 0, 0, m.Gf, n, nn);
        return gemm(m.w, g);
    }
    if (m.kronF) {
        // W1 = expK diag(d) G diag(1/d) column by column, then G = W1 expKinv row by row
        if (kron(m, m.w, 0, 0, m.G, m.w.W1, vec<double>(m.dvec + (long long)l * n, sV), vnone<double>(),
                 vec<double>(m.dinv + (long long)l * n, sV), pending))
            return 1;
        return kron(m, m.w, 1, 3, m.w.W1, m.G, vnone<double>(), vnone<double>(), vnone<double>());
    }
    GemmArgs<double> g = gemm_args<double>(n, n, n, R, m.expK, n, 0, 0, m.G, n, nn, 0, m.w.W1, n, nn);
    g.br = vec<double>(m.dvec + (long long)l * n, sV);
    g.bc = vec<double>(m.dinv + (long long)l * n, sV);
    if (gemm(m.w, g)) return 1;
    g = gemm_args<double>(n, n, n, R, m.w.W1, n, nn, 0, m.expKinv, n, 0, 0, m.G, n, nn);
    return gemm(m.w, g);
}

// FP32 B-block B_{lb+len-1} ... B_{lb} (green_build_Bblock, src/green.c:151-166)
// into F1 or F2; returns the buffer holding it.
static float *bblock_f32(Model &m, int lb, int len)
{
    const int n = m.n, R = m.nrep;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    hipLaunchKernelGGL(k_scale_cols_f, grid2(nn, R), dim3(256), 0, m.w.s, m.expKf,
                       m.dvecf + (long long)lb * n, sV, m.w.F1, nn, n, R);
    float *cur = m.w.F1, *oth = m.w.F2;
    for (int l = lb + 1; l < lb + len; l++) {
        GemmArgs<float> g = gemm_args<float>(n, n, n, R, m.expKf, n, 0, 0, cur, n, nn, 0, oth, n, nn);
        g.br = vec<float>(m.dvecf + (long long)l * n, sV);
        g.status = m.w.status;
        gemm_launch(g, m.w.s);
        std::swap(cur, oth);
    }
    return cur;
}

// out = B_{lb+len-1} ... B_{lb} (FP64, left-accumulated as green_build_Bblock,
// src/green.c:151-166): B_lb = expK diag(d_lb), then X <- expK diag(d_l) X.
// The product alternates between `out` and W2 so that it ends in `out`.
static int bblock_build(Model &m, int lb, int len, double *out)
{
    const int n = m.n, R = m.nrep;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    double *bufs[2] = {out, m.w.W2};
    int cur = (len - 1) % 2; // the first buffer so that the last product lands in out
    hipLaunchKernelGGL(k_scale_cols_d, grid2(nn, R), dim3(256), 0, m.w.s, m.expK, m.dvec + (long long)lb * n, sV,
                       bufs[cur], nn, n, R);
    if (launch_check("k_scale_cols_d")) return 1;
    for (int l = lb + 1; l < lb + len; l++) {
        if (m.kronF) {
            if (kron(m, m.w, 0, 0, bufs[cur], bufs[cur ^ 1], vec<double>(m.dvec + (long long)l * n, sV),
                     vnone<double>(), vnone<double>()))
                return 1;
        } else {
            GemmArgs<double> g =
                gemm_args<double>(n, n, n, R, m.expK, n, 0, 0, bufs[cur], n, nn, 0, bufs[cur ^ 1], n, nn);
            g.br = vec<double>(m.dvec + (long long)l * n, sV);
            if (gemm(m.w, g)) return 1;
        }
        cur ^= 1;
    }
    return 0;
}

// every block product from the current fields (after the host wrote fields
// or a single step changed them)
static int bblock_refresh(Model &m)
{
    if (!m.BB) return 0;
    const long long nn = (long long)m.n * m.n;
    for (int j = 0; j < m.M; j++) {
        const int lb = j * m.stab;
        if (bblock_build(m, lb, std::min(m.stab, m.L - lb), m.BB + (long long)j * m.nrep * nn)) return 1;
    }
    m.bb_valid = true;
    return 0;
}

// recentre a UDV's D (into `to`, which may be `from`): udv_recenter
static int udv_recenter(Model &m, const Udv &from, const Udv &to)
{
    hipLaunchKernelGGL(k_udv_recenter, dim3(m.nrep), dim3(256), 0, m.w.s, from.D, from.sD, from.off, from.so, to.D,
                       to.sD, to.off, to.so, m.n, m.nrep, m.w.status, m.w.reason);
    return launch_check("k_udv_recenter");
}

// the update's margin check of the centered mode on B (and T for rmul)
static int update_margin(Model &m, const double *B, long long sB, const double *T, long long sT, const double *D,
                         long long sD, int left, int trans = 0)
{
    hipLaunchKernelGGL(k_update_margin, dim3(m.nrep), dim3(256), 0, m.w.s, B, sB, T, sT, D, sD, m.n, left, trans,
                       m.nrep, m.w.status, m.w.reason);
    return launch_check("k_update_margin");
}

// out <- B_{lb+len-1} ... B_{lb} in   (udv_lmul_work, src/linalg.c:515-642);
// `out` may be `in` (in place) or another UDV (the prefix stack's next entry);
// Bpre, when given, is that block product already formed (a cached one)
static int stage_lmul(dqmc_gpu_ctx *ctx, Model &m, const Udv &in, const Udv &out, int lb, int len,
                      const double *Bpre = nullptr)
{
    ProfScope ps_(DQMC_GPU_PROF_LMUL, m.w.s);
    const int n = m.n, R = m.nrep;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    double *cur;
    // centered mode (udv_lmul_centered_work, src/linalg.c:540-560): the
    // input's D is recentred into `out` and the update reads it from there
    Udv src = in;
    if (m.centered) {
        if (udv_recenter(m, in, out)) return 1;
        src.D = out.D;
        src.sD = out.sD;
    }
    const bool aligned = lb % m.stab == 0 && len == std::min(m.stab, m.L - lb);
    if (chain_f32(ctx)) {
        float *Bf = bblock_f32(m, lb, len);
        hipLaunchKernelGGL(k_f2d, grid2(nn, R), dim3(256), 0, m.w.s, Bf, nn, m.w.W3, nn, nn, R);
        GemmArgs<double> g = gemm_args<double>(n, n, n, R, m.w.W3, n, nn, 0, in.U, n, in.sU, 0, m.w.W1, n, nn);
        g.bc = vec(src.D, src.sD);
        if (gemm(m.w, g)) return 1;
        cur = m.w.W1;
    } else if (Bpre || (m.BB && (aligned || m.centered))) {
        // Bblock_j = B_{lb+len-1} ... B_{lb}, left-accumulated (src/green.c:151-166),
        // kept for the next sweep's suffix stack (a block not on the stack
        // grid, centered only, goes to W3); then (Bblock U) diag(D)
        const double *Bj = Bpre;
        if (!Bj) {
            double *Bw = aligned ? m.BB + (long long)(lb / m.stab) * R * nn : m.w.W3;
            if (bblock_build(m, lb, len, Bw)) return 1;
            Bj = Bw;
        }
        if (m.centered && update_margin(m, Bj, nn, nullptr, 0, src.D, src.sD, 1)) return 1;
        GemmArgs<double> g = gemm_args<double>(n, n, n, R, Bj, n, nn, 0, in.U, n, in.sU, 0, m.w.W1, n, nn);
        g.bc = vec(src.D, src.sD);
        if (gemm(m.w, g)) return 1;
        cur = m.w.W1;
    } else {
        // (B U) diag(D) with B applied one slice at a time
        GemmArgs<double> g = gemm_args<double>(n, n, n, R, m.expK, n, 0, 0, in.U, n, in.sU, 0, m.w.W1, n, nn);
        g.br = vec<double>(m.dvec + (long long)lb * n, sV);
        g.bc = vec(in.D, in.sD);
        if (gemm(m.w, g)) return 1;
        cur = m.w.W1;
        double *oth = m.w.W2;
        for (int l = lb + 1; l < lb + len; l++) {
            g = gemm_args<double>(n, n, n, R, m.expK, n, 0, 0, cur, n, nn, 0, oth, n, nn);
            g.br = vec<double>(m.dvec + (long long)l * n, sV);
            if (gemm(m.w, g)) return 1;
            std::swap(cur, oth);
        }
    }
    // QR: U' = Q, D' = diag R, T' = (D'^-1 R) P^T T
    if (qr_update(m.w, cur, out) || update_T(m.w, in, out)) return 1;
    return m.centered ? udv_recenter(m, out, out) : 0;
}

// S[j] = S[j+1] B_{b_{j+1}-1} ... B_{b_j}   (green_stack_build_suffix +
// udv_rmul, src/green.c:203-226, src/linalg.c:662-784), kept transposed:
// stack entry j holds U D T = S[j]^T, so the step is the left update
// S[j]^T = Bblock_j^T S[j+1]^T: M = (Bblock_j^T U) diag(D), QR(M) = Q R,
// U' = Q, D' = diag R, T' = (D'^-1 R) T.  Its QR input is graded by
// columns, as in the left chain, where the blocked Householder QR keeps
// the small entries to working accuracy; the reference's D (T B) is graded
// by rows, where a blocked (compact WY) trailing update loses digits in the
// small rows.  S[M] is the identity either way.
static int stage_rmul(dqmc_gpu_ctx *ctx, Model &m, int j, Work &w)
{
    ProfScope ps_(DQMC_GPU_PROF_STACK, w.s);
    const int n = m.n, R = m.nrep;
    const long long nn = (long long)n * n;
    const int bj = std::min(j * m.stab, m.L), bj1 = std::min((j + 1) * m.stab, m.L);
    const Udv s1 = m.stack(j + 1), s0 = m.stack(j);
    const bool cached = m.BB && m.bb_valid && !chain_f32(ctx);
    if (!cached && &w != &m.w) {
        set_err("suffix update without cached block products runs on the main stream only");
        return 1;
    }
    // the block product Bblock_j: cached, or formed into W3 (main stream)
    const double *Bj;
    if (cached) {
        Bj = m.BB + (long long)j * R * nn;
    } else if (chain_f32(ctx)) {
        float *Bf = bblock_f32(m, bj, bj1 - bj);
        hipLaunchKernelGGL(k_f2d, grid2(nn, R), dim3(256), 0, w.s, Bf, nn, w.W3, nn, nn, R);
        if (launch_check("k_f2d")) return 1;
        Bj = w.W3;
    } else {
        if (bblock_build(m, bj, bj1 - bj, w.W3)) return 1;
        Bj = w.W3;
    }
    // centered mode (udv_rmul_centered, src/linalg.c:687-704): S[j+1]'s D
    // recentred into S[j], from where the update reads it
    const double *Dsrc = s1.D;
    long long sDsrc = s1.sD;
    if (m.centered) {
        if (udv_recenter(m, s1, s0)) return 1;
        Dsrc = s0.D;
        sDsrc = s0.sD;
        if (update_margin(m, Bj, nn, nullptr, 0, Dsrc, sDsrc, 1, 1)) return 1;
    }
    GemmArgs<double> g = gemm_args<double>(n, n, n, R, Bj, n, nn, 1, s1.U, n, s1.sU, 0, w.W1, n, nn);
    g.bc = vec(Dsrc, sDsrc);
    if (gemm(w, g)) return 1; // W1 = Bblock_j^T U diag(D)
    // U' = Q, D' = diag R, T' = (D'^-1 R) P^T T
    if (qr_update(w, w.W1, s0) || update_T(w, s1, s0)) return 1;
    return m.centered ? udv_recenter(m, s0, s0) : 0;
}

static int udv_identity(Model &m, const Udv &u)
{
    const long long nn = (long long)m.n * m.n;
    hipLaunchKernelGGL(k_udv_identity, grid2(nn, m.nrep), dim3(256), 0, m.w.s, u.U, u.sU, u.D, u.sD, u.T,
                       u.sU, u.q, u.sq, m.n, m.nrep, u.off, u.so);
    return launch_check("k_udv_identity");
}

// G <- Gnew after a rebuild with the FP32 copy, the sign reset, the drift
// guard and, when record is set, the distance max|G_wrap - Gnew| as a
// drift sample (DQMC_GPU_DRIFT_REBUILD, and DQMC_GPU_DRIFT_SCRATCH where
// the rebuild is itself the from-scratch G of tau).  A non-finite wrapped
// G (check_old: every rebuild of a sweep) fails the replica
// (DQMC_GPU_FAIL_NONFINITE), as does a non-finite rebuilt one.
static int stage_accept_rebuild(dqmc_gpu_ctx *ctx, Model &m, int record, int tau, int check_old = 1)
{
    ProfScope ps_(DQMC_GPU_PROF_ACCEPT, m.w.s);
    const long long nn = (long long)m.n * m.n;
    const double guard = ctx->opts.drift_guard > 0.0 ? ctx->opts.drift_guard : 0.0;
    const int rec = record && ctx->opts.drift != DQMC_GPU_DRIFT_OFF ? 1 : 0;
    const int measure = rec || guard > 0.0 ? 1 : 0;
    hipLaunchKernelGGL(k_drift_copy, grid2(nn / 4, m.nrep), dim3(256), 0, m.w.s, m.G, m.Gf, m.Gnew, nn, m.n, m.nrep,
                       measure && check_old, check_old, is_mixed(ctx) ? 1 : 0, 1, m.dtmax, m.dtbad, m.w.status);
    if (launch_check("k_drift_copy")) return 1;
    hipLaunchKernelGGL(k_drift_finish, dim3((m.nrep + 255) / 256), dim3(256), 0, m.w.s, m.nrep, rec && check_old,
                       check_old ? guard : 0.0, m.dtmax,
                       m.dtbad, m.dmax, m.dsum, m.dcnt, m.w.status, m.w.reason, m.sign, tau, m.sweeps, m.dtaumax,
                       m.dsweepmax);
    return launch_check("k_drift_finish");
}

// dqmc_record_stab_drift (src/dqmc.c:64-98) with DQMC_GPU_DRIFT_SCRATCH:
// G(tau) from scratch, B_{tau-1} ... B_0 B_{L-1} ... B_tau accumulated in
// blocks of stab slices from tau as green_from_scratch does
// (src/green.c:262-306), into Gnew from the drift UDV, and the distance to
// the wrapped G recorded as a drift sample; G itself is left alone.  tau is
// a block boundary and every cached block product is current (the fields of
// each block are final when its product is formed).
static int stage_drift_scratch(dqmc_gpu_ctx *ctx, Model &m, int tau)
{
    if (ctx->opts.drift != DQMC_GPU_DRIFT_SCRATCH) return 0;
    if (!m.XU || !m.BB || !m.bb_valid) {
        set_err("from-scratch drift needs the FP64 block products");
        return 1;
    }
    const long long nn = (long long)m.n * m.n;
    const Udv x = m.xdrift();
    if (udv_identity(m, x)) return 1;
    const int j0 = tau / m.stab;
    for (int k = 0; k < m.M; k++) {
        const int j = (j0 + k) % m.M, lb = j * m.stab;
        if (stage_lmul(ctx, m, x, x, lb, std::min(m.stab, m.L - lb), m.BB + (long long)j * m.nrep * nn)) return 1;
    }
    if (udv_inv_one_plus(m.w, x, m.Gnew, nn)) return 1;
    hipLaunchKernelGGL(k_drift_copy, grid2(nn / 4, m.nrep), dim3(256), 0, m.w.s, m.G, m.Gf, m.Gnew, nn, m.n, m.nrep,
                       1, 0, is_mixed(ctx) ? 1 : 0, 0, m.dtmax, m.dtbad, m.w.status);
    if (launch_check("k_drift_copy")) return 1;
    hipLaunchKernelGGL(k_drift_finish, dim3((m.nrep + 255) / 256), dim3(256), 0, m.w.s, m.nrep, 1, 0.0, m.dtmax,
                       m.dtbad, m.dmax, m.dsum, m.dcnt, m.w.status, m.w.reason, (double *)nullptr, tau, m.sweeps,
                       m.dtaumax, m.dsweepmax);
    return launch_check("k_drift_finish");
}

// green_from_scratch(G, l0) (src/green.c:262-306): B_{l0-1} ... B_0 B_{L-1}
// ... B_{l0} accumulated into the left UDV in blocks of at most stab slices.
static int stage_from_scratch(dqmc_gpu_ctx *ctx, Model &m, int l0)
{
    const Udv lf = m.left();
    if (udv_identity(m, lf)) return 1;
    int segs[2][2] = {{l0, m.L}, {0, l0}};
    for (auto &sg : segs)
        for (int lb = sg[0]; lb < sg[1]; lb += m.stab)
            if (stage_lmul(ctx, m, lf, lf, lb, std::min(m.stab, sg[1] - lb))) return 1;
    if (udv_inv_one_plus(m.w, lf, m.Gnew, (long long)m.n * m.n)) return 1;
    // G before this rebuild may be anything (unset at init)
    return stage_accept_rebuild(ctx, m, 0, l0, 0);
}

// G <- B_l^-1 G B_l = diag(1/d_l) expKinv G expK diag(d_l)
// (green_wrap_backward, src/green.c:529-546)
static int stage_wrap_back(dqmc_gpu_ctx *ctx, Model &m, int l, bool pending = false)
{
    ProfScope ps_(DQMC_GPU_PROF_WRAP, m.w.s);
    const int n = m.n, R = m.nrep;
    const long long nn = (long long)n * n, sV = (long long)m.L * n;
    if (is_mixed(ctx)) {
        GemmArgs<float> g = gemm_args<float>(n, n, n, R, m.expKinvf, n, 0, 0, m.Gf, n, nn, 0, m.w.F1, n, nn);
        g.ar = vec<float>(m.dinvf + (long long)l * n, sV);
        if (gemm(m.w, g)) return 1;
        g = gemm_args<float>(n, n, n, R, m.w.F1, n, nn, 0, m.expKf, n, 0, 0, m.Gf, n, nn);
        g.bc = vec<float>(m.dvecf + (long long)l * n, sV);
        return gemm(m.w, g);
    }
    if (m.kronF) {
        // W1 = diag(1/d) expKinv G column by column, then G = W1 expK diag(d) row by row
        if (kron(m, m.w, 0, 2, m.G, m.w.W1, vnone<double>(), vec<double>(m.dinv + (long long)l * n, sV),
                 vnone<double>(), pending))
            return 1;
        return kron(m, m.w, 1, 1, m.w.W1, m.G, vnone<double>(), vec<double>(m.dvec + (long long)l * n, sV),
                    vnone<double>());
    }
    GemmArgs<double> g = gemm_args<double>(n, n, n, R, m.expKinv, n, 0, 0, m.G, n, nn, 0, m.w.W1, n, nn);
    g.ar = vec<double>(m.dinv + (long long)l * n, sV);
    if (gemm(m.w, g)) return 1;
    g = gemm_args<double>(n, n, n, R, m.w.W1, n, nn, 0, m.expK, n, 0, 0, m.G, n, nn);
    g.bc = vec<double>(m.dvec + (long long)l * n, sV);
    return gemm(m.w, g);
}

// G at a block boundary from a prefix and a (transposed) suffix UDV
// (green_from_boundary_factors, src/green.c:308-345), after the drift
// sample against the from-scratch G (DQMC_GPU_DRIFT_SCRATCH), then the
// copy with the guard
static int stage_boundary_rebuild(dqmc_gpu_ctx *ctx, Model &m, const Udv &prefix, const Udv &suffix, int tau)
{
    const long long nn = (long long)m.n * m.n;
    if (stage_drift_scratch(ctx, m, tau)) return 1;
    if (m.two_sided) {
        if (udv_inv_two_sided(m.w, prefix, suffix, m.Gnew, nn)) return 1;
    } else {
        if (udv_combine(m.w, prefix, suffix, m.comb(), 1)) return 1;
        if (udv_inv_one_plus(m.w, m.comb(), m.Gnew, nn)) return 1;
    }
    return stage_accept_rebuild(ctx, m, ctx->opts.drift == DQMC_GPU_DRIFT_REBUILD, tau);
}

// Sweep variants: the plain forward sweep, and the two legs of the
// alternating mode, each with or without rebuilding the stack it reads.
enum { SW_FORWARD = 0, SW_FWD_BUILD, SW_FWD_CARRIED, SW_BWD_BUILD, SW_BWD_CARRIED, SW_N };

static int sweep_variant(const Model &m)
{
    if (!m.alternating) return SW_FORWARD;
    if (m.dir == 0) return m.suffix_ok ? SW_FWD_CARRIED : SW_FWD_BUILD;
    return m.prefix_ok ? SW_BWD_CARRIED : SW_BWD_BUILD;
}

// direction and carried stacks after a sweep (src/dqmc.c:806-810, 945-948)
static void sweep_advance(Model &m, int v)
{
    if (v == SW_FORWARD) return;
    const bool fwd = v == SW_FWD_BUILD || v == SW_FWD_CARRIED;
    m.prefix_ok = fwd;
    m.suffix_ok = !fwd;
    m.dir = fwd ? 1 : 0;
}

// One forward sweep (dqmc_sweep_forward, PH branch, src/dqmc.c:609-813).
// carry: the alternating mode's forward leg, whose left chain runs through
// the prefix stack, A[j] = Bblock_{j-1} A[j-1] (stored at every boundary,
// src/dqmc.c:704-706), so that the next backward sweep reads it.
static int sweep_forward(dqmc_gpu_ctx *ctx, Model &m, bool build_suffix, bool carry)
{
    const long long nn = (long long)m.n * m.n;
    // The suffix stack S[M-1] .. S[1] reads only the cached block products
    // Bblock_j, j >= 1, and S[M], and nothing of it is read before the first
    // boundary rebuild; block 0's Metropolis, wraps and left update touch
    // G, the fields, the left UDV and Bblock_0 only.  So the stack goes to
    // the second stream (its own scratch) and runs beside block 0.  Not in
    // centered or FP32-chain modes (their helpers use the main scratch), nor
    // in a profiled sweep, whose per-stage events assume one stream.
    const bool ovl = build_suffix && !carry && m.w2.s && !m.centered && !chain_f32(ctx) && m.BB && m.bb_valid &&
                     m.M >= 2 && g_prof == nullptr;
    if (build_suffix) {
        if (ovl) {
            HIPCHK(hipEventRecord(m.ev_fork, m.w.s));
            HIPCHK(hipStreamWaitEvent(m.w2.s, m.ev_fork, 0));
        }
        for (int j = m.M - 1; j >= 1; j--)
            if (stage_rmul(ctx, m, j, ovl ? m.w2 : m.w)) return 1;
        if (ovl) HIPCHK(hipEventRecord(m.ev_join, m.w2.s));
    }
    bool joined = !ovl;
    const Udv lf = m.left();
    if (!carry && udv_identity(m, lf)) return 1;
    int tau_left = 0;
    const bool fuse = fuse_flush(ctx, m);
    for (int l = 0; l < m.L; l++) {
        if (stage_metro(ctx, m, l, fuse)) return 1;
        if (stage_wrap(ctx, m, l, fuse)) return 1;
        if (((l + 1) % m.stab) == 0 && (l + 1) < m.L) {
            const int tau = l + 1, j = tau / m.stab;
            const Udv cur = carry ? m.alt(j) : lf;
            if (stage_lmul(ctx, m, carry ? m.alt(j - 1) : lf, cur, tau_left, tau - tau_left)) return 1;
            tau_left = tau;
            if (!joined) {
                HIPCHK(hipStreamWaitEvent(m.w.s, m.ev_join, 0));
                joined = true;
            }
            if (stage_boundary_rebuild(ctx, m, cur, m.stack(j), tau)) return 1;
        }
    }
    if (!joined) HIPCHK(hipStreamWaitEvent(m.w.s, m.ev_join, 0));
    // tau = L: the whole chain B_{L-1} ... B_0 (A[M] in the alternating mode,
    // src/dqmc.c:769-771)
    const Udv fin = carry ? m.alt(m.M) : lf;
    if (m.L - tau_left > 0) {
        if (stage_lmul(ctx, m, carry ? m.alt(tau_left / m.stab) : lf, fin, tau_left, m.L - tau_left)) return 1;
    }
    // the end rebuild reads the whole chain B_{L-1} ... B_0 accumulated from
    // the identity in blocks of stab slices: it is the from-scratch G(0) of
    // the drift samples as well
    if (udv_inv_one_plus(m.w, fin, m.Gnew, nn)) return 1;
    return stage_accept_rebuild(ctx, m, 1, 0);
}

// S[j] = S[j+1] Bblock_j in a backward sweep, after the Metropolis of
// block j: its fields are final, so the block product is formed here
// (green_build_Bblock + udv_rmul, src/dqmc.c:828-830, 924-928)
static int stage_suffix_block(dqmc_gpu_ctx *ctx, Model &m, int j)
{
    if (m.BB && !chain_f32(ctx)) {
        const long long nn = (long long)m.n * m.n;
        const int lb = j * m.stab;
        ProfScope ps_(DQMC_GPU_PROF_STACK, m.w.s);
        if (bblock_build(m, lb, std::min(m.stab, m.L - lb), m.BB + (long long)j * m.nrep * nn)) return 1;
    }
    return stage_rmul(ctx, m, j, m.w);
}

// One backward sweep (dqmc_sweep_backward + dqmc_backward_boundary,
// src/dqmc.c:815-950): slices L-1 .. 0 with G carried down by the backward
// wrap; at tau = j s the suffix S[j] = B_{L-1} ... B_tau is accumulated
// into the suffix stack (read by the next forward sweep) and G is rebuilt
// from A[j] and S[j]; at tau = 0, G = (1 + S[0])^-1.  S[M] is the identity.
static int sweep_backward(dqmc_gpu_ctx *ctx, Model &m, bool build_prefix)
{
    const long long nn = (long long)m.n * m.n;
    if (build_prefix) {
        // green_stack_build_prefix (src/green.c:228-250); A[M] is never read
        // by the backward sweep
        for (int j = 1; j < m.M; j++) {
            const int lb = (j - 1) * m.stab;
            if (stage_lmul(ctx, m, m.alt(j - 1), m.alt(j), lb, std::min(m.stab, m.L - lb))) return 1;
        }
    }
    if (stage_wrap_back(ctx, m, m.L - 1)) return 1;
    const bool fuse = fuse_flush(ctx, m);
    for (int l = m.L - 1; l >= 0; l--) {
        // the flush rides on the next backward wrap unless a rebuild (which
        // reads the wrapped G for its drift) or the sweep's end comes first
        const bool f = fuse && l > 0 && l % m.stab != 0;
        if (stage_metro(ctx, m, l, f)) return 1;
        if (l == 0) break;
        if (l % m.stab == 0) {
            const int j = l / m.stab;
            if (stage_suffix_block(ctx, m, j)) return 1;
            if (stage_boundary_rebuild(ctx, m, m.alt(j), m.stack(j), l)) return 1;
        }
        if (stage_wrap_back(ctx, m, l - 1, f)) return 1;
    }
    if (stage_suffix_block(ctx, m, 0)) return 1;
    // G(0) = (1 + S[0])^-1 with S[0] stored transposed: the two-sided form
    // with the identity prefix A[0]
    if (stage_drift_scratch(ctx, m, 0)) return 1;
    if (udv_inv_two_sided(m.w, m.alt(0), m.stack(0), m.Gnew, nn)) return 1;
    return stage_accept_rebuild(ctx, m, ctx->opts.drift == DQMC_GPU_DRIFT_REBUILD, 0);
}

static int sweep_body(dqmc_gpu_ctx *ctx, Model &m, int v)
{
    int rc;
    switch (v) {
    case SW_FORWARD: rc = sweep_forward(ctx, m, true, false); break;
    case SW_FWD_BUILD: rc = sweep_forward(ctx, m, true, true); break;
    case SW_FWD_CARRIED: rc = sweep_forward(ctx, m, false, true); break;
    case SW_BWD_BUILD: rc = sweep_backward(ctx, m, true); break;
    case SW_BWD_CARRIED: rc = sweep_backward(ctx, m, false); break;
    default: set_err("unknown sweep variant %d", v); return 1;
    }
    if (rc) return 1;
    hipLaunchKernelGGL(k_sweep_tick, dim3(1), dim3(64), 0, m.w.s, m.sweeps);
    return launch_check("k_sweep_tick");
}

// One sweep of a model (dqmc_sweep, src/dqmc.c:952-963); each variant is
// captured once as a hipGraph and replayed.
static int launch_sweep(dqmc_gpu_ctx *ctx, Model &m)
{
    if (m.BB && !m.bb_valid && bblock_refresh(m)) return 1;
    const int v = sweep_variant(m);
    if (!ctx->opts.use_graph) {
        if (sweep_body(ctx, m, v)) return 1;
        sweep_advance(m, v);
        return 0;
    }
    if (!m.gexec[v]) {
        HIPCHK(hipStreamBeginCapture(m.w.s, hipStreamCaptureModeThreadLocal));
        const int rc = sweep_body(ctx, m, v);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(m.w.s, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return 1;
        }
        if (e != hipSuccess) {
            set_err("hipStreamEndCapture: %s", hipGetErrorString(e));
            return 1;
        }
        m.graph[v] = gr;
        HIPCHK(hipGraphInstantiate(&m.gexec[v], m.graph[v], nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(m.gexec[v], m.w.s));
    sweep_advance(m, v);
    return 0;
}

static int stage_measure(Model &m, int bin, hipStream_t st = nullptr)
{
    if (!st) st = m.w.s;
    ProfScope ps_(DQMC_GPU_PROF_MEASURE, st);
    MeasureArgs a;
    memset(&a, 0, sizeof a);
    a.n = m.n; a.Lx = m.Lx; a.Ly = m.Ly; a.nq = m.nq; a.batch = m.nrep;
    a.G = m.G; a.sG = (long long)m.n * m.n;
    a.thop = m.thop; a.bipart = m.bipart; a.phase = m.phase;
    a.U = m.U; a.mu = m.mu;
    a.bin = bin; a.nbin = std::max(m.nbin, 1);
    a.bins = m.bins; a.counts = m.counts; a.bszz = m.bszz; a.bsperp = m.bsperp;
    a.sign = m.sign; a.acc = m.acc; a.att = m.att; a.acc_prev = m.acc_prev; a.att_prev = m.att_prev;
    a.status = m.w.status; a.reason = m.w.reason;
    hipLaunchKernelGGL(k_measure, dim3(m.nrep), dim3(256), measure_lds_bytes(m.n), st, a);
    return launch_check("k_measure");
}

// ------------------------------------------------------ persistent sweeps
//
// Models with n <= 64 in FP64, forward sweep order, the two-sided rebuild
// and drift off or against the rebuild run their whole sweep in
// k_sweep_persist (persist.hpp): one launch per sweep for all of them, on
// the stream of the first, with its ticks and measurements after it.

static bool persist_eligible(const dqmc_gpu_ctx *ctx, const Model &m)
{
    return m.persist_ok && m.nrep > 0 && m.allocated && m.n <= PS_NP && !is_mixed(ctx) && !m.alternating &&
           !m.centered && m.two_sided && m.sub == 0 && m.BB != nullptr &&
           ctx->opts.drift != DQMC_GPU_DRIFT_SCRATCH;
}

// the launch set and its device descriptors, rebuilt when the set changes
static int persist_prepare(dqmc_gpu_ctx *ctx)
{
    std::vector<Model *> set;
    for (Model *m : ctx->models)
        if (persist_eligible(ctx, *m)) set.push_back(m);
    // most work first, so that the longest sweeps start first
    std::stable_sort(set.begin(), set.end(), [](const Model *a, const Model *b) {
        return (double)a->n * a->n * a->n * a->L > (double)b->n * b->n * b->n * b->L;
    });
    if (set == ctx->pset) return 0;
    for (Model *m : ctx->models) m->in_pset = false;
    ctx->pset = set;
    ctx->pblocks = 0;
    if (set.empty()) return 0;
    std::vector<PersistModel> pm(set.size());
    for (size_t k = 0; k < set.size(); k++) {
        Model &m = *set[k];
        m.in_pset = true;
        PersistModel &d = pm[k];
        memset(&d, 0, sizeof d);
        d.n = m.n; d.L = m.L; d.stab = m.stab; d.M = m.M; d.first = ctx->pblocks; d.nrep = m.nrep;
        d.pivot = m.pivot ? 1 : 0;
        d.record = ctx->opts.drift == DQMC_GPU_DRIFT_REBUILD ? 1 : 0;
        d.guard = ctx->opts.drift_guard > 0.0 ? ctx->opts.drift_guard : 0.0;
        for (int q = 0; q < 2; q++) {
            d.e[q] = m.expv[q]; d.einv[q] = m.expvinv[q]; d.nu[q] = m.nu[q]; d.nd[q] = m.nd[q];
        }
        d.expK = m.expK; d.expKinv = m.expKinv; d.G = m.G; d.field = m.field; d.dvec = m.dvec; d.dinv = m.dinv;
        d.rng = m.rng; d.BB = m.BB; d.SU = m.SU; d.SD = m.SD; d.ST = m.ST; d.Sq = m.Sq;
        d.LU = m.LU_; d.LD = m.LD; d.LT = m.LT; d.Lq = m.Lq;
        d.acc = m.acc; d.att = m.att; d.sign = m.sign; d.status = m.w.status; d.reason = m.w.reason;
        d.dmax = m.dmax; d.dsum = m.dsum; d.dcnt = m.dcnt; d.dtaumax = m.dtaumax; d.dsweepmax = m.dsweepmax;
        d.sweeps = m.sweeps; d.rlog = m.rlog;
        ctx->pblocks += m.nrep;
    }
    const size_t bytes = pm.size() * sizeof(PersistModel);
    if (bytes > ctx->d_pm_cap) {
        if (ctx->d_pm) HIPCHK(hipFree(ctx->d_pm));
        ctx->d_pm = nullptr;
        ctx->d_pm_cap = 0;
        HIPCHK(hipMalloc(&ctx->d_pm, bytes));
        ctx->d_pm_cap = bytes;
    }
    HIPCHK(hipMemcpy(ctx->d_pm, pm.data(), bytes, hipMemcpyHostToDevice));
    if (!ctx->pev) HIPCHK(hipEventCreateWithFlags(&ctx->pev, hipEventDisableTiming));
    return 0;
}

// one sweep of every model of the launch set (bin >= 0: measured into bin)
static int persist_sweep(dqmc_gpu_ctx *ctx, int bin)
{
    if (ctx->pset.empty()) return 0;
    hipStream_t st = ctx->pset[0]->w.s;
    for (Model *m : ctx->pset)
        if (!m->bb_valid) {
            // block products from fields written by the host or by single steps
            if (bblock_refresh(*m)) return 1;
            HIPCHK(hipStreamSynchronize(m->w.s));
        }
    {
        ProfScope ps_(DQMC_GPU_PROF_PERSIST, st);
        // a profiled sweep also takes the phase timers of workgroup 0
        unsigned long long *tm = nullptr;
        if (g_prof) {
            if (!ctx->ptimers) HIPCHK(hipMalloc(&ctx->ptimers, PS_NTIMER * sizeof(unsigned long long)));
            HIPCHK(hipMemsetAsync(ctx->ptimers, 0, PS_NTIMER * sizeof(unsigned long long), st));
            tm = ctx->ptimers;
        }
        hipLaunchKernelGGL(k_sweep_persist<PS_W>, dim3(ctx->pblocks), dim3(PS_THREADS), ps_lds_bytes(), st, ctx->d_pm,
                           (int)ctx->pset.size(), tm);
        if (launch_check("k_sweep_persist")) return 1;
    }
    for (Model *m : ctx->pset) {
        hipLaunchKernelGGL(k_sweep_tick, dim3(1), dim3(64), 0, st, m->sweeps);
        if (launch_check("k_sweep_tick")) return 1;
        if (bin >= 0 && stage_measure(*m, bin, st)) return 1;
    }
    return 0;
}

// the other models of the launch set wait for its stream
static int persist_join(dqmc_gpu_ctx *ctx)
{
    if (ctx->pset.size() < 2) return 0;
    HIPCHK(hipEventRecord(ctx->pev, ctx->pset[0]->w.s));
    for (size_t k = 1; k < ctx->pset.size(); k++) HIPCHK(hipStreamWaitEvent(ctx->pset[k]->w.s, ctx->pev, 0));
    return 0;
}

static int set_lds_limits()
{
    // function attributes are set once per device (the current one)
    static bool done_dev[64] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_err("device %d beyond the engine's table of 64", dev);
        return 1;
    }
    if (done_dev[dev]) return 0;
    // 160 KiB of LDS per CU on gfx950, less each kernel's static LDS
    const void *kernels[] = {(const void *)k_sweep_persist<PS_W>, (const void *)k_metro_sub<1, 8>, (const void *)k_metro_sub<2, 12>,
                             (const void *)k_metro_slice<double, 16>,
                             (const void *)k_metro_slice<double, 8>,  (const void *)k_metro_slice<float, 32>,
                             (const void *)k_metro_slice<float, 16>,  (const void *)k_metro_slice<float, 8>,
                             (const void *)k_qr_udv<1>, (const void *)k_qr_udv<2>, (const void *)k_qr_udv<4>, (const void *)k_qr_udv<8>,
                             (const void *)k_trsm_unit_upper, (const void *)k_measure, (const void *)k_qr_formq,
                             (const void *)k_qr_formq_reg<16>, (const void *)k_qr_formq_reg<32>,
                             (const void *)k_qr_formq_reg<64>,
                             (const void *)k_gemm_dl<0, 0>, (const void *)k_gemm_dl<0, 1>,
                             (const void *)k_gemm_dl<1, 0>, (const void *)k_gemm_dl<1, 1>,
                             (const void *)k_gemm_big<double, 0, 0>, (const void *)k_gemm_big<double, 1, 0>,
                             (const void *)k_gemm_big<double, 0, 1>, (const void *)k_gemm_big<double, 1, 1>,
                             (const void *)k_gemm_big<float, 0, 0>, (const void *)k_gemm_big<float, 1, 0>,
                             (const void *)k_gemm_big<float, 0, 1>, (const void *)k_gemm_big<float, 1, 1>};
    for (const void *k : kernels) {
        hipFuncAttributes at;
        HIPCHK(hipFuncGetAttributes(&at, k));
        const int dyn = 160 * 1024 - (int)at.sharedSizeBytes;
        HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    }
    done_dev[dev] = true;
    return 0;
}

// ------------------------------------------------------------- autotuner
//
// opts.autotune: after a model's initial rebuild, its Metropolis launch shape
// (delay window W and the number of site segments per slice, each segment
// ending in a rank-P flush) and, on a 16 x 16 lattice, the B-matrix product
// path (separable 16 x 16 transforms or dense GEMMs) are timed on the
// model's own replicas and state, and the fastest is kept.  Every candidate
// starts from the same saved state (G, fields and their diagonals, RNG
// streams, counters, sign), which is restored afterwards, so tuning leaves
// the Markov chains exactly where the initial rebuild put them.  n_wrap
// (stab) is not tuned: it is an input of the reference (the rebuild
// interval), and the drift it allows is part of the result.

struct TuneSave {
    std::vector<void *> bufs;
    struct Item {
        void *dst, *src;
        size_t bytes;
    };
    std::vector<Item> items;
    ~TuneSave()
    {
        for (void *p : bufs) (void)hipFree(p);
    }
    int add(void *live, size_t bytes)
    {
        if (!live || !bytes) return 0;
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, bytes));
        bufs.push_back(p);
        items.push_back({live, p, bytes});
        HIPCHK(hipMemcpy(p, live, bytes, hipMemcpyDeviceToDevice));
        return 0;
    }
    int restore(hipStream_t s)
    {
        for (const Item &it : items) HIPCHK(hipMemcpyAsync(it.dst, it.src, it.bytes, hipMemcpyDeviceToDevice, s));
        return 0;
    }
};

// milliseconds of `reps` slices of f() on the model's stream, from the saved state
template <typename F> static int tune_time(Model &m, TuneSave &sv, int reps, F f, float *ms)
{
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int rc = sv.restore(m.w.s);
    for (int l = 0; l < reps && !rc; l++) rc = f(l); // warm-up pass
    if (!rc) rc = sv.restore(m.w.s);
    if (!rc && hipEventRecord(a, m.w.s) != hipSuccess) rc = 1;
    for (int l = 0; l < reps && !rc; l++) rc = f(l);
    if (!rc && (hipEventRecord(b, m.w.s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(ms, a, b) != hipSuccess))
        rc = 1;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc && g_err.empty()) set_err("autotune timing failed");
    return rc;
}

static int tune_model(dqmc_gpu_ctx *ctx, Model &m, int model_id)
{
    const size_t nn = (size_t)m.n * m.n, R = (size_t)m.nrep, Ln = (size_t)m.L * m.n;
    // the saved state must be the initial rebuild's, which runs on the
    // model's (non-blocking) stream
    HIPCHK(hipStreamSynchronize(m.w.s));
    TuneSave sv;
    if (sv.add(m.G, nn * R * 8) || sv.add(m.Gf, m.Gf ? nn * R * 4 : 0) || sv.add(m.field, Ln * R) ||
        sv.add(m.dvec, Ln * R * 8) || sv.add(m.dinv, Ln * R * 8) || sv.add(m.dvecf, m.dvecf ? Ln * R * 4 : 0) ||
        sv.add(m.dinvf, m.dinvf ? Ln * R * 4 : 0) || sv.add(m.rng, 32 * R) || sv.add(m.acc, 8 * R) ||
        sv.add(m.att, 8 * R) || sv.add(m.sign, 8 * R))
        return 1;
    const int reps = std::min(m.L, 4);
    const bool mixed = is_mixed(ctx);
    const size_t elem = mixed ? sizeof(float) : sizeof(double);
    const int Wmax = metro_window(m.n, mixed ? 32 : 16, elem);
    int bestW = mixed ? m.Wf : m.W, bestS = m.msplit;
    // the candidates are windowed shapes: the submatrix path
    // (DQMC_GPU_METRO_SUB) measured slower at every size, and tuning
    // replaces it
    m.sub = 0;
    double *const kf = m.kronF;
    bool bestK = kf != nullptr;
    float best = 1e30f;
    std::string tried;
    // a slice as the sweep runs it: Metropolis (the last segment's flush
    // left to the wrap where that applies), then the wrap
    auto slice = [&](int l) {
        const bool f = fuse_flush(ctx, m);
        return stage_metro(ctx, m, l, f) || stage_wrap(ctx, m, l, f);
    };
    // product paths to try: separable (when expK splits) and dense; the
    // mixed modes' wrap is a dense FP32 GEMM either way, so they keep theirs
    std::vector<int> paths;
    if (kf) paths.push_back(1);
    if (!kf || !mixed) paths.push_back(0);
    for (int useK : paths) {
        m.kronF = useK ? kf : nullptr;
        for (int W : {8, 16, 32}) {
            if (W > Wmax) continue;
            const int nwin = (m.n + W - 1) / W;
            for (int sp = 1; sp <= 4 && sp <= nwin; sp++) {
                (mixed ? m.Wf : m.W) = W;
                m.msplit = sp;
                float ms = 0.0f;
                if (tune_time(m, sv, reps, slice, &ms)) {
                    m.kronF = kf;
                    return 1;
                }
                char buf[64];
                snprintf(buf, sizeof buf, " %sW%d/S%d %.1f", kf ? (useK ? "sep:" : "dense:") : "", W, sp,
                         1e3 * ms / reps);
                tried += buf;
                if (ms < best) {
                    best = ms;
                    bestW = W;
                    bestS = sp;
                    bestK = useK == 1;
                }
            }
        }
    }
    (mixed ? m.Wf : m.W) = bestW;
    m.msplit = bestS;
    m.kronF = bestK ? kf : nullptr; // an unused factor buffer stays owned by the model
    const char *prod = bestK ? "separable" : "dense";
    m.tuned_separable = m.kronF != nullptr;
    if (sv.restore(m.w.s)) return 1;
    HIPCHK(hipStreamSynchronize(m.w.s));
    fprintf(stderr,
            "dqmc_gpu autotune: model %d (n=%d, L=%d, %d replicas): window %d, %d site segment(s) per slice, %s "
            "B-matrix products; us per slice (Metropolis and wrap):%s\n",
            model_id, m.n, m.L, m.nrep, bestW, bestS, prod, tried.c_str());
    return 0;
}

// ------------------------------------------------------------------- ABI

static uint64_t splitmix_next(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double host_xoshiro(uint64_t s[4])
{
    const uint64_t x = s[1] * 5;
    const uint64_t out = ((x << 7) | (x >> 57)) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
    return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

extern "C" {

const char *dqmc_gpu_last_error(void) { return g_err.c_str(); }

int dqmc_gpu_create(dqmc_gpu_ctx **out, int device, const dqmc_gpu_opts *o)
{
    if (!out) {
        set_err("null output pointer");
        return 1;
    }
    *out = nullptr;
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) {
        set_err("device %d not present (%d devices)", device, count);
        return 1;
    }
    HIPCHK(hipSetDevice(device));
    if (set_lds_limits()) return 1;
    dqmc_gpu_ctx *c = new dqmc_gpu_ctx();
    c->device = device;
    c->opts.precision = DQMC_GPU_FP64;
    c->opts.use_graph = 1;
    c->opts.drift = 0;
    c->opts.autotune = 0;
    c->opts.drift_guard = 0.0;
    if (o) c->opts = *o;
    if (c->opts.precision < DQMC_GPU_FP64 || c->opts.precision > DQMC_GPU_MIXED_FP32_CHAIN) {
        set_err("unknown precision %d", c->opts.precision);
        delete c;
        return 1;
    }
    if (c->opts.drift < DQMC_GPU_DRIFT_OFF || c->opts.drift > DQMC_GPU_DRIFT_REBUILD) {
        set_err("unknown drift mode %d", c->opts.drift);
        delete c;
        return 1;
    }
    if (!std::isfinite(c->opts.drift_guard) || c->opts.drift_guard < 0.0) {
        set_err("drift_guard must be a finite non-negative bound");
        delete c;
        return 1;
    }
    // the FP32 wrap and delayed updates can leave G far from the rebuilt one
    // (at 16x16, U=4, beta=8 by O(10)): a mixed context needs the guard
    if (c->opts.precision != DQMC_GPU_FP64 && !(c->opts.drift_guard > 0.0)) {
        set_err("mixed precision needs a drift guard (opts.drift_guard > 0)");
        delete c;
        return 1;
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        set_err("hipEventCreate failed");
        delete c;
        return 1;
    }
    *out = c;
    return 0;
}

// M = F_y (x) F_x for a 16 x 16 lattice (site x + 16 y, column-major M):
// F_x from the (0,0) block, F_y from the (0,0) entries of the blocks over
// M[0][0]; kept when F_y (x) F_x reproduces M to 1e-12 of its largest entry
static bool kron_split(const std::vector<double> &M, std::vector<double> &fx, std::vector<double> &fy)
{
    const int L = KRON_L, n = KRON_N;
    if ((int)M.size() != n * n || !(M[0] != 0.0)) return false;
    fx.assign(L * L, 0.0);
    fy.assign(L * L, 0.0);
    for (int r = 0; r < L; r++)
        for (int c = 0; c < L; c++) {
            fx[r * L + c] = M[r + (size_t)c * n];
            fy[r * L + c] = M[L * r + (size_t)(L * c) * n] / M[0];
        }
    double mx = 0.0, err = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double v = M[i + (size_t)j * n];
            mx = std::max(mx, std::fabs(v));
            err = std::max(err, std::fabs(fy[(i / L) * L + j / L] * fx[(i % L) * L + j % L] - v));
        }
    return std::isfinite(err) && err <= 1e-12 * mx;
}

int dqmc_gpu_add_model(dqmc_gpu_ctx *ctx, int n, int L, int stab, int delay_k, const double *expK,
                       const double *expKinv, const double *thop, const int *bipart, double U, double dtau,
                       double mu, double lambda, int Lx, int Ly, int *model_id)
{
    if (!ctx || !expK || !expKinv || !thop || !bipart || !model_id) {
        set_err("null argument");
        return 1;
    }
    if (ctx->initialised) {
        set_err("models must be added before dqmc_gpu_init");
        return 1;
    }
    if (n <= 0 || n > METRO_MAX_N || L <= 0 || stab <= 0 || Lx <= 0 || Ly <= 0 || Lx * Ly != n) {
        set_err("invalid model sizes n=%d L=%d stab=%d Lx=%d Ly=%d", n, L, stab, Lx, Ly);
        return 1;
    }
    if (!(dtau > 0.0) || !(U >= 0.0) || !std::isfinite(lambda)) {
        set_err("invalid U=%g dtau=%g lambda=%g", U, dtau, lambda);
        return 1;
    }
    Model *m = new Model();
    const size_t nn = (size_t)n * n;
    m->n = n; m->L = L; m->stab = stab; m->M = (L + stab - 1) / stab;
    m->Lx = Lx; m->Ly = Ly; m->U = U; m->dtau = dtau; m->mu = mu; m->lambda = lambda;
    m->cap = delay_k;
    m->W = metro_window(n, delay_k > 0 ? delay_k : 16, sizeof(double));
    m->msplit = n >= 128 ? 2 : 1;
    if (const char *e = getenv("DQMC_GPU_METRO_SPLIT")) m->msplit = std::max(1, std::min(8, atoi(e)));
    if (const char *e = getenv("DQMC_GPU_REBUILD")) m->two_sided = strcmp(e, "combine") != 0;
    if (const char *e = getenv("DQMC_GPU_PIVOT")) m->pivot = e[0] != '0';
    if (const char *e = getenv("DQMC_GPU_FUSE")) m->fuse = e[0] != '0';
    if (const char *e = getenv("DQMC_GPU_RATIO_LOG")) m->rlog_on = e[0] == '1';
    if (const char *e = getenv("DQMC_GPU_FLUSH64")) m->flush64 = e[0] != '0';
    if (const char *e = getenv("DQMC_GPU_PERSIST")) m->persist_ok = e[0] != '0';
    // FP64: DQMC_GPU_METRO_SUB=k runs the Metropolis on the submatrix path
    // with k site segments (at least enough for segments of MSUB_MAX sites);
    // the autotuner also times it
    if (ctx->opts.precision == DQMC_GPU_FP64) {
        if (const char *e = getenv("DQMC_GPU_METRO_SUB")) {
            const int v = atoi(e);
            m->sub = v <= 0 ? 0 : std::min(n, std::max((n + MSUB_MAX - 1) / MSUB_MAX, v));
        }
    }
    m->Wf = metro_window(n, delay_k > 0 ? delay_k : 32, sizeof(float));
    // expv for sigma = +1 (src/green.c:63-67), N = exp(-2 lambda sigma s) - 1 (src/field.c:10-17)
    const double c = -dtau * U / 2.0;
    m->expv[0] = std::exp(-lambda + c);
    m->expv[1] = std::exp(lambda + c);
    m->expvinv[0] = 1.0 / m->expv[0];
    m->expvinv[1] = 1.0 / m->expv[1];
    m->nu[0] = std::exp(-2.0 * lambda * 1.0 * -1.0) - 1.0;
    m->nu[1] = std::exp(-2.0 * lambda * 1.0 * 1.0) - 1.0;
    m->nd[0] = std::exp(-2.0 * lambda * -1.0 * -1.0) - 1.0;
    m->nd[1] = std::exp(-2.0 * lambda * -1.0 * 1.0) - 1.0;
    m->h_expK.assign(expK, expK + nn);
    m->h_expKinv.assign(expKinv, expKinv + nn);
    {
        const char *e = getenv("DQMC_GPU_KRON");
        std::vector<double> ex, ey, fx, fy;
        if (!(e && e[0] == '0') && Lx == KRON_L && Ly == KRON_L && n == KRON_N && kron_split(m->h_expK, ex, ey) &&
            kron_split(m->h_expKinv, fx, fy)) {
            auto tr = [](const std::vector<double> &a) {
                std::vector<double> t(a.size());
                for (int r = 0; r < KRON_L; r++)
                    for (int c = 0; c < KRON_L; c++) t[c * KRON_L + r] = a[r * KRON_L + c];
                return t;
            };
            for (const std::vector<double> &f : {ex, ey, tr(ex), tr(ey), fx, fy, tr(fx), tr(fy)})
                m->h_kron.insert(m->h_kron.end(), f.begin(), f.end());
        }
    }
    m->h_thop.assign(thop, thop + nn);
    m->h_bipart.assign(bipart, bipart + n);
    *model_id = (int)ctx->models.size();
    ctx->models.push_back(m);
    return 0;
}

int dqmc_gpu_add_replicas(dqmc_gpu_ctx *ctx, int model_id, int nrep, const unsigned long long *seeds,
                          const signed char *fields, int *first_replica_id)
{
    if (!ctx || !seeds || !first_replica_id || model_id < 0 || model_id >= (int)ctx->models.size() ||
        nrep <= 0) {
        set_err("invalid add_replicas arguments");
        return 1;
    }
    if (ctx->initialised) {
        set_err("replicas must be added before dqmc_gpu_init");
        return 1;
    }
    Model &m = *ctx->models[model_id];
    if (m.nrep > 0 && ctx->models.size() > 1 && model_id != (int)ctx->models.size() - 1) {
        set_err("replicas of a model must be added before the next model");
        return 1;
    }
    if (m.nrep == 0) {
        int first = 0;
        for (Model *o : ctx->models)
            if (o != &m) first += o->nrep;
        m.first = first;
    }
    const size_t Ln = (size_t)m.L * m.n;
    for (int r = 0; r < nrep; r++) {
        uint64_t st[4], x = seeds[r];
        for (int q = 0; q < 4; q++) st[q] = splitmix_next(&x); // rng_seed, src/rng.c:16-21
        for (size_t k = 0; k < Ln; k++) {                       // field_init, src/field.c:20-22
            const double u = host_xoshiro(st);
            if (fields) {
                const signed char v = fields[(size_t)r * Ln + k];
                if (v != 1 && v != -1) {
                    set_err("field values must be +-1");
                    return 1;
                }
                m.h_fields.push_back(v);
            } else {
                m.h_fields.push_back(u < 0.5 ? -1 : 1);
            }
        }
        for (int q = 0; q < 4; q++) m.h_rng.push_back(st[q]);
        m.seeds.push_back(seeds[r]);
    }
    *first_replica_id = m.first + m.nrep;
    m.nrep += nrep;
    ctx->nrep_total += nrep;
    return 0;
}

int dqmc_gpu_set_measure(dqmc_gpu_ctx *ctx, int model_id, int nq, const double *phase_cos, int nbin)
{
    if (!ctx || model_id < 0 || model_id >= (int)ctx->models.size() || nq < 0 || nbin < 0 ||
        (nq > 0 && !phase_cos)) {
        set_err("invalid set_measure arguments");
        return 1;
    }
    if (ctx->initialised) {
        set_err("measurement must be configured before dqmc_gpu_init");
        return 1;
    }
    Model &m = *ctx->models[model_id];
    m.nq = nq;
    m.nbin = nbin;
    m.h_phase.assign(phase_cos, phase_cos + (size_t)m.n * nq);
    return 0;
}

int dqmc_gpu_set_sweep_order(dqmc_gpu_ctx *ctx, int model_id, int order)
{
    if (!ctx || model_id < 0 || model_id >= (int)ctx->models.size() ||
        (order != DQMC_GPU_SWEEP_FORWARD && order != DQMC_GPU_SWEEP_ALTERNATING)) {
        set_err("invalid set_sweep_order arguments");
        return 1;
    }
    Model &m = *ctx->models[model_id];
    m.alternating = order == DQMC_GPU_SWEEP_ALTERNATING;
    m.dir = 0;
    m.prefix_ok = m.suffix_ok = false;
    if (m.alternating && m.allocated) {
        HIPCHK(hipSetDevice(ctx->device));
        if (alt_alloc(m)) return 1;
        HIPCHK(hipStreamSynchronize(m.w.s));
    }
    return 0;
}

int dqmc_gpu_set_rebuild(dqmc_gpu_ctx *ctx, int model_id, int mode)
{
    if (!ctx || model_id < 0 || model_id >= (int)ctx->models.size() ||
        (mode != DQMC_GPU_REBUILD_COMBINE && mode != DQMC_GPU_REBUILD_TWO_SIDED &&
         mode != DQMC_GPU_REBUILD_CENTERED)) {
        set_err("invalid set_rebuild arguments");
        return 1;
    }
    if (ctx->initialised) {
        set_err("the rebuild mode must be set before dqmc_gpu_init");
        return 1;
    }
    if (mode == DQMC_GPU_REBUILD_CENTERED && ctx->opts.precision == DQMC_GPU_MIXED_FP32_CHAIN) {
        set_err("the centered rebuild needs the FP64 B-chain");
        return 1;
    }
    Model &m = *ctx->models[model_id];
    m.centered = mode == DQMC_GPU_REBUILD_CENTERED;
    if (m.centered) m.two_sided = true; // centered factors meet in the two-sided inverse
    return 0;
}

static int sync_all(dqmc_gpu_ctx *ctx)
{
    for (Model *m : ctx->models) HIPCHK(hipStreamSynchronize(m->w.s));
    return 0;
}

int dqmc_gpu_init(dqmc_gpu_ctx *ctx)
{
    if (!ctx) {
        set_err("null context");
        return 1;
    }
    HIPCHK(hipSetDevice(ctx->device));
    for (Model *m : ctx->models) {
        if (m->nrep == 0) continue;
        if (!m->allocated && model_alloc(ctx, *m)) return 1;
        if (stage_from_scratch(ctx, *m, 0)) return 1; // also forms every block product
        m->bb_valid = m->BB != nullptr;
    }
    if (persist_prepare(ctx)) return 1;
    if (ctx->opts.autotune)
        for (size_t k = 0; k < ctx->models.size(); k++) {
            Model *m = ctx->models[k];
            if (m->nrep > 0 && m->in_pset) {
                fprintf(stderr,
                        "dqmc_gpu autotune: model %zu (n=%d, L=%d, %d replicas): persistent sweep kernel "
                        "(one workgroup per replica, G in LDS, 16-site windows)\n",
                        k, m->n, m->L, m->nrep);
                continue;
            }
            if (m->nrep > 0 && tune_model(ctx, *m, (int)k)) return 1;
        }
    if (sync_all(ctx)) return 1;
    ctx->initialised = true;
    return 0;
}

static int require_init(dqmc_gpu_ctx *ctx)
{
    if (!ctx) {
        set_err("null context");
        return 1;
    }
    if (!ctx->initialised) {
        set_err("dqmc_gpu_init has not been called");
        return 1;
    }
    HIPCHK(hipSetDevice(ctx->device));
    return 0;
}

int dqmc_gpu_sweep(dqmc_gpu_ctx *ctx, int nsweep, int bin)
{
    if (require_init(ctx)) return 1;
    for (Model *m : ctx->models)
        if (bin >= 0 && (m->nbin <= 0 || bin >= m->nbin)) {
            set_err("bin %d outside [0, %d)", bin, m->nbin);
            return 1;
        }
    if (sync_all(ctx)) return 1;
    if (persist_prepare(ctx)) return 1;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->models.empty() ? nullptr : ctx->models[0]->w.s));
    for (int s = 0; s < nsweep; s++) {
        if (persist_sweep(ctx, bin)) return 1;
        for (Model *m : ctx->models) {
            if (m->nrep == 0 || m->in_pset) continue;
            if (launch_sweep(ctx, *m)) return 1;
            if (bin >= 0 && stage_measure(*m, bin)) return 1;
        }
    }
    if (persist_join(ctx)) return 1;
    if (bin < 0)
        for (Model *m : ctx->models) {
            if (m->nrep == 0) continue;
            hipLaunchKernelGGL(k_acc_mark, dim3((m->nrep + 255) / 256), dim3(256), 0, m->w.s, m->acc, m->att,
                               m->acc_prev, m->att_prev, m->nrep);
            if (launch_check("k_acc_mark")) return 1;
        }
    for (size_t k = 1; k < ctx->models.size(); k++) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHK(hipEventRecord(e, ctx->models[k]->w.s));
        HIPCHK(hipStreamWaitEvent(ctx->models[0]->w.s, e, 0));
        HIPCHK(hipEventDestroy(e));
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->models.empty() ? nullptr : ctx->models[0]->w.s));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    return sync_all(ctx);
}

int dqmc_gpu_run(dqmc_gpu_ctx *ctx, int nwarm, int nmeas_per_bin, int nbin)
{
    if (require_init(ctx)) return 1;
    double total = 0.0;
    if (nwarm > 0) {
        if (dqmc_gpu_sweep(ctx, nwarm, -1)) return 1;
        total += ctx->last_ms;
    }
    for (int b = 0; b < nbin; b++) {
        if (dqmc_gpu_sweep(ctx, nmeas_per_bin, b)) return 1;
        total += ctx->last_ms;
    }
    ctx->last_ms = total;
    return 0;
}

int dqmc_gpu_last_elapsed_ms(dqmc_gpu_ctx *ctx, double *ms)
{
    if (!ctx || !ms) return 1;
    *ms = ctx->last_ms;
    return 0;
}

int dqmc_gpu_read_bins(dqmc_gpu_ctx *ctx, int replica, double *bins8, int *counts, double *szz, double *sperp)
{
    if (require_init(ctx)) return 1;
    Model *m;
    int r;
    if (model_for_replica(ctx, replica, &m, &r)) return 1;
    if (sync_all(ctx)) return 1;
    const int nbin = std::max(m->nbin, 1);
    if (bins8) HIPCHK(hipMemcpy(bins8, m->bins + (size_t)8 * nbin * r, (size_t)8 * nbin * 8, hipMemcpyDeviceToHost));
    if (counts) HIPCHK(hipMemcpy(counts, m->counts + (size_t)nbin * r, (size_t)nbin * sizeof(int), hipMemcpyDeviceToHost));
    if (m->nq > 0) {
        if (szz) HIPCHK(hipMemcpy(szz, m->bszz + (size_t)m->nq * nbin * r, (size_t)m->nq * nbin * 8, hipMemcpyDeviceToHost));
        if (sperp) HIPCHK(hipMemcpy(sperp, m->bsperp + (size_t)m->nq * nbin * r, (size_t)m->nq * nbin * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}

int dqmc_gpu_read_status(dqmc_gpu_ctx *ctx, int *status, int *failure_reason, unsigned long long *accepted,
                         unsigned long long *attempts, double *sign)
{
    if (require_init(ctx)) return 1;
    if (sync_all(ctx)) return 1;
    for (Model *m : ctx->models) {
        const int R = m->nrep, f = m->first;
        if (status) HIPCHK(hipMemcpy(status + f, m->w.status, R * sizeof(int), hipMemcpyDeviceToHost));
        if (failure_reason) HIPCHK(hipMemcpy(failure_reason + f, m->w.reason, R * sizeof(int), hipMemcpyDeviceToHost));
        if (accepted) HIPCHK(hipMemcpy(accepted + f, m->acc, R * 8, hipMemcpyDeviceToHost));
        if (attempts) HIPCHK(hipMemcpy(attempts + f, m->att, R * 8, hipMemcpyDeviceToHost));
        if (sign) HIPCHK(hipMemcpy(sign + f, m->sign, R * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}

int dqmc_gpu_read_drift(dqmc_gpu_ctx *ctx, double *max_inf, double *mean_inf, unsigned long long *samples)
{
    if (require_init(ctx)) return 1;
    if (sync_all(ctx)) return 1;
    for (Model *m : ctx->models) {
        const int R = m->nrep, f = m->first;
        std::vector<double> mx(R), sm(R);
        std::vector<unsigned long long> ct(R);
        HIPCHK(hipMemcpy(mx.data(), m->dmax, R * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(sm.data(), m->dsum, R * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ct.data(), m->dcnt, R * 8, hipMemcpyDeviceToHost));
        for (int r = 0; r < R; r++) {
            if (max_inf) max_inf[f + r] = mx[r];
            if (mean_inf) mean_inf[f + r] = ct[r] ? sm[r] / (double)ct[r] : 0.0;
            if (samples) samples[f + r] = ct[r];
        }
    }
    return 0;
}

int dqmc_gpu_read_drift_at(dqmc_gpu_ctx *ctx, int *max_tau, unsigned long long *max_sweep)
{
    if (require_init(ctx)) return 1;
    if (sync_all(ctx)) return 1;
    for (Model *m : ctx->models) {
        const int R = m->nrep, f = m->first;
        std::vector<int> tt(R);
        std::vector<unsigned long long> ss(R);
        HIPCHK(hipMemcpy(tt.data(), m->dtaumax, R * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ss.data(), m->dsweepmax, R * 8, hipMemcpyDeviceToHost));
        for (int r = 0; r < R; r++) {
            if (max_tau) max_tau[f + r] = tt[r];
            if (max_sweep) max_sweep[f + r] = ss[r];
        }
    }
    return 0;
}

int dqmc_gpu_read_tuning(dqmc_gpu_ctx *ctx, int model_id, int *window, int *segments, int *separable)
{
    if (!ctx || model_id < 0 || model_id >= (int)ctx->models.size()) {
        set_err("invalid read_tuning arguments");
        return 1;
    }
    const Model &m = *ctx->models[model_id];
    const bool sub = !is_mixed(ctx) && m.sub > 0;
    if (window) *window = sub ? 0 : is_mixed(ctx) ? m.Wf : m.W;
    if (segments) *segments = sub ? m.sub : m.msplit;
    if (separable) *separable = m.kronF != nullptr ? 1 : 0;
    return 0;
}

int dqmc_gpu_read_ratios(dqmc_gpu_ctx *ctx, int replica, double *ratios)
{
    Model *m;
    int r;
    if (!ratios || model_for_replica(ctx, replica, &m, &r)) return 1;
    if (!m->rlog) {
        set_err("ratio log off: set DQMC_GPU_RATIO_LOG=1 before adding the model");
        return 1;
    }
    if (sync_all(ctx)) return 1;
    HIPCHK(hipMemcpy(ratios, m->rlog + (size_t)r * m->n, sizeof(double) * m->n, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_gpu_read_green(dqmc_gpu_ctx *ctx, int replica, double *g)
{
    if (require_init(ctx)) return 1;
    Model *m;
    int r;
    if (!g || model_for_replica(ctx, replica, &m, &r)) return 1;
    if (sync_all(ctx)) return 1;
    const size_t nn = (size_t)m->n * m->n;
    if (is_mixed(ctx)) {
        // the live G of the slice loop is the FP32 copy
        std::vector<float> f(nn);
        HIPCHK(hipMemcpy(f.data(), m->Gf + nn * r, nn * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nn; k++) g[k] = (double)f[k];
        return 0;
    }
    HIPCHK(hipMemcpy(g, m->G + nn * r, nn * 8, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_gpu_write_green(dqmc_gpu_ctx *ctx, int replica, const double *g)
{
    if (require_init(ctx)) return 1;
    Model *m;
    int r;
    if (!g || model_for_replica(ctx, replica, &m, &r)) return 1;
    if (sync_all(ctx)) return 1;
    const size_t nn = (size_t)m->n * m->n;
    HIPCHK(hipMemcpy(m->G + nn * r, g, nn * 8, hipMemcpyHostToDevice));
    if (m->Gf) {
        std::vector<float> f(g, g + nn);
        HIPCHK(hipMemcpy(m->Gf + nn * r, f.data(), nn * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

int dqmc_gpu_read_field(dqmc_gpu_ctx *ctx, int replica, signed char *s)
{
    if (require_init(ctx)) return 1;
    Model *m;
    int r;
    if (!s || model_for_replica(ctx, replica, &m, &r)) return 1;
    if (sync_all(ctx)) return 1;
    const size_t Ln = (size_t)m->L * m->n;
    HIPCHK(hipMemcpy(s, m->field + Ln * r, Ln, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_gpu_write_field(dqmc_gpu_ctx *ctx, int replica, const signed char *s)
{
    if (require_init(ctx)) return 1;
    Model *m;
    int r;
    if (!s || model_for_replica(ctx, replica, &m, &r)) return 1;
    if (sync_all(ctx)) return 1;
    const size_t Ln = (size_t)m->L * m->n;
    std::vector<double> dv(Ln), di(Ln);
    for (size_t k = 0; k < Ln; k++) {
        if (s[k] != 1 && s[k] != -1) {
            set_err("field values must be +-1");
            return 1;
        }
        dv[k] = m->expv[s[k] > 0];
        di[k] = m->expvinv[s[k] > 0];
    }
    HIPCHK(hipMemcpy(m->field + Ln * r, s, Ln, hipMemcpyHostToDevice));
    m->bb_valid = false;
    m->prefix_ok = m->suffix_ok = false; // dqmc_invalidate_carried (src/dqmc.c:31-35)
    HIPCHK(hipMemcpy(m->dvec + Ln * r, dv.data(), Ln * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->dinv + Ln * r, di.data(), Ln * 8, hipMemcpyHostToDevice));
    if (m->dvecf) {
        std::vector<float> fv(dv.begin(), dv.end()), fi(di.begin(), di.end());
        HIPCHK(hipMemcpy(m->dvecf + Ln * r, fv.data(), Ln * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m->dinvf + Ln * r, fi.data(), Ln * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

int dqmc_gpu_read_rng(dqmc_gpu_ctx *ctx, int replica, unsigned long long *s4)
{
    if (require_init(ctx)) return 1;
    Model *m;
    int r;
    if (!s4 || model_for_replica(ctx, replica, &m, &r)) return 1;
    if (sync_all(ctx)) return 1;
    HIPCHK(hipMemcpy(s4, m->rng + 4 * (size_t)r, 32, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_gpu_step(dqmc_gpu_ctx *ctx, int step, int arg)
{
    if (require_init(ctx)) return 1;
    for (Model *m : ctx->models) {
        if (m->nrep == 0) continue;
        if ((step == DQMC_GPU_STEP_PROPOSE_SLICE || step == DQMC_GPU_STEP_WRAP) && (arg < 0 || arg >= m->L)) {
            set_err("slice %d outside [0, %d)", arg, m->L);
            return 1;
        }
        int rc;
        switch (step) {
        case DQMC_GPU_STEP_PROPOSE_SLICE:
            rc = stage_metro(ctx, *m, arg);
            m->bb_valid = false;
            m->prefix_ok = m->suffix_ok = false;
            break;
        case DQMC_GPU_STEP_WRAP: rc = stage_wrap(ctx, *m, arg); break;
        case DQMC_GPU_STEP_REBUILD:
            if (arg < 0 || arg >= m->L) {
                set_err("tau %d outside [0, %d)", arg, m->L);
                return 1;
            }
            rc = stage_from_scratch(ctx, *m, arg);
            break;
        default: set_err("unknown step %d", step); return 1;
        }
        if (rc) return 1;
    }
    return sync_all(ctx);
}

int dqmc_gpu_profile_sweep(dqmc_gpu_ctx *ctx, int nsweep, double *ms, int *launches)
{
    if (require_init(ctx)) return 1;
    if (!ms || !launches || nsweep < 1) {
        set_err("invalid profile_sweep arguments");
        return 1;
    }
    if (sync_all(ctx)) return 1;
    if (persist_prepare(ctx)) return 1;
    std::vector<ProfRec> recs;
    g_prof = &recs;
    int rc = 0;
    // phase times of the persistent kernel's workgroup 0 (100 MHz ticks)
    unsigned long long ptm[PS_NTIMER] = {};
    for (int s = 0; s < nsweep && !rc; s++) {
        rc = persist_sweep(ctx, -1);
        if (!rc && !ctx->pset.empty()) {
            unsigned long long t[PS_NTIMER];
            HIPCHK(hipStreamSynchronize(ctx->pset[0]->w.s));
            HIPCHK(hipMemcpy(t, ctx->ptimers, sizeof t, hipMemcpyDeviceToHost));
            for (int k = 0; k < PS_NTIMER; k++) ptm[k] += t[k];
        }
        for (Model *m : ctx->pset)
            if (!rc && m->nbin > 0) rc = stage_measure(*m, 0, ctx->pset[0]->w.s);
        for (Model *m : ctx->models) {
            if (rc) break;
            if (m->nrep == 0 || m->in_pset) continue;
            rc = (m->BB && !m->bb_valid) ? bblock_refresh(*m) : 0;
            const int v = sweep_variant(*m);
            if (!rc) rc = sweep_body(ctx, *m, v);
            if (!rc) sweep_advance(*m, v);
            if (!rc && m->nbin > 0) rc = stage_measure(*m, 0);
            if (rc) break;
        }
    }
    g_prof = nullptr;
    if (!rc) rc = persist_join(ctx);
    const int sync_rc = sync_all(ctx);
    for (int c = 0; c < DQMC_GPU_PROF_N; c++) {
        ms[c] = 0.0;
        launches[c] = 0;
    }
    for (ProfRec &r : recs) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            ms[r.cls] += t;
            launches[r.cls] += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    // the persistent kernel's phases (workgroup 0, in-kernel wall clock)
    // into the stage classes they correspond to; launches stay those of the
    // batched stages
    if (!ctx->pset.empty()) {
        const double tick_ms = 1e-5; // 100 MHz
        ms[DQMC_GPU_PROF_STACK] += ptm[PT_SUFFIX] * tick_ms;
        ms[DQMC_GPU_PROF_METRO] += ptm[PT_METRO] * tick_ms;
        ms[DQMC_GPU_PROF_WRAP] += (ptm[PT_PRODUCT] + ptm[PT_WRAP]) * tick_ms;
        ms[DQMC_GPU_PROF_LMUL] += ptm[PT_LMUL] * tick_ms;
        ms[DQMC_GPU_PROF_INV] += (ptm[PT_REBUILD_QR] + ptm[PT_REBUILD_SOLVE] + ptm[PT_REBUILD_G]) * tick_ms;
        if (getenv("DQMC_GPU_PERSIST_PHASES")) {
            static const char *nm[PS_NTIMER] = {"suffix", "metro", "product", "wrap", "lmul", "rebuild_qr",
                                                "rebuild_solve", "rebuild_g"};
            fprintf(stderr, "dqmc_gpu persistent phases (ms, workgroup 0, %d sweeps):", nsweep);
            for (int k = 0; k < PS_NTIMER; k++) fprintf(stderr, " %s %.4f", nm[k], ptm[k] * tick_ms);
            fprintf(stderr, "\n");
        }
    }
    return rc || sync_rc;
}

void dqmc_gpu_destroy(dqmc_gpu_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (Model *m : ctx->models) {
        if (m->w.s) (void)hipStreamSynchronize(m->w.s);
        for (int v = 0; v < SW_N; v++) {
            if (m->gexec[v]) (void)hipGraphExecDestroy(m->gexec[v]);
            if (m->graph[v]) (void)hipGraphDestroy(m->graph[v]);
        }
        for (void *p : m->owned) (void)hipFree(p);
        if (m->w2.s) (void)hipStreamSynchronize(m->w2.s);
        m->w.release();
        if (m->w.s) (void)hipStreamDestroy(m->w.s);
        m->w2.release();
        if (m->w2.s) (void)hipStreamDestroy(m->w2.s);
        if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
        if (m->ev_join) (void)hipEventDestroy(m->ev_join);
        delete m;
    }
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->pev) (void)hipEventDestroy(ctx->pev);
    if (ctx->d_pm) (void)hipFree(ctx->d_pm);
    if (ctx->ptimers) (void)hipFree(ctx->ptimers);
    delete ctx;
}

// ------------------------------------------- batched linear algebra on host

struct HostBatch {
    Work w;
    std::vector<double *> bufs;
    ~HostBatch()
    {
        for (double *p : bufs) (void)hipFree(p);
        w.release();
        if (w.s) (void)hipStreamDestroy(w.s);
    }
};

static int hb_init(HostBatch &h, int device, int n, int batch)
{
    if (n <= 0 || n > METRO_MAX_N || batch <= 0) {
        set_err("invalid n=%d batch=%d", n, batch);
        return 1;
    }
    HIPCHK(hipSetDevice(device));
    if (set_lds_limits()) return 1;
    HIPCHK(hipStreamCreateWithFlags(&h.w.s, hipStreamNonBlocking));
    return h.w.alloc(n, batch, 0);
}

static int hb_upload(HostBatch &h, const double *src, size_t count, double **dst)
{
    if (dalloc(dst, count)) return 1;
    h.bufs.push_back(*dst);
    if (src) HIPCHK(hipMemcpy(*dst, src, count * 8, hipMemcpyHostToDevice));
    return 0;
}

static int hb_int(HostBatch &h, int count, int **dst)
{
    double *p;
    if (dalloc(&p, (size_t)count)) return 1;
    h.bufs.push_back(p);
    *dst = (int *)p;
    return 0;
}

static int hb_status(HostBatch &h)
{
    HIPCHK(hipStreamSynchronize(h.w.s));
    std::vector<int> st(h.w.nb), rs(h.w.nb);
    HIPCHK(hipMemcpy(st.data(), h.w.status, h.w.nb * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rs.data(), h.w.reason, h.w.nb * sizeof(int), hipMemcpyDeviceToHost));
    for (int b = 0; b < h.w.nb; b++)
        if (st[b]) {
            set_err("batch entry %d failed (reason %d)", b, rs[b]);
            return 1;
        }
    return 0;
}

static int hb_lr_mul(int device, int n, int batch, double *U, double *D, double *T, const double *B, int left)
{
    if (!U || !D || !T || !B) {
        set_err("null argument");
        return 1;
    }
    HostBatch h;
    if (hb_init(h, device, n, batch)) return 1;
    const size_t nn = (size_t)n * n;
    double *dU, *dD, *dT, *dB;
    int *q;
    if (hb_upload(h, U, nn * batch, &dU) || hb_upload(h, D, (size_t)n * batch, &dD) ||
        hb_upload(h, T, nn * batch, &dT) || hb_upload(h, B, nn * batch, &dB) || hb_int(h, batch, &q))
        return 1;
    Work &w = h.w;
    const long long snn = (long long)nn;
    if (left) {
        // M = (B U) diag(D); QR -> U' = Q, D' = diag R, T' = (D'^-1 R) T
        GemmArgs<double> g = gemm_args<double>(n, n, n, batch, dB, n, snn, 0, dU, n, snn, 0, w.W1, n, snn);
        g.bc = vec<double>(dD, n);
        if (gemm(w, g) || qr(w, w.W1, dU, snn, w.W3, snn, dD, n, q, 1)) return 1;
        g = gemm_args<double>(n, n, n, batch, w.W3, n, snn, 0, dT, n, snn, 0, w.W4, n, snn);
        if (gemm(w, g) || copy_nn(w, w.W4, snn, dT, snn, snn)) return 1;
    } else {
        // M = diag(D) (T B); QR -> U' = U Q, D' = diag R, T' = D'^-1 R
        GemmArgs<double> g = gemm_args<double>(n, n, n, batch, dT, n, snn, 0, dB, n, snn, 0, w.W1, n, snn);
        g.cr = vec<double>(dD, n);
        if (gemm(w, g) || qr(w, w.W1, w.W2, snn, dT, snn, dD, n, q, 1)) return 1;
        g = gemm_args<double>(n, n, n, batch, dU, n, snn, 0, w.W2, n, snn, 0, w.W4, n, snn);
        if (gemm(w, g) || copy_nn(w, w.W4, snn, dU, snn, snn)) return 1;
    }
    if (hb_status(h)) return 1;
    HIPCHK(hipMemcpy(U, dU, nn * batch * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(D, dD, (size_t)n * batch * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(T, dT, nn * batch * 8, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_gpu_udv_lmul(int device, int n, int batch, double *U, double *D, double *T, const double *B)
{
    return hb_lr_mul(device, n, batch, U, D, T, B, 1);
}

int dqmc_gpu_udv_rmul(int device, int n, int batch, double *U, double *D, double *T, const double *B)
{
    return hb_lr_mul(device, n, batch, U, D, T, B, 0);
}

// sgn det A = sgn det Q * prod sgn R_ii from a QR of A, one workgroup per
// batch entry; 0 when a diagonal entry of R is zero or not finite
__global__ void k_qr_det_sign(const int *qsign, const double *D, int n, int *out, int batch, const int *status)
{
    const int b = blockIdx.x;
    if (b >= batch || status[b]) return;
    __shared__ int neg, bad;
    if (threadIdx.x == 0) {
        neg = 0;
        bad = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = D[(long long)b * n + i];
        if (d < 0.0) atomicXor(&neg, 1);
        if (!(d != 0.0) || !isfinite(d)) bad = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[b] = bad ? 0 : qsign[b] * (neg ? -1 : 1);
}

// Orthogonal U from outside the engine carries no tracked determinant sign:
// it is taken on the device from a QR of U (the reference takes it from an
// LU, la_logdet, src/linalg.c:255-286); U is copied first, the QR destroys
// its input.
static int device_det_sign(Work &w, const double *U, int *out)
{
    const long long nn = (long long)w.n * w.n;
    if (copy_nn(w, U, nn, w.W4, nn, nn)) return 1;
    if (qr(w, w.W4, nullptr, 0, nullptr, 0, w.V5, w.n, w.I2, 1)) return 1;
    hipLaunchKernelGGL(k_qr_det_sign, dim3(w.nb), dim3(256), 0, w.s, w.I2, w.V5, w.n, out, w.nb, w.status);
    return launch_check("k_qr_det_sign");
}

int dqmc_gpu_udv_combine(int device, int n, int batch, const double *Ul, const double *Dl, const double *Tl,
                         const double *Ur, const double *Dr, const double *Tr, double *Uo, double *Do,
                         double *To)
{
    if (!Ul || !Dl || !Tl || !Ur || !Dr || !Tr || !Uo || !Do || !To) {
        set_err("null argument");
        return 1;
    }
    HostBatch h;
    if (hb_init(h, device, n, batch)) return 1;
    const size_t nn = (size_t)n * n;
    double *a[9];
    int *q[3];
    const double *src[9] = {Ul, Dl, Tl, Ur, Dr, Tr, nullptr, nullptr, nullptr};
    const size_t cnt[9] = {nn, (size_t)n, nn, nn, (size_t)n, nn, nn, (size_t)n, nn};
    for (int k = 0; k < 9; k++)
        if (hb_upload(h, src[k], cnt[k] * batch, &a[k])) return 1;
    for (int k = 0; k < 3; k++)
        if (hb_int(h, batch, &q[k])) return 1;
    if (device_det_sign(h.w, a[0], q[0])) return 1;
    const Udv l{a[0], a[1], a[2], q[0], (long long)nn, n, 1};
    const Udv r{a[3], a[4], a[5], q[1], (long long)nn, n, 1};
    const Udv o{a[6], a[7], a[8], q[2], (long long)nn, n, 1};
    if (udv_combine(h.w, l, r, o) || hb_status(h)) return 1;
    HIPCHK(hipMemcpy(Uo, a[6], nn * batch * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Do, a[7], (size_t)n * batch * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(To, a[8], nn * batch * 8, hipMemcpyDeviceToHost));
    return 0;
}

static int inv_one_plus_batch(int device, int n, int batch, const double *U, const double *D, const double *T,
                              const double *off, double *g, int *sign, int *reason)
{
    if (!U || !D || !T || !g) {
        set_err("null argument");
        return 1;
    }
    HostBatch h;
    if (hb_init(h, device, n, batch)) return 1;
    const size_t nn = (size_t)n * n;
    double *dU, *dD, *dT, *dG, *dO = nullptr;
    int *q;
    if (hb_upload(h, U, nn * batch, &dU) || hb_upload(h, D, (size_t)n * batch, &dD) ||
        hb_upload(h, T, nn * batch, &dT) || hb_upload(h, nullptr, nn * batch, &dG) || hb_int(h, batch, &q))
        return 1;
    if (off && hb_upload(h, off, (size_t)batch, &dO)) return 1;
    if (device_det_sign(h.w, dU, q)) return 1;
    const Udv u{dU, dD, dT, q, (long long)nn, n, 1, dO, 1};
    if (udv_inv_one_plus(h.w, u, dG, (long long)nn)) return 1;
    if (reason) {
        HIPCHK(hipStreamSynchronize(h.w.s));
        HIPCHK(hipMemcpy(reason, h.w.reason, batch * sizeof(int), hipMemcpyDeviceToHost));
    }
    if (hb_status(h)) return 1;
    HIPCHK(hipMemcpy(g, dG, nn * batch * 8, hipMemcpyDeviceToHost));
    if (sign) HIPCHK(hipMemcpy(sign, h.w.det_sign, batch * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_gpu_udv_inv_one_plus(int device, int n, int batch, const double *U, const double *D, const double *T,
                              double *g, int *sign)
{
    return inv_one_plus_batch(device, n, batch, U, D, T, nullptr, g, sign, nullptr);
}

int dqmc_gpu_udv_inv_one_plus_off(int device, int n, int batch, const double *U, const double *D, const double *T,
                                  const double *log_offset, double *g, int *sign, int *reason)
{
    if (!log_offset) {
        set_err("null argument");
        return 1;
    }
    return inv_one_plus_batch(device, n, batch, U, D, T, log_offset, g, sign, reason);
}

int dqmc_gpu_measure(int device, int Lx, int Ly, int batch, const double *gu, const double *thop,
                     const int *bipart, double U, int nq, const double *phase_cos, double *scal5, double *szz,
                     double *sperp)
{
    const int n = Lx * Ly;
    if (!gu || !thop || !bipart || !scal5 || (nq > 0 && (!phase_cos || !szz || !sperp)) || nq < 0) {
        set_err("null argument");
        return 1;
    }
    HostBatch h;
    if (hb_init(h, device, n, batch)) return 1;
    const size_t nn = (size_t)n * n;
    double *dG, *dth, *dph, *ds5, *dz, *dp;
    int *dbp;
    if (hb_upload(h, gu, nn * batch, &dG) || hb_upload(h, thop, nn, &dth) ||
        hb_upload(h, phase_cos, (size_t)n * std::max(nq, 1), &dph) || hb_upload(h, nullptr, (size_t)5 * batch, &ds5) ||
        hb_upload(h, nullptr, (size_t)std::max(nq, 1) * batch, &dz) ||
        hb_upload(h, nullptr, (size_t)std::max(nq, 1) * batch, &dp) || hb_int(h, n, &dbp))
        return 1;
    HIPCHK(hipMemcpy(dbp, bipart, n * sizeof(int), hipMemcpyHostToDevice));
    MeasureArgs a;
    memset(&a, 0, sizeof a);
    a.n = n; a.Lx = Lx; a.Ly = Ly; a.nq = nq; a.batch = batch;
    a.G = dG; a.sG = (long long)nn; a.thop = dth; a.bipart = dbp; a.phase = dph; a.U = U;
    a.scal5 = ds5; a.szz = dz; a.sperp = dp; a.bin = -1;
    a.status = h.w.status; a.reason = h.w.reason;
    hipLaunchKernelGGL(k_measure, dim3(batch), dim3(256), measure_lds_bytes(n), h.w.s, a);
    if (launch_check("k_measure") || hb_status(h)) return 1;
    HIPCHK(hipMemcpy(scal5, ds5, (size_t)5 * batch * 8, hipMemcpyDeviceToHost));
    if (nq > 0) {
        HIPCHK(hipMemcpy(szz, dz, (size_t)nq * batch * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(sperp, dp, (size_t)nq * batch * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}

// ---------------------------------------- inverse without allocations
//
// dqmc_gpu_udv_inv_one_plus_noalloc: the batched (1 + U D T)^-1 on device
// buffers the caller owns, for callers that cannot allocate (code under
// graph capture, serving loops, harnesses that forbid hipMalloc).  Its
// workspace is the file-scope device array `scratch`, NOALLOC_BYTES long;
// each device holds its own copy, whose address hipGetSymbolAddress gives
// once per device (noalloc_scratch) and which host code then uses in
// place of the array's name.  All calls on one device share that copy:
// calls that overlap on one device race.

#define NOALLOC_BYTES (64u << 20)
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static void *g_noalloc_ptr[64];
static bool g_noalloc_ready[64];

static int noalloc_scratch(void **ptr)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_err("device %d beyond the engine's table of 64", dev);
        return 1;
    }
    if (!g_noalloc_ready[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return 1;
        }
        if (set_lds_limits()) return 1;
        g_noalloc_ptr[dev] = p;
        g_noalloc_ready[dev] = true;
    }
    *ptr = g_noalloc_ptr[dev];
    return 0;
}

// bytes of scratch for n, batch, each piece 256-byte aligned
static size_t noalloc_layout(int n, int batch, Work *w, int **q, unsigned char *base)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> unsigned char * {
        unsigned char *p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t nn = (size_t)n * n * batch, nv = (size_t)n * batch;
    const int npan = (n + QR_NB - 1) / QR_NB;
    double **mats[] = {&w->W1, &w->W2, &w->W3, &w->W4};
    for (double **m : mats) *m = (double *)take(nn * 8);
    double **vecs[] = {&w->V1, &w->V2, &w->V3, &w->V4, &w->V5};
    for (double **v : vecs) *v = (double *)take(nv * 8);
    w->Tws = (double *)take((size_t)npan * QR_NB * QR_NB * batch * 8);
    int **ints[] = {&w->I1, &w->I2, &w->status, &w->reason, &w->det_sign, q};
    for (int **v : ints) *v = (int *)take((size_t)batch * 4);
    w->perm = (int *)take(nv * 4);
    return off;
}

// status and reason cleared by a kernel (no memset node under capture)
__global__ void k_clear_status(int *status, int *reason, int batch)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch) {
        status[b] = 0;
        reason[b] = 0;
    }
}

// sign 0 for a failed entry; the reason copied out when asked for
__global__ void k_noalloc_out(const int *det_sign, const int *status, const int *reason, int *sign, int *reason_out,
                              int batch)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    sign[b] = status[b] ? 0 : det_sign[b];
    if (reason_out) reason_out[b] = status[b] ? reason[b] : 0;
}

int dqmc_gpu_udv_inv_one_plus_noalloc(int n, int batch, const double *U, const double *D, const double *T, double *g,
                                      int *sign, int *reason, void *stream)
{
    if (!U || !D || !T || !g || !sign) {
        set_err("null argument");
        return 1;
    }
    if (n <= 0 || n > METRO_MAX_N || batch <= 0) {
        set_err("invalid n=%d batch=%d", n, batch);
        return 1;
    }
    Work w;
    int *q = nullptr;
    const size_t need = noalloc_layout(n, batch, &w, &q, nullptr);
    if (need > NOALLOC_BYTES) {
        set_err("n=%d batch=%d needs %zu bytes of workspace, above the %u of dqmc_gpu_udv_inv_one_plus_noalloc", n,
                batch, need, NOALLOC_BYTES);
        return 1;
    }
    void *base = nullptr;
    if (noalloc_scratch(&base)) return 1;
    noalloc_layout(n, batch, &w, &q, (unsigned char *)base);
    w.n = n;
    w.nb = batch;
    w.s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_clear_status, dim3((batch + 255) / 256), dim3(256), 0, w.s, w.status, w.reason, batch);
    if (launch_check("k_clear_status")) return 1;
    if (device_det_sign(w, U, q)) return 1;
    const Udv u{const_cast<double *>(U), const_cast<double *>(D), const_cast<double *>(T), q, (long long)n * n, n, 1};
    if (udv_inv_one_plus(w, u, g, (long long)n * n)) return 1;
    hipLaunchKernelGGL(k_noalloc_out, dim3((batch + 255) / 256), dim3(256), 0, w.s, w.det_sign, w.status, w.reason,
                       sign, reason, batch);
    return launch_check("k_noalloc_out");
}

} // extern "C"
