// Persistent sweep for lattices of at most 64 sites: one 512-thread
// workgroup per replica runs a whole forward sweep (dqmc_sweep_forward, PH
// branch, src/dqmc.c:609-813) in a single launch, with G, expK, expK^-1 and
// the running block product resident in LDS.  One launch covers the
// replicas of every such model of a context, whatever their n, L and stab
// (uneven batches): each workgroup reads its model's descriptor.
//
// Per sweep and replica:
//   suffix stack S[M-1] .. S[1] (src/green.c:203-226) from the block
//     products of the previous sweep, kept transposed as in the batched
//     path: S[j]^T = Bblock_j^T S[j+1]^T as a left update with a column
//     pre-pivoted QR;
//   for every slice l: Metropolis over the n sites (src/dqmc.c:656-679) in
//     windows of 16: wave 0 decides the window's proposals on the 16 x 16
//     block of G in its registers (green_delay_ratio / _accept,
//     src/green.c:384-460), every thread forms the window's pairs from G,
//     and G -= C V^T is applied in LDS on MFMA before the next window (the
//     delayed update flushed every 16 sites, green_delay_flush,
//     src/green.c:462-475); then the block product P <- B_l P
//     (green_build_Bblock, src/green.c:151-166) and the wrap
//     G <- B_l G B_l^-1 (green_wrap, src/green.c:509-527), three in-place
//     LDS GEMMs on FP64 MFMA;
//   at each block boundary tau = j s: P goes to the block-product cache,
//     the left UDV is updated (udv_lmul, src/linalg.c:515-642, pre-pivoted
//     QR) and G is rebuilt from it and S[j] by the two-sided inverse
//     (src/linalg.c:796-1115 in the one-QR form of the batched path), with
//     the drift sample, the guard and the sign reset of the batched path;
//     at tau = L the suffix is the identity (udv_inv_one_plus).
// The QR, the explicit Q, Q^T applied to a matrix and the unit-triangular
// solve run on matrices held in registers, one column per lane and eight
// rows per wave, the column steps synchronised through small LDS partial
// sums.
#pragma once
#include "common.hpp"

#define PS_NP 64                 // largest lattice
#define PS_LD (PS_NP + 1)        // LDS leading dimension (odd: row reads hit distinct banks)
#define PS_MAT (PS_NP * PS_LD)   // doubles per LDS matrix
#define PS_THREADS 512
#define PS_W 16                  // Metropolis window (delay depth)
// phase timers of a profiled launch: suffix stack, Metropolis, spins and
// block product, wrap, left update, rebuild QR, rebuild solve, new G with
// drift and reloads
#define PS_NTIMER 8
enum { PT_SUFFIX, PT_METRO, PT_PRODUCT, PT_WRAP, PT_LMUL, PT_REBUILD_QR, PT_REBUILD_SOLVE, PT_REBUILD_G };

// One model of the launch (device memory, read-only during the sweep).
struct __align__(8) PersistModel {
    int n, L, stab, M, first, nrep;  // workgroups [first, first + nrep) run this model
    int pivot;                       // column pre-pivoted UDV updates
    int record;                      // drift samples at rebuilds (DQMC_GPU_DRIFT_REBUILD)
    double guard;                    // drift guard (0: none)
    double e[2], einv[2], nu[2], nd[2];
    const double *expK, *expKinv;    // n x n
    double *G;                       // R x n^2
    signed char *field;              // R x L x n
    double *dvec, *dinv;             // R x L x n
    uint64_t *rng;                   // R x 4
    double *BB;                      // M x R x n^2 block products
    double *SU, *SD, *ST;            // suffix stack, R x (M+1) entries
    int *Sq;
    double *LU, *LD, *LT;            // left UDV, R entries
    int *Lq;
    unsigned long long *acc, *att;
    double *sign;
    int *status, *reason;
    double *dmax, *dsum;
    unsigned long long *dcnt;
    int *dtaumax;
    unsigned long long *dsweepmax;
    const unsigned long long *sweeps;
    double *rlog;                    // null, or n ratios per replica (latest slice)
};

// What a replica's sweep carries from one phase to the next, in LDS:
// thread 0's RNG stream, running sign and counters, which slot holds G,
// expK, expK^-1 and the block product, and the failure flag of a phase
// (1: non-finite reflector, 2: zero or non-finite D, 3: non-finite G,
// 4: drift above the guard).
struct PsState {
    uint64_t rng[4];
    double sgn;
    unsigned long long nacc, natt;
    int g, k, ki, p;
    int flag;
    int timing;                  // phase timers on (workgroup 0 of a profiled launch)
    unsigned long long t0;       // the last phase boundary (wall clock, 100 MHz)
    unsigned long long tm[PS_NTIMER];
};

// LDS carve-up (doubles unless noted), about 162.5 KB
struct PsLds {
    double *slot[4];
    double *Cw, *Vw;     // window pairs, q-major (PS_W x PS_NP each); the QR
                         // phases reuse them as partial-sum and row buffers
    double *cwin, *vwin, *facw;
    double *draws;       // two slices' uniform draws (slice l in draws + (l & 1) PS_NP)
    double *dl, *dli;    // diagonal of the slice's B and its inverse
    double *DL, *DR;     // D of the left UDV, of the suffix entry
    double *lbi, *ls, *rbi, *rs; // Db^-1 and Ds of DL, DR
    double *DM, *tau, *keys, *scl;
    double *rk;          // 2 x 64: row k of a QR step
    double *red;         // 16 reduction slots
    int *perm;
    signed char *fs, *flip; // spins of two slices (slice l at fs + (l & 1) PS_NP), flips
    unsigned *mask;      // the accepts of a window, two alternating words
    int *flag;
    PsState *st;
    PersistModel *pm;    // the workgroup's model, copied from device memory
};

__host__ __device__ static inline size_t ps_lds_bytes()
{
    return sizeof(double) * (4 * (size_t)PS_MAT + 2 * PS_W * PS_NP + 2 * PS_W * PS_W + PS_W + 2 * PS_NP + 12 * PS_NP +
                             128 + 16) +
           sizeof(int) * PS_NP + 3 * PS_NP + 16 + sizeof(PsState) + sizeof(PersistModel);
}


__device__ __forceinline__ PsLds ps_carve(unsigned char *raw)
{
    PsLds s;
    double *p = (double *)raw;
    for (int k = 0; k < 4; k++) {
        s.slot[k] = p;
        p += PS_MAT;
    }
    s.Cw = p; p += PS_W * PS_NP;
    s.Vw = p; p += PS_W * PS_NP;
    s.cwin = p; p += PS_W * PS_W;
    s.vwin = p; p += PS_W * PS_W;
    s.facw = p; p += PS_W;
    s.draws = p; p += 2 * PS_NP;
    s.dl = p; p += PS_NP;
    s.dli = p; p += PS_NP;
    s.DL = p; p += PS_NP;
    s.DR = p; p += PS_NP;
    s.lbi = p; p += PS_NP;
    s.ls = p; p += PS_NP;
    s.rbi = p; p += PS_NP;
    s.rs = p; p += PS_NP;
    s.DM = p; p += PS_NP;
    s.tau = p; p += PS_NP;
    s.keys = p; p += PS_NP;
    s.scl = p; p += PS_NP;
    s.rk = p; p += 128;
    s.red = p; p += 16;
    s.perm = (int *)p;
    s.fs = (signed char *)(s.perm + PS_NP);
    s.flip = s.fs + 2 * PS_NP;
    s.mask = (unsigned *)(s.flip + PS_NP);
    s.flag = (int *)(s.mask + 2);
    s.st = (PsState *)(s.mask + 4);
    s.flag = &s.st->flag;
    s.pm = (PersistModel *)(s.st + 1);
    return s;
}

// close a phase of the profiled workgroup: its time since the last
// boundary is added to timer k (called after a barrier, so the whole
// workgroup has reached it)
__device__ __forceinline__ void ps_tick(const PsLds &s, int k)
{
    if (threadIdx.x == 0 && s.st->timing) {
        const unsigned long long t = wall_clock64();
        s.st->tm[k] += t - s.st->t0;
        s.st->t0 = t;
    }
}

// the wave index as a uniform (scalar) value
__device__ __forceinline__ int ps_wave() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
// a double equal in every lane, as a uniform value
__device__ __forceinline__ double ps_unid(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// a double of a wave-uniform lane
__device__ __forceinline__ double ps_rl(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
// the same, moved into vector registers at once (values that stay live
// across a barrier would otherwise crowd the scalar registers)
__device__ __forceinline__ double ps_rlv(double v, int lane)
{
    double r = ps_rl(v, lane);
    asm volatile("" : "+v"(r));
    return r;
}

// a 0/1 factor for a wave-uniform condition (a scalar value, so that a
// row of a register array is chosen by arithmetic, not by an index the
// compiler could turn into a memory access)
__device__ __forceinline__ double ps_ind(bool cnd) { return cnd ? 1.0 : 0.0; }

// C = A B over the 16 x 16 tiles of an n x n product (at most 4 x 4 tiles,
// two per wave), FP64 MFMA 16x16x4.  fa(i, k) and fb(k, j) give operand
// entries, zero outside n x n; fo(i, j, c) stores an entry with i, j < n.
// inplace: every tile is computed before any is stored, so the output may
// overwrite an input.  Ends with a barrier.
template <class FA, class FB, class FO>
__device__ __forceinline__ void ps_gemm(int n, FA fa, FB fb, FO fo, bool inplace)
{
    const int lane = threadIdx.x & 63, wave = ps_wave();
    const int lc = lane & 15, lr = lane >> 4;
    const int nt = (n + 15) >> 4, K = (n + 3) & ~3;
    f64x4 acc[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
        const int tile = wave + 8 * t;
        if (tile < nt * nt) {
            const int i = (tile % nt) * 16 + lc, j = (tile / nt) * 16 + lc;
#pragma unroll 4
            for (int k0 = 0; k0 < K; k0 += 4) acc[t] = Mfma<double>::mma(fa(i, k0 + lr), fb(k0 + lr, j), acc[t]);
        }
    }
    if (inplace) __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int tile = wave + 8 * t;
        if (tile < nt * nt) {
            const int i0 = (tile % nt) * 16, j = (tile / nt) * 16 + lc;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + lr + 4 * v;
                if (i < n && j < n) fo(i, j, acc[t][v]);
            }
        }
    }
    __syncthreads();
}

// The same with two products summed by fo(i, j, c1, c2); the output must
// not alias an input.
template <class FA1, class FB1, class FA2, class FB2, class FO>
__device__ __forceinline__ void ps_gemm2(int n, FA1 fa1, FB1 fb1, FA2 fa2, FB2 fb2, FO fo)
{
    const int lane = threadIdx.x & 63, wave = ps_wave();
    const int lc = lane & 15, lr = lane >> 4;
    const int nt = (n + 15) >> 4, K = (n + 3) & ~3;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int tile = wave + 8 * t;
        if (tile < nt * nt) {
            const int i0 = (tile % nt) * 16, j0 = (tile / nt) * 16;
            f64x4 a1 = f64x4{0.0, 0.0, 0.0, 0.0}, a2 = a1;
#pragma unroll 4
            for (int k0 = 0; k0 < K; k0 += 4) {
                a1 = Mfma<double>::mma(fa1(i0 + lc, k0 + lr), fb1(k0 + lr, j0 + lc), a1);
                a2 = Mfma<double>::mma(fa2(i0 + lc, k0 + lr), fb2(k0 + lr, j0 + lc), a2);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + lr + 4 * v, j = j0 + lc;
                if (i < n && j < n) fo(i, j, a1[v], a2[v]);
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------- copies
// LDS matrices have leading dimension PS_LD; global ones n.
__device__ __forceinline__ void ps_load(double *S, const double *g, int n)
{
    for (int idx = threadIdx.x; idx < n * n; idx += PS_THREADS) S[idx % n + (idx / n) * PS_LD] = g[idx];
}
__device__ __forceinline__ void ps_store(double *g, const double *S, int n)
{
    for (int idx = threadIdx.x; idx < n * n; idx += PS_THREADS) g[idx] = S[idx % n + (idx / n) * PS_LD];
}
__device__ __forceinline__ void ps_identity(double *S, int n)
{
    for (int idx = threadIdx.x; idx < n * n; idx += PS_THREADS) S[idx % n + (idx / n) * PS_LD] = (idx % n == idx / n) ? 1.0 : 0.0;
}

// ------------------------------------------------- register-resident QR
// A matrix in registers: lane c holds column c, wave w rows 8w .. 8w+7 in
// a[0..7]; entries outside n x n are zero.

// Load an LDS matrix into that layout, column c taken from column col[c]
// of S (col = null: column c).
__device__ __forceinline__ void ps_to_regs(double (&a)[8], const double *S, int n, const int *col)
{
    const int c = threadIdx.x & 63, w = ps_wave();
    const int cc = (c < n) ? (col ? col[c] : c) : 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int r = 8 * w + m;
        a[m] = (c < n && r < n) ? S[r + cc * PS_LD] : 0.0;
    }
}
__device__ __forceinline__ void ps_from_regs(double *S, const double (&a)[8], int n)
{
    const int c = threadIdx.x & 63, w = ps_wave();
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int r = 8 * w + m;
        if (c < n && r < n) S[r + c * PS_LD] = a[m];
    }
}

// Householder QR, column by column as LAPACK dgeqr2 with dlarfg's
// reflectors (beta = -sign(alpha) ||(alpha, x)||, tau = (beta - alpha) /
// beta, v = (1, x / (alpha - beta))).  On return a[] holds R on and above
// the diagonal and the reflectors below it, tau[k] is in LDS.  Returns the
// number of reflectors with tau != 0 (sgn det Q = (-1)^that); *bad is set
// (by thread 0) when a reflector is not finite.  The input must be
// equilibrated (entries far below 1e150, as the chain's inputs are after
// ps_pivot's power-of-two column scales, and the rebuild's bounded M).
//
// One barrier per column: before it every wave writes, for every column c,
// its partial of d_c = x^T a_c over its rows below the diagonal (x the part
// of column k below it, so d_k = ||x||^2), and the wave holding row k writes
// that row; after it every thread has alpha, ||x||^2, d_c and a_kc, forms
// the reflector and w_c = v^T a_c = a_kc + d_c / (alpha - beta), and
// updates its rows of column c.  The two buffers alternate between columns.
__device__ int ps_qr(double (&a)[8], int n, double *tau, double *pd, double *rk, double *Vs, int *bad)
{
    const int c = threadIdx.x & 63, w = ps_wave();
    int ntau = 0;
    for (int k = 0; k < n; k++) {
        double *pb = pd + (k & 1) * 512, *rb = rk + (k & 1) * 64;
        // rows 8w + m of this wave lie below the diagonal for m > rel: all of
        // them for rel < 0, none for rel > 6 (wave-uniform branches)
        const int rel = k - 8 * w;
        // x: this wave's rows of column k below the diagonal (from lane k,
        // kept in vector registers; rows past n hold 0); the partial dot
        // products with every column, two accumulators
        double x[8];
        double p0 = 0.0, p1 = 0.0;
        if (rel < 0) {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                x[m] = ps_rlv(a[m], k);
                if (m & 1) p1 += x[m] * a[m];
                else p0 += x[m] * a[m];
            }
        } else if (rel < 7) {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                x[m] = ps_rlv(a[m], k) * ps_ind(m > rel);
                if (m & 1) p1 += x[m] * a[m];
                else p0 += x[m] * a[m];
            }
        } else {
#pragma unroll
            for (int m = 0; m < 8; m++) x[m] = 0.0;
        }
        pb[w * 64 + c] = p0 + p1;
        if (rel >= 0 && rel < 8) {
            double rv = 0.0;
#pragma unroll
            for (int m = 0; m < 8; m++) rv += ps_ind(m == rel) * a[m];
            rb[c] = rv;
        }
        __syncthreads();
        double q[8], e[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            q[u] = pb[u * 64 + k];
            e[u] = pb[u * 64 + c];
        }
        const double xn2 = ps_unid(((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])));
        const double dc = ((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7]));
        const double alpha = ps_unid(rb[k]), akc = rb[c];
        double t = 0.0, sc = 0.0, beta = alpha;
        if (xn2 != 0.0) {
            // tau = (beta - alpha) / beta and sc = 1 / (alpha - beta) from one
            // reciprocal of (alpha - beta) beta
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            const double d = alpha - beta, inv = 1.0 / (d * beta);
            t = -(d * d) * inv;
            sc = beta * inv;
        }
        if (threadIdx.x == 0) {
            tau[k] = t;
            if (!isfinite(t) || !isfinite(sc) || !isfinite(xn2)) *bad = 1;
        }
        if (t != 0.0) ntau++;
        if (t != 0.0 && rel < 8) {
            // the trailing columns (lanes k < c < n) take the reflector:
            // a_c -= t v (v^T a_c), v = (1, sc x), v^T a_c = a_kc + sc d_c
            const bool upd = c > k && c < n;
            const double tw = upd ? t * (akc + sc * dc) : 0.0, tsw = tw * sc;
            if (rel < 0) {
#pragma unroll
                for (int m = 0; m < 8; m++) a[m] -= tsw * x[m];
            } else {
#pragma unroll
                for (int m = 0; m < 8; m++) a[m] -= tsw * x[m] + ps_ind(m == rel) * tw;
            }
        }
        // lane k: the reflector below the diagonal (its x scaled), beta on it
        if (c == k) {
#pragma unroll
            for (int m = 0; m < 8; m++)
                a[m] = a[m] * ((m > rel) ? sc : ps_ind(m < rel)) + ps_ind(m == rel) * beta;
        }
    }
    // the reflectors into Vs: column c is v_c, unit entry, zeros above, the
    // stored entries below (rows past n stay 0)
    if (c < n) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int r = 8 * w + m;
            if (r < n) Vs[r + c * PS_LD] = (r > c) ? a[m] : ((r == c) ? 1.0 : 0.0);
        }
    }
    __syncthreads();
    return ntau;
}

// x <- H_k x for k in order (forward: H_0 first, i.e. x <- Q^T x; backward:
// H_{n-1} first, i.e. x <- Q x), the reflectors in the LDS matrix Vs (as
// ps_qr leaves them: column k is v_k with its unit entry and zeros above,
// rows past n zero) and tau in LDS.  One barrier per reflector; the partial-sum buffer
// alternates so that a step's reads never meet the next step's writes.
__device__ void ps_apply_h(const double *Vs, double (&x)[8], int n, const double *tau, double *pd, bool forward)
{
    const int c = threadIdx.x & 63, w = ps_wave();
    int ph = 0;
    // reflector k's tau and this wave's eight rows of it (zero past n), the
    // next one's read while the current one is applied
    int k = forward ? 0 : n - 1;
    double t = ps_unid(tau[k]);
    double v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = Vs[8 * w + m + k * PS_LD];
    for (int s = 0; s < n; s++) {
        const int kn = forward ? k + 1 : k - 1;
        const bool more = s + 1 < n;
        double tn = 0.0, vn[8];
        if (more) {
            tn = ps_unid(tau[kn]);
#pragma unroll
            for (int m = 0; m < 8; m++) vn[m] = Vs[8 * w + m + kn * PS_LD];
        }
        if (t != 0.0) {
            double p0 = 0.0, p1 = 0.0;
#pragma unroll
            for (int m = 0; m < 8; m++) {
                if (m & 1) p1 += v[m] * x[m];
                else p0 += v[m] * x[m];
            }
            double *pb = pd + ph * 512;
            ph ^= 1;
            pb[w * 64 + c] = p0 + p1;
            __syncthreads();
            double e[8];
#pragma unroll
            for (int u = 0; u < 8; u++) e[u] = pb[u * 64 + c];
            const double tw = t * (((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7])));
#pragma unroll
            for (int m = 0; m < 8; m++) x[m] -= tw * v[m];
        }
        if (more) {
            t = tn;
#pragma unroll
            for (int m = 0; m < 8; m++) v[m] = vn[m];
        }
        k = kn;
    }
    __syncthreads();
}

// Q = H_0 ... H_{n-1} applied to the identity (dorg2r)
__device__ __forceinline__ void ps_formq(const double *Vs, double (&q)[8], int n, const double *tau, double *pd)
{
    const int c = threadIdx.x & 63, w = ps_wave();
#pragma unroll
    for (int m = 0; m < 8; m++) q[m] = (8 * w + m == c && c < n) ? 1.0 : 0.0;
    ps_apply_h(Vs, q, n, tau, pd, false);
}

// x <- Rn^-1 x, Rn = D^-1 R unit upper triangular (R on and above the
// diagonal of a[], D = diag R): the 8-row block of each wave is solved in
// its registers (the entries of R come from lane k by readlane), its final
// rows go to the blocks above through LDS (two alternating buffers).
__device__ void ps_trsm(const double (&a)[8], double (&x)[8], int n, double *rowbuf)
{
    const int c = threadIdx.x & 63, w = ps_wave();
    double di[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int r = 8 * w + m;
        const double d = ps_rl(a[m], r < 64 ? r : 63);
        di[m] = r < n ? 1.0 / d : 0.0;
    }
    const int nb = (n + 7) >> 3;
    for (int b = nb - 1; b >= 0; b--) {
        double *rb = rowbuf + (b & 1) * 512;
        if (w == b) {
#pragma unroll
            for (int m = 7; m >= 0; m--) {
                const int k = 8 * b + m;
                if (k < n) {
#pragma unroll
                    for (int m2 = 0; m2 < m; m2++) x[m2] -= ps_rl(a[m2], k) * di[m2] * x[m];
                }
            }
#pragma unroll
            for (int m = 0; m < 8; m++) rb[m * 64 + c] = x[m];
        }
        __syncthreads();
        if (w < b) {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                const int k = 8 * b + m;
                if (k < n) {
                    const double xk = rb[m * 64 + c];
#pragma unroll
                    for (int m2 = 0; m2 < 8; m2++) x[m2] -= ps_rl(a[m2], k) * di[m2] * xk;
                }
            }
        }
    }
    __syncthreads();
}

// Rn = D^-1 R (unit diagonal, zero below) from a[] into an LDS matrix, D
// into Dv, undoing the column scales of the QR input: column c of the
// factored matrix was column col[c] of the input times sc[col[c]] (powers of
// two), so D_c = R'_cc / sc[col[c]] and Rn_rc = Rn'_rc sc[col[r]] /
// sc[col[c]] (sc = null: no scales).  Sets *bad = 2 for a zero or
// non-finite D.
__device__ __forceinline__ void ps_store_rn(double *S, double *Dv, const double (&a)[8], int n, int *bad,
                                            const double *sc, const int *col)
{
    const int c = threadIdx.x & 63, w = ps_wave();
    const double scc = (sc && c < n) ? sc[col[c]] : 1.0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int r = 8 * w + m;
        const double d = ps_rl(a[m], r);
        const double scr = (sc && r < n) ? sc[col[r]] : 1.0;
        if (r < n && c < n) S[r + c * PS_LD] = (r < c) ? (a[m] / d) * (scr / scc) : (r == c ? 1.0 : 0.0);
        if (r < n && c == r) {
            const double dt = d / scc;
            Dv[r] = dt;
            if (!isfinite(dt) || dt == 0.0) *bad = 2;
        }
    }
}

// Column pre-pivoting and equilibration of the LDS matrix S: perm = its
// columns in order of decreasing 2-norm, ties by index (the order of
// k_col_keys / k_col_sort; pivot = 0: the identity), and sc[c] = 2^-e with
// e the binary exponent of max_r |S_rc| (1 for a zero column).  Norms are
// compared as logarithms, each wave's partial sum of squares taken on its
// entries scaled by a power of two, so nothing overflows and nothing is
// divided.  A non-finite entry sets *bad = 1.
__device__ void ps_pivot(const double *S, int n, int *perm, double *keys, double *sc, double *pd, int pivot,
                         int *bad)
{
    const int c = threadIdx.x & 63, w = ps_wave();
    double x[8], mx = 0.0;
    bool fin = true;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int r = 8 * w + m;
        x[m] = (r < n && c < n) ? S[r + c * PS_LD] : 0.0;
        mx = fmax(mx, fabs(x[m]));
        fin = fin && isfinite(x[m]);
    }
    const int e = (mx > 0.0 && isfinite(mx)) ? ilogb(mx) : 0;
    double s2 = 0.0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const double t = ldexp(x[m], -e);
        s2 += t * t;
    }
    pd[w * 64 + c] = mx > 0.0 ? (double)e : -1e300;
    pd[512 + w * 64 + c] = s2;
    if (!fin) *bad = 1;
    __syncthreads();
    if (w == 0 && c < n) {
        double emax = -1e300;
#pragma unroll
        for (int u = 0; u < 8; u++) emax = fmax(emax, pd[u * 64 + c]);
        double key = -INFINITY, scale = 1.0;
        if (emax > -1e299) {
            const int em = (int)emax;
            double S2 = 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double eu = pd[u * 64 + c];
                if (eu > -1e299) S2 += ldexp(pd[512 + u * 64 + c], 2 * ((int)eu - em));
            }
            key = 0.5 * log(S2) + em * 0.69314718055994530942;
            if (isnan(key)) key = -INFINITY;
            scale = ldexp(1.0, -em);
        }
        keys[c] = key;
        sc[c] = scale;
    }
    __syncthreads();
    if (w == 0 && c < n) {
        if (!pivot) {
            perm[c] = c;
        } else {
            const double key = keys[c];
            int rank = 0;
            for (int k = 0; k < n; k++) {
                const double kk = keys[k];
                rank += (kk > key || (kk == key && k < c)) ? 1 : 0;
            }
            perm[rank] = c;
        }
    }
    __syncthreads();
}

// One UDV update of the chain from an LDS input M (n x n): M P = Q R with
// the pre-pivoting order P (QR of the equilibrated M P); Q into Uout (LDS,
// may be M), D into Dv, Rn = D^-1 R into Rs (LDS, which holds the
// reflectors meanwhile), the order into perm.  Returns sgn det Q; *bad as in ps_qr / ps_store_rn.
__device__ int ps_update_qr(const double *M, double *Uout, double *Rs, double *Dv, int n, const PsLds &s,
                            int pivot, int *bad)
{
    ps_pivot(M, n, s.perm, s.keys, s.scl, s.Cw, pivot, bad);
    double a[8], q[8];
    {
        const int c = threadIdx.x & 63, w = ps_wave();
        const int cc = c < n ? s.perm[c] : 0;
        const double f = c < n ? s.scl[cc] : 0.0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int r = 8 * w + m;
            a[m] = (c < n && r < n) ? M[r + cc * PS_LD] * f : 0.0;
        }
    }
    // the reflectors go to Rs until Rn replaces them
    const int ntau = ps_qr(a, n, s.tau, s.Cw, s.rk, Rs, bad);
    ps_formq(Rs, q, n, s.tau, s.Cw);
    ps_from_regs(Uout, q, n);
    __syncthreads();
    ps_store_rn(Rs, Dv, a, n, bad, s.scl, s.perm);
    __syncthreads();
    return (ntau & 1) ? -1 : 1;
}

// Db/Ds split without a log offset (d_split, src/linalg.c:966-986):
// Db = D where |D| > 1 (else 1), Ds = D / Db; bi = 1 / Db
__device__ __forceinline__ void ps_split(double d, double &bi, double &sm)
{
    const bool big = fabs(d) > 1.0;
    bi = big ? 1.0 / d : 1.0;
    sm = big ? 1.0 : d;
}

// the dynamic LDS of k_sweep_persist, seen by its phase functions
extern __shared__ __align__(16) unsigned char ps_lds_raw[];

// LDS matrix k (0..3) by arithmetic on the LDS base, so that the compiler
// keeps the LDS address space (a pointer picked out of an array by a
// run-time index would turn every access into a flat memory instruction)
__device__ __forceinline__ double *ps_slot(int k) { return (double *)ps_lds_raw + k * PS_MAT; }

// a pointer or an int known to be equal in every lane, as the compiler's
// uniform value (the phase functions receive their arguments in vector
// registers)
__device__ __forceinline__ int ps_uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T> __device__ __forceinline__ T *ps_uni(T *p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

// The suffix stack S[M-1] .. S[1] of replica r, kept transposed:
// S[j]^T = Bblock_j^T S[j+1]^T as a left update, M = (Bblock_j^T U) diag(D),
// M P = Q R, U' = Q, D' = diag R, T' = (D'^-1 R) P^T T (the batched path's
// stage_rmul); S[M] is the identity.  Slots 1, 2 carry U, T of S[j+1],
// slot 0 the block product, slot 3 the QR input.  Returns the flag.
__device__ __attribute__((noinline)) int ps_suffix(int r_)
{
    const int r = ps_uni(r_);
    const PsLds s = ps_carve(ps_lds_raw);
    const PersistModel *pm = s.pm;
    const int tid = threadIdx.x;
    const int n = ps_uni(pm->n), M = ps_uni(pm->M), R = ps_uni(pm->nrep);
    const long long nn = (long long)n * n;
    double *SUr = pm->SU + (long long)r * (M + 1) * nn, *STr = pm->ST + (long long)r * (M + 1) * nn;
    double *SDr = pm->SD + (long long)r * (M + 1) * n;
    int *Sqr = pm->Sq + (long long)r * (M + 1);
    double *Us = ps_slot(1), *Ts = ps_slot(2), *Bs = ps_slot(0), *Ms = ps_slot(3);
    ps_identity(Us, n);
    ps_identity(Ts, n);
    if (tid < n) s.DR[tid] = 1.0;
    for (int j = M - 1; j >= 1; j--) {
        ps_load(Bs, pm->BB + ((long long)j * R + r) * nn, n);
        __syncthreads();
        ps_gemm(
            n, [=](int i, int k) { return Bs[k + i * PS_LD]; }, [=](int k, int c) { return Us[k + c * PS_LD]; },
            [=](int i, int c, double v) { Ms[i + c * PS_LD] = v * s.DR[c]; }, false);
        const int qs = ps_update_qr(Ms, Us, Bs, s.DR, n, s, pm->pivot, s.flag);
        ps_gemm(
            n, [=](int i, int k) { return Bs[i + k * PS_LD]; },
            [=](int k, int c) { return k < n ? Ts[s.perm[k] + c * PS_LD] : 0.0; },
            [=](int i, int c, double v) { Ts[i + c * PS_LD] = v; }, true);
        ps_store(SUr + (long long)j * nn, Us, n);
        ps_store(STr + (long long)j * nn, Ts, n);
        if (tid < n) SDr[(long long)j * n + tid] = s.DR[tid];
        if (tid == 0) Sqr[j] = qs;
        __syncthreads();
        if (*s.flag) break;
    }
    const int f = *s.flag;
    __syncthreads();
    ps_tick(s, PT_SUFFIX);
    return f;
}

// Draws i0 .. i1-1 of the replica's stream (xoshiro256**, src/rng.c:23-37)
// into out, by one whole wave on wave-uniform values (scalar instructions),
// lane 0 storing; the state lives in the LDS sweep state.
__device__ __forceinline__ void ps_draws(PsState *st, double *out, int i0, int i1)
{
    uint64_t x[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t v = st->rng[q];
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
        x[q] = ((uint64_t)hi << 32) | lo;
    }
    for (int i = i0; i < i1; i++) {
        const double d = xoshiro_double(x);
        if ((threadIdx.x & 63) == 0) out[i] = d;
    }
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 4; q++) st->rng[q] = x[q];
}

// Slice l of replica r: Metropolis over its n sites in windows of W (G in
// slot st->g), the slice's spins and B diagonal written back, then the
// block product P <- B_l P in slot st->p and the wrap of G.
template <int W>
__device__ __forceinline__ void ps_slice(const PersistModel *pm, int r, int l)
{
    constexpr int RPL = W * W / 64;
    const PsLds s = ps_carve(ps_lds_raw);
    const int tid = threadIdx.x, lane = tid & 63, wave = ps_wave();
    const int lc = lane & 15, lr = lane >> 4;
    const int n = ps_uni(pm->n), L = ps_uni(pm->L), stab = ps_uni(pm->stab);
    double *Gs = ps_slot(ps_uni(s.st->g)), *Ks = ps_slot(ps_uni(s.st->k)), *Kis = ps_slot(ps_uni(s.st->ki)),
           *Ps = ps_slot(ps_uni(s.st->p));
    signed char *fld = pm->field + ((long long)r * L + l) * n;
    // the site constants as uniform values, read once
    const double nu0 = pm->nu[0], nu1 = pm->nu[1], nd0 = pm->nd[0], nd1 = pm->nd[1];
    const double e0 = pm->e[0], e1 = pm->e[1], ei0 = pm->einv[0], ei1 = pm->einv[1];
    // this slice's spins were fetched during the previous slice (slice 0:
    // here); the next slice's are fetched now and stored at the end
    signed char *fs = s.fs + (l & 1) * PS_NP, *fsn = s.fs + ((l & 1) ^ 1) * PS_NP;
    if (l == 0 && tid < n) fs[tid] = fld[tid];
    signed char fnext = 0;
    if (l + 1 < L && tid < n) fnext = fld[n + tid];
    if (tid < n) s.flip[tid] = 0;
    // The slice's n draws: taken ahead by wave 1 during the previous slice's
    // proposals, except for the first slice of a block, whose draws are
    // taken here (so a sweep that fails at a rebuild has taken no draw of
    // the slices after it, as on the batched path).
    double *dcur = s.draws + (l & 1) * PS_NP, *dnext = s.draws + ((l & 1) ^ 1) * PS_NP;
    if (l % stab == 0 && wave == 1) ps_draws(s.st, dcur, 0, n);
    const bool ahead = l + 1 < L && (l + 1) % stab != 0;
    __syncthreads();
    int negflips = 0, naccept = 0;
    for (int i0 = 0; i0 < n; i0 += W) {
        const int wn = min(W, n - i0);
        // the mask word alternates between windows: a window without
        // accepts has no barrier after its read
        unsigned *mword = s.mask + ((i0 / W) & 1);
        if (wave == 0) {
            const int c = lane % W, h = lane / W;
            double bw[RPL];
#pragma unroll
            for (int k = 0; k < RPL; k++) {
                const int rr = h * RPL + k;
                bw[k] = (rr < wn && c < wn) ? Gs[(i0 + rr) + (i0 + c) * PS_LD] : 0.0;
            }
            double dr_l = 0.0, nu_l = 0.0, nd_l = 0.0;
            if (lane < wn) {
                const bool sp = fs[i0 + lane] > 0;
                dr_l = dcur[i0 + lane];
                nu_l = sp ? nu1 : nu0;
                nd_l = sp ? nd1 : nd0;
            }
            double *rlog = pm->rlog;
            unsigned mask = 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
                if (w < wn) {
                    const int i = i0 + w;
                    const double gii = ps_rl(bw[w % RPL], (w / RPL) * W + w);
                    const double Nu = ps_rl(nu_l, w), Nd = ps_rl(nd_l, w);
                    const double Ru = 1.0 + (1.0 - gii) * Nu;
                    const double Rd = 1.0 + gii * Nd;
                    const double Rt = Ru * Rd;
                    if (rlog && lane == 0) rlog[(long long)r * n + i] = Rt;
                    // N / R_up of an accept, started before the decision
                    const double fac = Nu / Ru;
                    if (ps_rl(dr_l, w) < fabs(Rt)) {
                        // column w of the block to every lane of each 16-lane
                        // row (DPP row broadcast) gives c; row w, held by
                        // 16-lane row w / RPL, to all four rows (row swaps)
                        // gives v; G~ -= c v^T in registers.  The window
                        // pairs also go to LDS for the pair formation after
                        // the next barrier.
                        double cw[RPL];
#pragma unroll
                        for (int k = 0; k < RPL; k++)
                            cw[k] = fac * ((h * RPL + k == w ? 1.0 : 0.0) - row_bcast(bw[k], w));
                        const double vc = row_to_all(bw[w % RPL], w / RPL);
                        if (c == w) {
#pragma unroll
                            for (int k = 0; k < RPL; k++) s.cwin[w * W + h * RPL + k] = cw[k];
                        }
                        if (h == w / RPL) s.vwin[w * W + c] = vc;
                        if (lane == 0) {
                            s.facw[w] = fac;
                            s.flip[i] = 1;
                        }
                        mask |= 1u << w;
                        naccept++;
                        if (Rt < 0.0) negflips ^= 1;
#pragma unroll
                        for (int k = 0; k < RPL; k++) bw[k] -= cw[k] * vc;
                    }
                }
            }
            if (lane == 0) *mword = mask;
        } else if (wave == 1 && ahead) {
            ps_draws(s.st, dnext, i0, i0 + wn);
        }
        __syncthreads();
        const unsigned mask = __builtin_amdgcn_readfirstlane(*mword);
        const int p = __builtin_popcount(mask);
        if (p == 0) continue;
        // the window's pairs over all sites: thread j < n forms c_q(j) from
        // column i0 + w of G, thread n + j forms v_q(j) from row i0 + w, the
        // accepts in order, each final pair entry removed at once from the
        // later window sites (right-looking: the coefficients of one accept
        // are one row of the window pairs, read together)
        {
            const bool vside = tid >= n;
            const int jj = vside ? tid - n : tid;
            if (jj < n) {
                const double *coef = vside ? s.cwin : s.vwin;
                double *dst = vside ? s.Vw : s.Cw;
                double gc[W];
#pragma unroll
                for (int w = 0; w < W; w++) {
                    const int ww = min(w, wn - 1);
                    const double x = vside ? Gs[(i0 + ww) + jj * PS_LD] : Gs[jj + (i0 + ww) * PS_LD];
                    gc[w] = (w < wn) ? x : 0.0;
                }
#pragma unroll
                for (int u = 0; u < W; u++) {
                    if ((mask >> u) & 1u) {
                        gc[u] = vside ? gc[u] : s.facw[u] * ((jj == i0 + u ? 1.0 : 0.0) - gc[u]);
                        const int q = __builtin_popcount(mask & ((1u << u) - 1u));
                        dst[q * PS_NP + jj] = gc[u];
#pragma unroll
                        for (int w = u + 1; w < W; w++) gc[w] -= gc[u] * coef[u * W + w];
                    }
                }
            }
        }
        __syncthreads();
        // G -= C^T V over the window's p pairs
        {
            const int nt = (n + 15) >> 4;
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int tile = wave + 8 * t;
                if (tile < nt * nt) {
                    const int ti = (tile % nt) * 16, tj = (tile / nt) * 16;
                    f64x4 acc;
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[v] = Gs[(ti + lr + 4 * v) + (tj + lc) * PS_LD];
                    for (int q0 = 0; q0 < p; q0 += 4) {
                        const int q = q0 + lr;
                        const double ca = q < p ? s.Cw[q * PS_NP + ti + lc] : 0.0;
                        const double vb = q < p ? s.Vw[q * PS_NP + tj + lc] : 0.0;
                        acc = Mfma<double>::mma(-ca, vb, acc);
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int i = ti + lr + 4 * v, j = tj + lc;
                        if (i < n && j < n) Gs[i + j * PS_LD] = acc[v];
                    }
                }
            }
        }
        __syncthreads();
    }
    ps_tick(s, PT_METRO);
    // the slice's fields are final: spins, the diagonal of B_l
    if (tid < n) {
        const signed char sp = s.flip[tid] ? (signed char)-fs[tid] : fs[tid];
        fld[tid] = sp;
        if (l + 1 < L) fsn[tid] = fnext;
        const double dv = sp > 0 ? e1 : e0, di = sp > 0 ? ei1 : ei0;
        const long long o = ((long long)r * L + l) * n + tid;
        pm->dvec[o] = dv;
        pm->dinv[o] = di;
        s.dl[tid] = dv;
        s.dli[tid] = di;
    }
    if (tid == 0) {
        s.st->nacc += (unsigned long long)naccept;
        s.st->natt += (unsigned long long)n;
        if (negflips) s.st->sgn = -s.st->sgn;
    }
    __syncthreads();
    // block product P <- B_l P (B_lb = expK diag(d_lb) at a block start)
    if (l % stab == 0) {
        for (int idx = tid; idx < n * n; idx += PS_THREADS) {
            const int i = idx % n, k = idx / n;
            Ps[i + k * PS_LD] = Ks[i + k * PS_LD] * s.dl[k];
        }
        __syncthreads();
    } else {
        ps_gemm(
            n, [=](int i, int k) { return Ks[i + k * PS_LD]; },
            [=](int k, int c) { return s.dl[k] * Ps[k + c * PS_LD]; },
            [=](int i, int c, double v) { Ps[i + c * PS_LD] = v; }, true);
    }
    ps_tick(s, PT_PRODUCT);
    // wrap G <- expK diag(d) G diag(1/d) expKinv
    ps_gemm(
        n, [=](int i, int k) { return Ks[i + k * PS_LD]; },
        [=](int k, int c) { return s.dl[k] * Gs[k + c * PS_LD] * s.dli[c]; },
        [=](int i, int c, double v) { Gs[i + c * PS_LD] = v; }, true);
    ps_gemm(
        n, [=](int i, int k) { return Gs[i + k * PS_LD]; }, [=](int k, int c) { return Kis[k + c * PS_LD]; },
        [=](int i, int c, double v) { Gs[i + c * PS_LD] = v; }, true);
    ps_tick(s, PT_WRAP);
}

// The slices l0 .. l1-1 of one block (one call per block: the slice code
// lives in this function's registers, and its calls are few)
template <int W>
__device__ __attribute__((noinline)) void ps_block(int r_, int l0_, int l1_)
{
    const PersistModel *pm = ps_carve(ps_lds_raw).pm;
    const int r = ps_uni(r_), l0 = ps_uni(l0_), l1 = ps_uni(l1_);
    for (int l = l0; l < l1; l++) ps_slice<W>(pm, r, l);
}

// The block boundary after slice l (tau = l + 1): the block product goes to
// the cache, the left UDV is updated (lmul with a pre-pivoted QR, the left
// UDV being the identity before the first block), and G is rebuilt from it
// and S[tau / stab] (the identity at tau = L) by the two-sided inverse,
//   M = Dlb^-1 (Ul^T Ur) Drb^-1 + Dls (Tl Tr^T) Drs = Q_M D_M Rn_M,
//   G = Ur Drb^-1 Rn_M^-1 D_M^-1 Q_M^T Dlb^-1 Ul^T,
// then compared with the wrapped G (drift sample, guard, finiteness), the
// sign reset to +1 and G replaced.  Returns the flag; on a QR failure G is
// the wrapped one.
__device__ __attribute__((noinline)) int ps_boundary(int r_, int l_)
{
    const int r = ps_uni(r_), l = ps_uni(l_);
    const PsLds s = ps_carve(ps_lds_raw);
    const PersistModel *pm = s.pm;
    const int tid = threadIdx.x, lane = tid & 63, wave = ps_wave();
    const int n = ps_uni(pm->n), L = ps_uni(pm->L), stab = ps_uni(pm->stab), M = ps_uni(pm->M),
              R = ps_uni(pm->nrep);
    const long long nn = (long long)n * n;
    const int gi = ps_uni(s.st->g), ki = ps_uni(s.st->k), kii = ps_uni(s.st->ki), pi = ps_uni(s.st->p);
    double *Gs = ps_slot(gi), *Us = ps_slot(ki), *Ts = ps_slot(kii), *Ps = ps_slot(pi);
    double *LUr = pm->LU + r * nn, *LTr = pm->LT + r * nn, *LDr = pm->LD + (long long)r * n;
    const int tau = l + 1, jb = l / stab;
    ps_store(pm->BB + ((long long)jb * R + r) * nn, Ps, n);
    if (jb == 0) {
        ps_identity(Us, n);
        ps_identity(Ts, n);
        if (tid < n) s.DL[tid] = 1.0;
    } else {
        ps_load(Us, LUr, n);
        ps_load(Ts, LTr, n);
        if (tid < n) s.DL[tid] = LDr[tid];
    }
    __syncthreads();
    // (P U) diag(D) in place in Us; its QR: Q back into Us, Rn into Ps
    ps_gemm(
        n, [=](int i, int k) { return Ps[i + k * PS_LD]; }, [=](int k, int c) { return Us[k + c * PS_LD]; },
        [=](int i, int c, double v) { Us[i + c * PS_LD] = v * s.DL[c]; }, true);
    const int qs = ps_update_qr(Us, Us, Ps, s.DL, n, s, pm->pivot, s.flag);
    if (tid == 0) pm->Lq[r] = qs;
    // T' = Rn P^T T, in place in Ts
    ps_gemm(
        n, [=](int i, int k) { return Ps[i + k * PS_LD]; },
        [=](int k, int c) { return k < n ? Ts[s.perm[k] + c * PS_LD] : 0.0; },
        [=](int i, int c, double v) { Ts[i + c * PS_LD] = v; }, true);
    ps_store(LUr, Us, n);
    ps_store(LTr, Ts, n);
    if (tid < n) LDr[tid] = s.DL[tid];
    __syncthreads();
    ps_tick(s, PT_LMUL);
    if (*s.flag) {
        const int f = *s.flag;
        __syncthreads();
        return f;
    }
    // G = (1 + L R)^-1 with R = S[j]
    const int j = (tau == L) ? M : tau / stab;
    const double *Ur = pm->SU + ((long long)r * (M + 1) + j) * nn, *Tr = pm->ST + ((long long)r * (M + 1) + j) * nn;
    const double *Dr = pm->SD + ((long long)r * (M + 1) + j) * n;
    if (tid < n) {
        double bi, sm;
        ps_split(s.DL[tid], bi, sm);
        s.lbi[tid] = bi;
        s.ls[tid] = sm;
        ps_split(Dr[tid], bi, sm);
        s.rbi[tid] = bi;
        s.rs[tid] = sm;
    }
    __syncthreads();
    double *Ms = Ps;
    ps_gemm2(
        n, [=](int i, int k) { return Us[k + i * PS_LD]; },
        [=](int k, int c) { return (k < n && c < n) ? Ur[k + (long long)c * n] : 0.0; },
        [=](int i, int k) { return Ts[i + k * PS_LD]; },
        [=](int k, int c) { return (k < n && c < n) ? Tr[c + (long long)k * n] : 0.0; },
        [=](int i, int c, double v1, double v2) {
            Ms[i + c * PS_LD] = s.lbi[i] * v1 * s.rbi[c] + s.ls[i] * v2 * s.rs[c];
        });
    {
        double a[8], x[8];
        ps_to_regs(a, Ms, n, nullptr);
        __syncthreads();
        // the reflectors replace M in its slot
        (void)ps_qr(a, n, s.tau, s.Cw, s.rk, Ms, s.flag);
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int rr = 8 * wave + m;
            const double d = ps_rl(a[m], rr);
            if (rr < n && lane == rr) {
                s.DM[rr] = d;
                if (!isfinite(d) || d == 0.0) *s.flag = 2;
            }
        }
        __syncthreads();
        ps_tick(s, PT_REBUILD_QR);
        // X = Dlb^-1 Ul^T, then Q_M^T X, D_M^-1 and Rn_M^-1 on it
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int rr = 8 * wave + m;
            x[m] = (rr < n && lane < n) ? s.lbi[rr] * Us[lane + rr * PS_LD] : 0.0;
        }
        ps_apply_h(Ms, x, n, s.tau, s.Cw, true);
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int rr = 8 * wave + m;
            x[m] = rr < n ? x[m] / s.DM[rr] : 0.0;
        }
        ps_trsm(a, x, n, s.Vw);
        ps_from_regs(Ts, x, n);
        __syncthreads();
        ps_tick(s, PT_REBUILD_SOLVE);
    }
    if (*s.flag) {
        const int f = *s.flag;
        __syncthreads();
        return f;
    }
    // G_new = Ur Drb^-1 W into Us
    ps_gemm(
        n, [=](int i, int k) { return (i < n && k < n) ? Ur[i + (long long)k * n] * s.rbi[k] : 0.0; },
        [=](int k, int c) { return Ts[k + c * PS_LD]; }, [=](int i, int c, double v) { Us[i + c * PS_LD] = v; },
        false);
    // drift max|G_wrap - G_new| and finiteness of both
    double mxd = 0.0;
    int bad = 0;
    for (int idx = tid; idx < n * n; idx += PS_THREADS) {
        const int i = idx % n, c = idx / n;
        const double o = Gs[i + c * PS_LD], v = Us[i + c * PS_LD];
        if (!isfinite(v) || !isfinite(o)) bad = 1;
        mxd = fmax(mxd, isnan(o - v) ? INFINITY : fabs(o - v));
    }
    mxd = wave_max(mxd);
    const int wbad = __builtin_amdgcn_readfirstlane(__ballot(bad) != 0ull ? 1 : 0);
    if (lane == 0) {
        s.red[wave] = mxd;
        s.red[8 + wave] = wbad;
    }
    __syncthreads();
    if (tid == 0) {
        double mm = 0.0;
        int bb = 0;
        for (int u = 0; u < 8; u++) {
            mm = fmax(mm, s.red[u]);
            bb |= s.red[8 + u] != 0.0;
        }
        s.st->sgn = 1.0;
        if (bb) *s.flag = 3;
        else if (pm->guard > 0.0 && !(mm <= pm->guard)) *s.flag = 4;
        if (pm->record && !bb) {
            if (pm->dcnt[r] == 0 || mm > pm->dmax[r]) {
                pm->dmax[r] = mm;
                pm->dtaumax[r] = (tau == L) ? 0 : tau;
                pm->dsweepmax[r] = *pm->sweeps;
            }
            pm->dsum[r] += mm;
            pm->dcnt[r] += 1;
        }
        // the new G takes U's slot, the old G's slot and T's slot take expK
        // and expK^-1 again
        s.st->g = ki;
        s.st->k = gi;
        s.st->ki = kii;
    }
    __syncthreads();
    if (tau != L) {
        ps_load(Gs, pm->expK, n);
        ps_load(Ts, pm->expKinv, n);
    }
    const int f = *s.flag;
    __syncthreads();
    ps_tick(s, PT_REBUILD_G);
    return f;
}

template <int W>
// timers: null, or PS_NTIMER counters (100 MHz ticks) that workgroup 0
// accumulates its phases into
__global__ __launch_bounds__(PS_THREADS) void k_sweep_persist(const PersistModel *models, int nmodels,
                                                              unsigned long long *timers)
{
    static_assert(W == PS_W, "window of 16 sites");
    int mi = 0;
    while (mi + 1 < nmodels && (int)blockIdx.x >= models[mi + 1].first) mi++;
    const PersistModel *pm = models + mi;
    const int r = blockIdx.x - pm->first;
    if (r < 0 || r >= pm->nrep) return;
    if (pm->status[r]) return;
    const PsLds s = ps_carve(ps_lds_raw);
    const int tid = threadIdx.x;
    const int n = pm->n, L = pm->L, stab = pm->stab;
    const long long nn = (long long)n * n;
    // all of the LDS starts at zero, so the entries of matrices and vectors
    // outside n x n (never written) read as 0
    {
        unsigned *z = (unsigned *)ps_lds_raw;
        const int words = (int)(ps_lds_bytes() / 4);
        for (int idx = tid; idx < words; idx += PS_THREADS) z[idx] = 0u;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 0; q < 4; q++) s.st->rng[q] = pm->rng[4 * (long long)r + q];
        s.st->sgn = pm->sign[r];
        s.st->g = 0;
        s.st->k = 1;
        s.st->ki = 2;
        s.st->p = 3;
        s.st->timing = (timers && blockIdx.x == 0) ? 1 : 0;
        s.st->t0 = wall_clock64();
    }
    // the model's descriptor into LDS, where the phases read it
    {
        const unsigned long long *src = (const unsigned long long *)pm;
        unsigned long long *dst = (unsigned long long *)s.pm;
        for (int q = tid; q < (int)(sizeof(PersistModel) / 8); q += PS_THREADS) dst[q] = src[q];
    }
    __syncthreads();
    int f = ps_suffix(r);
    if (f) {
        if (tid == 0) {
            pm->status[r] = 1;
            pm->reason[r] = f == 1 ? 16 : 17;
        }
        return;
    }
    ps_load(ps_slot(0), pm->G + r * nn, n);
    ps_load(ps_slot(1), pm->expK, n);
    ps_load(ps_slot(2), pm->expKinv, n);
    __syncthreads();
    for (int l0 = 0; l0 < L; l0 += stab) {
        const int l1 = min(L, l0 + stab);
        ps_block<W>(r, l0, l1);
        f = ps_boundary(r, l1 - 1);
        if (f) break;
    }
    // f: 1, 2 a failed QR (G stays the wrapped one, the sign is not reset);
    // 3, 4 a rebuilt G that is not finite or drifted beyond the guard (kept,
    // as the batched copy keeps it)
    ps_store(pm->G + r * nn, ps_slot(ps_uni(s.st->g)), n);
    if (tid == 0) {
        if (f) {
            pm->status[r] = 1;
            pm->reason[r] = f == 1 ? 16 : f == 2 ? 17 : f == 3 ? 16 : 19;
        }
        for (int q = 0; q < 4; q++) pm->rng[4 * (long long)r + q] = s.st->rng[q];
        pm->acc[r] += s.st->nacc;
        pm->att[r] += s.st->natt;
        pm->sign[r] = s.st->sgn;
        if (s.st->timing)
            for (int k = 0; k < PS_NTIMER; k++) timers[k] += s.st->tm[k];
    }
}
