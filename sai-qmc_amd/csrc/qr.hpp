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
//   panel (QR_NB columns, register-resident, owned round-robin by the
//     waves): unblocked dgeqr2 with one wave reduction per norm and per
//     column update; the waves follow a published-reflector counter in LDS
//     rather than a barrier per column, and the owner of the next pivot
//     column publishes it first (lookahead);
//   T factor of the compact WY form (dlarft, forward, columnwise);
//   trailing update A <- (I - V T^T V^T) A and Q <- (I - V T V^T) Q on FP64
//     MFMA 16x16x4: each wave takes 16-column blocks, W = V^T A (QR_NB / 16
//     row tiles, K = panel height), W' = op(T) W, A -= V W'.
//
// k_qr_formq: explicit Q from the reflectors, one wave per 16-column block.
//
// k_trsm_unit_upper: X = U^-1 B for unit upper-triangular U (the
// reference inverts T with dtrtri, src/linalg.c:992-1004; here the
// triangular factor is applied by a solve), 32-column blocks of X per
// workgroup, back substitution over 32-row blocks with MFMA products.
#pragma once
#include "common.hpp"

#define QR_NB 32
#define QR_NT (QR_NB / 16) // 16-row tiles of a panel's reflectors
#define QR_THREADS 512
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
// 1 on the diagonal, stored values below, 0 above and outside the nb x m
// panel.  No divergent branch: the LDS read always happens (column and row
// clamped into the panel) and the entry is selected afterwards, so reads
// can be batched ahead of the MFMAs that consume them
__device__ __forceinline__ double panel_vb(const double *P, int ld, int c, int r, int nb, int m)
{
    const int cc = c < QR_NB - 1 ? c : QR_NB - 1;
    const int rr = r < 0 ? 0 : (r < m ? r : m - 1);
    const double raw = P[cc * ld + rr];
    return (c < nb && r < m) ? (r > c ? raw : (r == c ? 1.0 : 0.0)) : 0.0;
}

__device__ __forceinline__ double bcast_lane(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// X[p0:p0+m, j0:j0+16] <- (I - V op(T) V^T) X[...] for one 16-column block,
// one wave.  Rows are held in registers in the MFMA B-operand layout (lane
// (lr, lc) holds rows 4u + lr of column j0 + lc, u < NU); for FP64 16x16x4
// the B layout of four consecutive k-steps is exactly the accumulator
// layout of a 16-row tile, so the registers that feed W = V^T X then receive
// X - V W' in place.  Blocks of up to 4*NU rows are loaded once (64 loads in
// flight per lane); taller blocks take a second pass that reloads each
// chunk but the last, under the MFMAs of the chunk before.  V is the panel P (LDS), T is Tl (LDS, QR_NB x QR_NB,
// Tl[i + c*QR_NB], upper), Wl a wave-private LDS buffer of QR_NB x 16.
// Columns at or beyond jend are untouched.
#define WY_NU 32
// X is reached through a global-address-space pointer (global_load/store
// rather than flat, whose waits would also cover the LDS reads); the loads
// come from clamped rows and are zeroed afterwards, so all WY_NU are in
// flight at once
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ void wy_load(double (&xb)[WY_NU], const double *col, int r0, int m, bool jok, int lr)
{
    const gdouble *g = (const gdouble *)col;
#pragma unroll
    for (int u = 0; u < WY_NU; u++) xb[u] = g[min(r0 + 4 * u + lr, m - 1)];
#pragma unroll
    for (int u = 0; u < WY_NU; u++) {
        const int k = r0 + 4 * u + lr;
        xb[u] = (jok && k < m) ? xb[u] : 0.0;
    }
}

__device__ __forceinline__ void wy_accum_w(const double (&xb)[WY_NU], const double *P, int ld, int nb, int r0, int m,
                                           int lc, int lr, f64x4 (&w)[QR_NT])
{
    // eight k-steps of V fragments read ahead of their MFMAs
#pragma unroll
    for (int u0 = 0; u0 < WY_NU; u0 += 8) {
        if (r0 + 4 * u0 < m) {
            double a[QR_NT][8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int k = r0 + 4 * (u0 + u) + lr;
#pragma unroll
                for (int h = 0; h < QR_NT; h++) a[h][u] = panel_vb(P, ld, 16 * h + lc, k, nb, m);
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
#pragma unroll
                for (int h = 0; h < QR_NT; h++) w[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[h][u], xb[u0 + u], w[h], 0, 0, 0);
        }
    }
}

// X - V W' for the rows r0 .. r0 + 4 WY_NU - 1 held in xb, two 16-row tiles at
// a time (their V fragments read ahead of 16 MFMAs), each pair stored as
// soon as it is done; with rnext >= 0 the pair's registers are then
// refilled with the same rows of the chunk at rnext, so that chunk's loads
// run under the MFMAs of the tiles still to come.  Refilled rows past m are
// clamped copies: their stores are skipped, and W is not formed from them.
__device__ __forceinline__ void wy_update_store(double (&xb)[WY_NU], const double (&wv)[QR_NB / 4], const double *P,
                                                int ld, int nb, int r0, int m, double *col, bool jok, int lc, int lr,
                                                int rnext)
{
    gdouble *g = (gdouble *)col;
#pragma unroll
    for (int t = 0; t < WY_NU / 4; t += 2) {
        if (r0 + 16 * t < m) {
            double av[2][QR_NB / 4];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int r = r0 + 16 * (t + h) + lc;
#pragma unroll
                for (int kk = 0; kk < QR_NB / 4; kk++) av[h][kk] = -panel_vb(P, ld, 4 * kk + lr, r, nb, m);
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                f64x4 c = {xb[4 * (t + h)], xb[4 * (t + h) + 1], xb[4 * (t + h) + 2], xb[4 * (t + h) + 3]};
#pragma unroll
                for (int kk = 0; kk < QR_NB / 4; kk++) c = __builtin_amdgcn_mfma_f64_16x16x4f64(av[h][kk], wv[kk], c, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 4; v++) xb[4 * (t + h) + v] = c[v];
            }
        }
#pragma unroll
        for (int u = 4 * t; u < 4 * t + 8; u++) {
            const int k = r0 + 4 * u + lr;
            if (jok && k < m) g[k] = xb[u];
        }
        if (rnext >= 0) {
#pragma unroll
            for (int u = 4 * t; u < 4 * t + 8; u++) xb[u] = g[min(rnext + 4 * u + lr, m - 1)];
        }
    }
}

__device__ void wy_apply_block(const double *P, int ld, const double *Tl, int nb, int m, double *X, int n, int p0,
                               int j0, int jend, int transT, double *Wl)
{
    const int lane = threadIdx.x & 63;
    const int lc = lane & 15, lr = lane >> 4;
    const int j = j0 + lc;
    const bool jok = j < jend;
    double *col = X + (long long)(jok ? j : j0) * n + p0;
    double xb[WY_NU];
    f64x4 w[QR_NT];
#pragma unroll
    for (int h = 0; h < QR_NT; h++) w[h] = f64x4{0, 0, 0, 0};
    const int chunk = 4 * WY_NU;
    for (int r0 = 0; r0 < m; r0 += chunk) {
        wy_load(xb, col, r0, m, jok, lr);
        wy_accum_w(xb, P, ld, nb, r0, m, lc, lr, w);
    }
#pragma unroll
    for (int h = 0; h < QR_NT; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) Wl[(16 * h + Mfma<double>::row(lane, v)) * 16 + lc] = w[h][v];
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    f64x4 uw[QR_NT];
#pragma unroll
    for (int h = 0; h < QR_NT; h++) uw[h] = f64x4{0, 0, 0, 0};
#pragma unroll
    for (int k0 = 0; k0 < QR_NB; k0 += 4) {
        const int k = k0 + lr;
        const double bw = Wl[k * 16 + lc];
#pragma unroll
        for (int h = 0; h < QR_NT; h++) {
            const int i = 16 * h + lc;
            const double t = transT ? Tl[k + i * QR_NB] : Tl[i + k * QR_NB];
            uw[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(t, bw, uw[h], 0, 0, 0);
        }
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < QR_NT; h++)
#pragma unroll
        for (int v = 0; v < 4; v++) Wl[(16 * h + Mfma<double>::row(lane, v)) * 16 + lc] = uw[h][v];
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    double wv[QR_NB / 4];
#pragma unroll
    for (int kk = 0; kk < QR_NB / 4; kk++) wv[kk] = Wl[(4 * kk + lr) * 16 + lc];
    if (m <= chunk) {
        wy_update_store(xb, wv, P, ld, nb, 0, m, col, jok, lc, lr, -1);
    } else {
        // the last chunk of the W pass is still in registers: it goes
        // first, refilling its registers with the first chunk as it goes,
        // and so on (the chunks are disjoint rows)
        const int rlast = ((m - 1) / chunk) * chunk;
        wy_update_store(xb, wv, P, ld, nb, rlast, m, col, jok, lc, lr, 0);
        for (int r0 = 0; r0 < rlast; r0 += chunk)
            wy_update_store(xb, wv, P, ld, nb, r0, m, col, jok, lc, lr, r0 + chunk < rlast ? r0 + chunk : -1);
    }
    __builtin_amdgcn_wave_barrier();
}

// Reflector of a column from its diagonal entry alpha and the squared norm
// xn2 below it (LAPACK dlarfg): beta = -sign(alpha) sqrt(alpha^2 + xn2),
// tau = (beta - alpha) / beta, scale 1 / (alpha - beta); tau = 0 (identity)
// when xn2 = 0.
__device__ __forceinline__ void reflector(double alpha, double xn2, double &tau, double &sc, double &beta)
{
    tau = 0.0;
    sc = 0.0;
    beta = alpha;
    if (xn2 != 0.0) {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tau = (beta - alpha) / beta;
        sc = 1.0 / (alpha - beta);
    }
}

// Panel factorisation with register-resident columns: wave w owns panel
// columns c2 = QR_WAVES k + w (slot k), lane l holds rows l + 64 s (slot s).
// Reflector c is the unscaled pivot column in P (column c, written once by
// its owner) with tau and the scale 1/(alpha - beta) in s_tau / s_sc.  The
// waves synchronise through two counters instead of a barrier per column:
// s_col = c + 1 once column c is in P and s_pub = c + 1 once its scalars
// are (release stores; consumers acquire-load them, sleeping while they
// are behind).  For reflector c every wave applies it to its live columns
// in order; the owner of column c+1 starts its reduction x_c^T a_{c+1} as
// soon as column c is out, while the publisher of c still works on the
// square root and divisions of dlarfg, then updates column c+1 with the
// scalars and publishes it, at raised priority, and only then updates its
// other columns, so the pivot chain never waits for the bulk of the
// updates.  The spin is bounded: a
// wait that runs out flags the factorisation as failed and goes on.
__device__ __forceinline__ int pub_load(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void pub_store(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void pub_wait(const int *p, int need, int *bad)
{
    for (int it = 0; pub_load(p) < need; it++) {
        if (it >= (1 << 18)) {
            *bad = 1;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

template <int MR>
__device__ void qr_panel(double *A, int n, int p0, int nb, int m, double *P, int ld, int *s_col, int *s_pub,
                         double *s_tau, double *s_sc, double *s_beta, int *s_big, double *Tl, int *bad)
{
    constexpr int NS = QR_NB / QR_WAVES; // column slots per wave
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform: scalar branches
    double pc[NS][MR];
    const double *A0 = A + (long long)p0 * n + p0; // the panel's top-left entry
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const double *colp = A0 + (long long)min(QR_WAVES * k + wave, nb - 1) * n;
#pragma unroll
        for (int s = 0; s < MR; s++) pc[k][s] = colp[min(lane + 64 * s, m - 1)];
    }
    bool nonfinite = false;
#pragma unroll
    for (int k = 0; k < NS; k++)
#pragma unroll
        for (int s = 0; s < MR; s++) {
            pc[k][s] = (QR_WAVES * k + wave < nb && lane + 64 * s < m) ? pc[k][s] : 0.0;
            nonfinite |= !isfinite(pc[k][s]);
        }
    // publish column c2 = col (slot k of this wave) as reflector c2 in two
    // steps: the column itself (s_col = c2 + 1), then, after its norm, the
    // square root and the divisions of dlarfg, its scalars (s_pub = c2 + 1)
    auto publish = [&](const double (&col)[MR], int c2) {
        double q = 0.0;
#pragma unroll
        for (int s = 0; s < MR; s++) {
            const int r = lane + 64 * s;
            if (r < m) P[c2 * ld + r] = col[s];
            if (r > c2 && r < m) q += col[s] * col[s];
        }
        pub_store(s_col, c2 + 1);
        q = wave_sum(q);
        const double alpha = bcast_lane(col[0], c2);
        double tau, sc, beta;
        // squares of entries beyond ~1e150 overflow and those below ~1e-140
        // lose digits: then ||x|| is taken scaled by max|x_i| (dnrm2,
        // dlapy2), and the column is marked big so that its dot products
        // use the scaled reflector (wave-uniform branches)
        const bool plain = q <= 1e300 && q >= 1e-280 && fabs(alpha) <= 1e150;
        if (plain) {
            reflector(alpha, q, tau, sc, beta);
        } else {
            double mx = 0.0;
#pragma unroll
            for (int s = 0; s < MR; s++) {
                const int r = lane + 64 * s;
                if (r > c2 && r < m) mx = fmax(mx, fabs(col[s]));
            }
            mx = wave_max(mx);
            double qs = 0.0;
            if (mx > 0.0) {
#pragma unroll
                for (int s = 0; s < MR; s++) {
                    const int r = lane + 64 * s;
                    const double t = col[s] / mx;
                    if (r > c2 && r < m) qs += t * t;
                }
            }
            qs = wave_sum(qs);
            const double xn = mx * sqrt(qs);
            tau = 0.0;
            sc = 0.0;
            beta = alpha;
            if (xn != 0.0) {
                beta = -copysign(hypot(alpha, xn), alpha);
                tau = (beta - alpha) / beta;
                sc = 1.0 / (alpha - beta);
            }
        }
        if (lane == 0) {
            s_tau[c2] = tau;
            s_sc[c2] = sc;
            s_beta[c2] = beta;
            s_big[c2] = plain ? 0 : 1;
            Tl[c2 * (QR_NB + 1)] = tau;
            if (!isfinite(tau) || !isfinite(sc)) *bad = 1;
        }
        pub_store(s_pub, c2 + 1);
    };
    if (wave == 0) {
        __builtin_amdgcn_s_setprio(2);
        publish(pc[0], 0);
        __builtin_amdgcn_s_setprio(0);
    }
    for (int c = 0; c < nb; c++) {
        // columns of this wave still to update: c2 > c
        if (QR_WAVES * (NS - 1) + wave <= c) break;
        const double *xc = P + c * ld;
        double x[MR];
        auto read_x = [&]() {
#pragma unroll
            for (int s = 0; s < MR; s++) {
                const int r = lane + 64 * s;
                const double raw = xc[min(r, m - 1)];
                x[s] = (r > c && r < m) ? raw : 0.0;
            }
        };
        const int cl = c; // the pivot row c < QR_NB <= 64 sits in register slot 0
        const bool owner = c + 1 < nb && ((c + 1) % QR_WAVES) == wave;
        const int k1 = (c + 1) / QR_WAVES;
        double tau = 0.0, sc = 0.0;
        int bigc = 0;
        // column c2 <- column c2 - w v_c, w = tau (pcc + sc v_c^T x_below), v_c = e_c + sc x;
        // for a big column x holds sc x already (dsum = v_c^T x_below)
        double scd = 0.0; // sc, or 1 for a big column whose x is scaled already
        auto apply = [&](double (&col)[MR], double dsum) {
            const double pcc = bcast_lane(col[0], cl);
            const double w = tau * (pcc + scd * dsum);
            const double ws = w * scd;
#pragma unroll
            for (int s = 0; s < MR; s++) col[s] -= ws * x[s];
            if (lane == cl) col[0] -= w;
        };
        auto scale_x = [&]() {
#pragma unroll
            for (int s = 0; s < MR; s++) x[s] *= sc;
        };
        if (owner) {
            // lookahead: the reduction x_c^T a_{c+1} as soon as column c is
            // out (it overlaps the publisher's square root and divisions),
            // then the update with reflector c's scalars, then publish c+1
            __builtin_amdgcn_s_setprio(2);
            pub_wait(s_col, c + 1, bad);
            read_x();
#pragma unroll
            for (int k = 0; k < NS; k++) {
                if (k == k1) {
                    double acc = 0.0;
#pragma unroll
                    for (int s = 0; s < MR; s++) acc += x[s] * pc[k][s];
                    double dsum = wave_sum(acc);
                    pub_wait(s_pub, c + 1, bad);
                    tau = s_tau[c];
                    sc = s_sc[c];
                    bigc = __builtin_amdgcn_readfirstlane(s_big[c]);
                    scd = bigc ? 1.0 : sc;
                    if (bigc) {
                        scale_x();
                        acc = 0.0;
#pragma unroll
                        for (int s = 0; s < MR; s++) acc += x[s] * pc[k][s];
                        dsum = wave_sum(acc);
                    }
                    if (tau != 0.0) apply(pc[k], dsum);
                    publish(pc[k], c + 1);
                }
            }
            __builtin_amdgcn_s_setprio(0);
        } else {
            pub_wait(s_pub, c + 1, bad);
            read_x();
            tau = s_tau[c];
            sc = s_sc[c];
            bigc = __builtin_amdgcn_readfirstlane(s_big[c]);
            scd = bigc ? 1.0 : sc;
            if (bigc) scale_x();
        }
        if (tau != 0.0) {
            double d[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                double acc = 0.0;
#pragma unroll
                for (int s = 0; s < MR; s++) acc += x[s] * pc[k][s];
                d[k] = acc;
            }
            static_assert(NS == 4, "the four column slots of a wave reduced together");
            wave_sum4(d);
#pragma unroll
            for (int k = 0; k < NS; k++) {
                const int c2 = QR_WAVES * k + wave;
                if (c2 > c && c2 < nb && !(owner && k == k1)) apply(pc[k], d[k]);
            }
        }
    }
    __syncthreads();
    // final panel: R above and on the diagonal, scaled reflectors below;
    // into LDS (for T and the trailing update) and back to A
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int c2 = QR_WAVES * k + wave;
        if (c2 < nb) {
            const double sc = s_sc[c2], beta = s_beta[c2];
            double *colp = A + (long long)(p0 + c2) * n + p0;
#pragma unroll
            for (int s = 0; s < MR; s++) {
                const int r = lane + 64 * s;
                if (r < m) {
                    double v = pc[k][s];
                    if (r == c2) v = beta;
                    else if (r > c2 && sc != 0.0) v *= sc;
                    P[c2 * ld + r] = v;
                    colp[r] = v;
                }
            }
        }
    }
    if (nonfinite) *bad = 1;
    __syncthreads();
}

// T of the compact WY form Q = I - V T V^T of the panel (reflectors with
// tau = 0 are the identity: their row and column of T vanish).  Y = V^T V
// on MFMA: the QR_NT^2 16x16 tiles are spread over the waves, each tile's
// k range split over QR_WAVES / QR_NT^2 waves whose partial sums meet in
// Y; then wave 0 builds the two diagonal 16 x 16 blocks of T at once (a
// dependent chain of 16 column steps) and the off-diagonal block from
// them by two MFMA products.  Y holds QR_WAVES / QR_NT^2 blocks of
// QR_NB x QR_NB.
__device__ void qr_tfactor(const double *P, int ld, int nb, int m, double *Tl, double *Y)
{
    constexpr int KS = QR_WAVES / (QR_NT * QR_NT); // waves per tile
    static_assert(KS >= 2, "two partial sums at least");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    {
        const int tile = wave % (QR_NT * QR_NT), part = wave / (QR_NT * QR_NT);
        const int ti = tile / QR_NT, tj = tile % QR_NT;
        const int ci = ti * 16 + lc, cj = tj * 16 + lc;
        const int steps = (m + 3) >> 2, per = (steps + KS - 1) / KS;
        const int s0 = part * per, s1 = min(steps, s0 + per);
        f64x4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
        // batches of 8 k-steps: all 16 branch-free LDS reads first, then the
        // MFMAs; rows past this wave's range read as row m (zero)
        for (int st0 = s0; st0 < s1; st0 += 8) {
            double va[8], vb[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int st = st0 + u;
                const int k = st < s1 ? 4 * st + lr : m;
                va[u] = panel_vb(P, ld, ci, k, nb, m);
                vb[u] = panel_vb(P, ld, cj, k, nb, m);
            }
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(va[u], vb[u], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(va[u + 1], vb[u + 1], acc1, 0, 0, 0);
            }
        }
        double *dst = Y + part * QR_NB * QR_NB;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = ti * 16 + Mfma<double>::row(lane, v), c = tj * 16 + lc;
            dst[i + c * QR_NB] = acc0[v] + acc1[v];
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < QR_NB * QR_NB; idx += blockDim.x) {
        double sum = Y[idx];
#pragma unroll
        for (int p = 1; p < KS; p++) sum += Y[p * QR_NB * QR_NB + idx];
        Y[idx] = sum;
    }
    __syncthreads();
    static_assert(QR_NB == 32, "T in two 16-column halves");
    if (wave == 0) {
        // T = [T11 T12; 0 T22] over the two 16-column halves of the panel.
        // The diagonal blocks come from the column recurrence (dlarft,
        // forward, columnwise) restricted to their half, both at once: lane
        // 16 h + r (h < 2) builds row r of T_hh column by column, T(r, i) =
        // -tau_i sum_{k<i} T(r, k) Y(k, i), T(i, i) = tau_i, the sum as four
        // independent chains over k mod 4.  The terms with k before the half
        // are products with the exact zeros of T left of its diagonal, so
        // the blocks equal those of the full recurrence.  Then
        // (I - V1 T11 V1^T)(I - V2 T22 V2^T) = I - V1 T11 V1^T - V2 T22 V2^T
        // + V1 T11 (V1^T V2) T22 V2^T gives T12 = -T11 Y12 T22, two
        // 16 x 16 x 16 products on FP64 MFMA.  Lanes 32 .. 63 repeat the
        // recurrence of lanes 0 .. 31 and store nothing.
        const int h = (lane >> 4) & 1, r = lane & 15, base = 16 * h;
        double trow[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const double tau = Tl[(base + i) * (QR_NB + 1)];
            double s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < i; k++) s4[k & 3] += trow[k] * Y[(base + k) + (base + i) * QR_NB];
            const double sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
            trow[i] = (r < i) ? -tau * sum : (r == i ? tau : 0.0);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (lane < QR_NB) {
#pragma unroll
            for (int i = 0; i < 16; i++) Tl[(base + r) + (base + i) * QR_NB] = trow[i];
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        // M = Y12 T22: A[m][k] = Y(m, 16 + k), B[k][n] = T(16 + k, 16 + n)
        f64x4 mt = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = 4 * s + lr;
            mt = Mfma<double>::mma(Y[lc + (16 + k) * QR_NB], Tl[(16 + k) + (16 + lc) * QR_NB], mt);
        }
        // T12 = -T11 M: A[m][k] = T(m, k); the B operand of k-step s, M[4s +
        // lr][lc], is accumulator entry s of the same lane
        f64x4 o = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++) o = Mfma<double>::mma(Tl[lc + (4 * s + lr) * QR_NB], mt[s], o);
#pragma unroll
        for (int v = 0; v < 4; v++) Tl[Mfma<double>::row(lane, v) + (16 + lc) * QR_NB] = -o[v];
    }
    __syncthreads();
}

template <int MR>
__global__ __launch_bounds__(QR_THREADS) void k_qr_udv(QrArgs a)
{
    const int b = blockIdx.x;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int n = a.n, tid = threadIdx.x;
    const int wave = tid >> 6;
    double *A = a.A + (long long)b * a.sA;
    double *Tw = a.Tws + (long long)b * a.sT;

    extern __shared__ double lds[];
    const int ld = n + 1;                      // odd leading dimension: conflict-free column reads
    double *P = lds;                           // QR_NB x n panel, P[c*ld + r]
    double *Tl = P + (size_t)QR_NB * ld;       // QR_NB x QR_NB
    double *Wl = Tl + QR_NB * QR_NB;           // QR_WAVES x (QR_NB x 16); also Y and T scratch
    double *s_tau = Wl + QR_WAVES * QR_NB * 16;
    double *s_sc = s_tau + QR_NB;
    double *s_beta = s_sc + QR_NB;
    __shared__ int s_bad, s_ntau, s_col, s_pub;
    __shared__ int s_big[QR_NB];
    if (tid == 0) { s_bad = 0; s_ntau = 0; }
    __syncthreads();

    const int npan = (n + QR_NB - 1) / QR_NB;
    for (int pi = 0; pi < npan; pi++) {
        const int p0 = pi * QR_NB;
        const int nb = min(QR_NB, n - p0);
        const int m = n - p0;
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tl[idx] = 0.0;
        if (tid == 0) {
            s_col = 0;
            s_pub = 0;
        }
        __syncthreads();
        qr_panel<MR>(A, n, p0, nb, m, P, ld, &s_col, &s_pub, s_tau, s_sc, s_beta, s_big, Tl, &s_bad);
        if (wave == 0) { // non-trivial reflectors, counted with one ballot
            const bool nz = (tid < nb) && s_tau[min(tid, QR_NB - 1)] != 0.0;
            const unsigned long long bal = __ballot(nz);
            if (tid == 0) s_ntau += __popcll(bal);
        }
        qr_tfactor(P, ld, nb, m, Tl, Wl);
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tw[pi * QR_NB * QR_NB + idx] = Tl[idx];
        // trailing update A[p0:, p0+nb:] <- (I - V T^T V^T) A
        for (int j0 = p0 + nb + 16 * wave; j0 < n; j0 += 16 * QR_WAVES)
            wy_apply_block(P, ld, Tl, nb, m, A, n, p0, j0, n, 1, Wl + wave * QR_NB * 16);
        __syncthreads();
    }

    // D = diag R, Rn = D^-1 R (row i scaled by 1 / D_i, the reciprocals
    // kept in the panel buffer, which is free now)
    double *Dd = a.D + (long long)b * a.sD;
    double *rinv = P;
    for (int i = tid; i < n; i += blockDim.x) {
        const double d = A[i + (long long)i * n];
        Dd[i] = d;
        rinv[i] = 1.0 / d;
        if (!isfinite(d) || d == 0.0) s_bad = (s_bad == 1) ? 1 : 2;
    }
    __syncthreads();
    if (a.Rn && tid < (int)(blockDim.x / n) * n) {
        // thread (i, jc) scales row i of columns jc + cpb t; 32 columns in
        // flight (the registers are free here), loads from clamped addresses
        // (n <= blockDim.x)
        double *Rn = a.Rn + (long long)b * a.sR;
        const int cpb = blockDim.x / n, i = tid % n, jc = tid / n;
        const double ri = rinv[i];
        for (int j0 = jc; j0 < n; j0 += 32 * cpb) {
            double v[32];
#pragma unroll
            for (int u = 0; u < 32; u++) v[u] = A[i + (long long)min(j0 + u * cpb, n - 1) * n];
#pragma unroll
            for (int u = 0; u < 32; u++) {
                const int j = j0 + u * cpb;
                if (j < n) Rn[i + (long long)j * n] = i <= j ? v[u] * ri : 0.0;
            }
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
    return sizeof(double) * ((size_t)QR_NB * (n + 1) + QR_NB * QR_NB + QR_WAVES * QR_NB * 16 + 4 * QR_NB);
}

// Explicit Q = H_0 H_1 ... H_{npan-1} (dorgqr) from the reflectors and T
// factors left by k_qr_udv.  Column blocks of Q are independent: a block
// starts as columns of I and receives the block reflectors from the last
// panel back to the first, so each wave owns one 16-column block and the
// launch has n/16 waves per matrix (all CUs busy for a batch of 128).
// Columns j < p0 of panel p are still e_j below row p0 and are skipped.
#define FQ_WAVES 8

__global__ __launch_bounds__(FQ_WAVES * 64) void k_qr_formq(QrArgs a)
{
    const int n = a.n;
    const int nblk = (n + 15) / 16;
    const int wg_per = (nblk + FQ_WAVES - 1) / FQ_WAVES;
    const int b = blockIdx.x / wg_per, g = blockIdx.x % wg_per;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double *A = a.A + (long long)b * a.sA;
    const double *Tw = a.Tws + (long long)b * a.sT;
    double *Q = a.Q + (long long)b * a.sQ;
    extern __shared__ double lds[];
    const int ld = n + 1;
    double *P = lds;
    double *Tl = P + (size_t)QR_NB * ld;
    double *Wl = Tl + QR_NB * QR_NB;
    const int jb = g * FQ_WAVES + wave;
    const int j0 = 16 * jb;
    const int jlast = min(n, 16 * (g + 1) * FQ_WAVES) - 1; // last column of this workgroup
    if (jb < nblk) {
        for (int idx = lane; idx < n * 16; idx += 64) {
            const int r = idx % n, c = j0 + idx / n;
            if (c < n) Q[r + (long long)c * n] = (r == c) ? 1.0 : 0.0;
        }
    }
    const int npan = (n + QR_NB - 1) / QR_NB;
    for (int pi = npan - 1; pi >= 0; pi--) {
        const int p0 = pi * QR_NB;
        if (jlast < p0) continue; // uniform over the workgroup
        const int nb = min(QR_NB, n - p0);
        const int m = n - p0;
        __syncthreads();
        for (int base = tid; base < nb * m; base += 8 * blockDim.x) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * blockDim.x;
                v[u] = idx < nb * m ? A[(p0 + idx % m) + (long long)(p0 + idx / m) * n] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * blockDim.x;
                if (idx < nb * m) P[(idx / m) * ld + idx % m] = v[u];
            }
        }
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tl[idx] = Tw[pi * QR_NB * QR_NB + idx];
        __syncthreads();
        if (jb < nblk && j0 + 15 >= p0)
            wy_apply_block(P, ld, Tl, nb, m, Q, n, p0, j0, n, 0, Wl + wave * QR_NB * 16);
    }
}

// Register-resident variant for n <= 4*NU: each wave keeps its 16-column
// block of Q in registers for all panels (lane (lr, lc) holds rows 4u + lr
// of column j0 + lc, u < NU: the MFMA B layout, which for FP64 16x16x4 is
// also the accumulator layout of 16-row tiles), starting from the columns
// of I, so Q is read from nowhere and written once.  Per panel: W = V^T X
// on MFMA, W' = T W, X -= V W'.
template <int NU>
__global__ __launch_bounds__(FQ_WAVES * 64) void k_qr_formq_reg(QrArgs a)
{
    const int n = a.n;
    const int nblk = (n + 15) / 16;
    const int wg_per = (nblk + FQ_WAVES - 1) / FQ_WAVES;
    const int b = blockIdx.x / wg_per, g = blockIdx.x % wg_per;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const double *A = a.A + (long long)b * a.sA;
    const double *Tw = a.Tws + (long long)b * a.sT;
    double *Q = a.Q + (long long)b * a.sQ;
    extern __shared__ double lds[];
    const int ld = n + 1;
    double *P = lds;
    double *Tl = P + (size_t)QR_NB * ld;
    double *Wl = Tl + QR_NB * QR_NB + wave * QR_NB * 16;
    // column blocks interleaved over the matrix's workgroups (block jb in
    // workgroup jb % wg_per): a block needs the panels at or left of its
    // columns only, so blocks far right carry the most work, and each
    // workgroup gets its share of them
    const int jb = wave * wg_per + g;
    const int j0 = 16 * jb;
    const int j = j0 + lc;
    const int jlast = min(n, 16 * ((FQ_WAVES - 1) * wg_per + g + 1)) - 1; // last column of this workgroup
    double xb[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) xb[u] = (4 * u + lr == j) ? 1.0 : 0.0;
    const int npan = (n + QR_NB - 1) / QR_NB;
    for (int pi = npan - 1; pi >= 0; pi--) {
        const int p0 = pi * QR_NB;
        if (jlast < p0) continue; // uniform over the workgroup
        const int nb = min(QR_NB, n - p0);
        const int m = n - p0;
        __syncthreads();
        for (int base = tid; base < nb * m; base += 8 * blockDim.x) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * blockDim.x;
                v[u] = idx < nb * m ? A[(p0 + idx % m) + (long long)(p0 + idx / m) * n] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int idx = base + u * blockDim.x;
                if (idx < nb * m) P[(idx / m) * ld + idx % m] = v[u];
            }
        }
        for (int idx = tid; idx < QR_NB * QR_NB; idx += blockDim.x) Tl[idx] = Tw[pi * QR_NB * QR_NB + idx];
        __syncthreads();
        if (jb >= nblk || j0 < p0) continue; // columns j < p0 are still e_j below row p0
        // W = V^T X[p0:, block] (rows k >= p0 of X; V(k, c) = panel column c, row k - p0)
        f64x4 w[QR_NT];
#pragma unroll
        for (int h = 0; h < QR_NT; h++) w[h] = f64x4{0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < NU; u++) {
            if (4 * u >= p0 && 4 * u < n) {
                const int k = 4 * u + lr - p0;
#pragma unroll
                for (int h = 0; h < QR_NT; h++)
                    w[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(panel_vb(P, ld, 16 * h + lc, k, nb, m), xb[u], w[h], 0, 0, 0);
            }
        }
#pragma unroll
        for (int h = 0; h < QR_NT; h++)
#pragma unroll
            for (int v = 0; v < 4; v++) Wl[(16 * h + Mfma<double>::row(lane, v)) * 16 + lc] = w[h][v];
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        // W' = T W
        f64x4 uw[QR_NT];
#pragma unroll
        for (int h = 0; h < QR_NT; h++) uw[h] = f64x4{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < QR_NB; k0 += 4) {
            const int k = k0 + lr;
            const double bw = Wl[k * 16 + lc];
#pragma unroll
            for (int h = 0; h < QR_NT; h++)
                uw[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(Tl[(16 * h + lc) + k * QR_NB], bw, uw[h], 0, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int h = 0; h < QR_NT; h++)
#pragma unroll
            for (int v = 0; v < 4; v++) Wl[(16 * h + Mfma<double>::row(lane, v)) * 16 + lc] = uw[h][v];
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        double wv[QR_NB / 4];
#pragma unroll
        for (int kk = 0; kk < QR_NB / 4; kk++) wv[kk] = Wl[(4 * kk + lr) * 16 + lc];
        // X[16t : 16t+16] -= V W' for the row tiles t at or below p0
#pragma unroll
        for (int t = 0; t < NU / 4; t++) {
            if (16 * t >= p0 && 16 * t < n) {
                f64x4 c = {xb[4 * t], xb[4 * t + 1], xb[4 * t + 2], xb[4 * t + 3]};
                const int r = 16 * t + lc - p0;
#pragma unroll
                for (int kk = 0; kk < QR_NB / 4; kk++) {
                    const int k = 4 * kk + lr;
                    const double av = -panel_vb(P, ld, k, r, nb, m);
                    c = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wv[kk], c, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; v++) xb[4 * t + v] = c[v];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (jb < nblk && j < n) {
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int r = 4 * u + lr;
            if (r < n) Q[r + (long long)j * n] = xb[u];
        }
    }
}

static inline size_t qr_formq_lds_bytes(int n)
{
    return sizeof(double) * ((size_t)QR_NB * (n + 1) + QR_NB * QR_NB + FQ_WAVES * QR_NB * 16);
}

static inline int qr_formq_blocks(int n, int batch)
{
    const int nblk = (n + 15) / 16;
    return ((nblk + FQ_WAVES - 1) / FQ_WAVES) * batch;
}

// X = U^-1 B (or X = U^-T B) for unit upper-triangular U (the diagonal of
// U is not read): dtrsm('L','U','N','U') (or 'T'), B rows optionally scaled
// on load.  Column blocks of X are independent, so a
// workgroup takes 32 columns of one matrix (n/32 workgroups per matrix) and
// walks the 32-row blocks from the bottom: S = B_I - U[I, I+1:] X[I+1:] on
// MFMA (one 16x16 tile per wave, U streamed from global memory, X resident
// in LDS), then the unit solve U_II X_I = S, one column per lane of wave 0
// with U_II staged in LDS.
struct TrsmArgs {
    int n, batch;
    const double *U;
    long long sU;
    const double *B;
    long long sB;
    double *X;
    long long sX;
    const int *status;
    int trans;        // 0: X = U^-1 B; 1: X = U^-T B (forward substitution)
    const double *rs; // optional row scaling of B (B(i, :) * rs(i)), may be null
    long long srs;
};

#define TS_NB 32

static inline size_t trsm_lds_bytes(int n)
{
    return sizeof(double) * ((size_t)TS_NB * (n + 4) + TS_NB * (TS_NB + 1));
}

static inline int trsm_blocks(int n, int batch) { return ((n + TS_NB - 1) / TS_NB) * batch; }

__global__ __launch_bounds__(256) void k_trsm_unit_upper(TrsmArgs a)
{
    const int n = a.n;
    const int nblk = (n + TS_NB - 1) / TS_NB;
    // a matrix's column blocks on one XCD: they read the same U, row block
    // by row block, from that XCD's L2
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int b = logical / nblk, cb = logical % nblk;
    if (b >= a.batch) return;
    if (a.status && a.status[b]) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lc = lane & 15, lr = lane >> 4;
    const int c0 = cb * TS_NB, cw = min(TS_NB, n - c0);
    // global-space views: every load below comes from a clamped, valid
    // address and is zeroed afterwards, so a batch is in flight at once
    const gdouble *U = (const gdouble *)(a.U + (long long)b * a.sU);
    const gdouble *Bm = (const gdouble *)(a.B + (long long)b * a.sB);
    gdouble *X = (gdouble *)(a.X + (long long)b * a.sX);
    extern __shared__ double lds[];
    const int ldx = n + 4;              // Xs[c*ldx + r]: column-major, conflict-free fragment reads
    double *Xs = lds;
    double *Ud = Xs + (size_t)TS_NB * ldx; // U_II, Ud[i*(TS_NB+1) + k]
    // B block: thread t takes rows t, t + 256, ... of every column
    const double *rsb = a.rs ? a.rs + (long long)b * a.srs : nullptr;
    for (int r = tid; r < n; r += blockDim.x) {
        double v[TS_NB];
        const double sr = rsb ? rsb[r] : 1.0;
#pragma unroll
        for (int c = 0; c < TS_NB; c++) v[c] = Bm[r + (long long)(c0 + min(c, cw - 1)) * n];
#pragma unroll
        for (int c = 0; c < TS_NB; c++) Xs[c * ldx + r] = c < cw ? v[c] * sr : 0.0;
    }
    __syncthreads();
    const int ti = wave >> 1, tj = wave & 1;
    const bool tr = a.trans != 0;
    for (int It = 0; It < nblk; It++) {
        const int I = tr ? It : nblk - 1 - It; // blocks bottom-up, or top-down for U^T
        const int r0 = I * TS_NB, ib = min(TS_NB, n - r0);
        // U_II into LDS (loads issued before the product below)
        double ud[TS_NB * TS_NB / 256];
#pragma unroll
        for (int u = 0; u < TS_NB * TS_NB / 256; u++) {
            const int idx = tid + 256 * u;
            const int i = idx % TS_NB, k = idx / TS_NB;
            ud[u] = tr ? U[(r0 + min(k, ib - 1)) + (long long)(r0 + min(i, ib - 1)) * n]  // (U^T)_II(i, k)
                       : U[(r0 + min(i, ib - 1)) + (long long)(r0 + min(k, ib - 1)) * n];
        }
        // S = B_I - U[r0:r0+ib, r0+ib:n] X[r0+ib:n] (U^T: B_I - U[0:r0, r0:r0+ib]^T X[0:r0]):
        // tile (ti, tj), rows i, columns c
        {
            f64x4 acc0, acc1 = {0, 0, 0, 0};
            const int i0 = r0 + ti * 16, cc = tj * 16 + lc;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + Mfma<double>::row(lane, v);
                acc0[v] = (i < r0 + ib) ? Xs[cc * ldx + min(i, n - 1)] : 0.0;
            }
            const bool iok = i0 + lc < r0 + ib;
            const int ic = min(i0 + lc, n - 1);
            const int kbeg = tr ? 0 : r0 + ib, kend = tr ? r0 : n;
            // U chunks of 64 k are double-buffered: the next chunk's loads
            // are in flight while this chunk's MFMAs run
            auto load_u = [&](double (&ua)[16], int k0) {
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    const int kc = min(k0 + 4 * u + lr, n - 1);
                    ua[u] = tr ? U[kc + (long long)ic * n] : U[ic + (long long)kc * n];
                }
            };
            auto mma_chunk = [&](double (&ua)[16], int k0) {
                double xb[16];
#pragma unroll
                for (int u = 0; u < 16; u++) xb[u] = Xs[cc * ldx + min(k0 + 4 * u + lr, n - 1)];
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    const bool ok = k0 + 4 * u + lr < kend;
                    ua[u] = (ok && iok) ? -ua[u] : 0.0;
                    xb[u] = ok ? xb[u] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 16; u += 2) {
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[u], xb[u], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[u + 1], xb[u + 1], acc1, 0, 0, 0);
                }
            };
            if (kbeg < kend) {
                double ua0[16], ua1[16];
                load_u(ua0, kbeg);
                for (int k0 = kbeg; k0 < kend; k0 += 128) {
                    if (k0 + 64 < kend) load_u(ua1, k0 + 64);
                    mma_chunk(ua0, k0);
                    if (k0 + 64 < kend) {
                        if (k0 + 128 < kend) load_u(ua0, k0 + 128);
                        mma_chunk(ua1, k0 + 64);
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + Mfma<double>::row(lane, v);
                if (i < r0 + ib) Xs[cc * ldx + i] = acc0[v] + acc1[v];
            }
        }
#pragma unroll
        for (int u = 0; u < TS_NB * TS_NB / 256; u++) {
            const int idx = tid + 256 * u;
            const int i = idx % TS_NB, k = idx / TS_NB;
            Ud[i * (TS_NB + 1) + k] = (i < ib && k < ib) ? ud[u] : 0.0;
        }
        __syncthreads();
        {
            // the unit solve of the 32 columns: every 16-lane row takes one
            // column (4 per wave, two passes over the 32), lane l of the row
            // holds rows l and 16 + l; right-looking, so once x_i is final
            // it is broadcast along the row (DPP) and removed from the rows
            // still open, with each entry updated in the same order as a
            // column-per-lane substitution
            const int rl = lane & 15;
#pragma unroll
            for (int pass = 0; pass < 2; pass++) {
                const int c = pass * 16 + wave * 4 + (lane >> 4);
                double *col = Xs + c * ldx + r0;
                double sv[2];
#pragma unroll
                for (int h = 0; h < 2; h++) sv[h] = col[min(16 * h + rl, ib - 1)];
                if (!tr) {
#pragma unroll
                    for (int i = TS_NB - 1; i >= 1; i--) {
                        if (i < ib) {
                            const double xi = row_bcast(sv[i >> 4], i);
#pragma unroll
                            for (int h = 0; h < 2; h++) {
                                const int k = 16 * h + rl;
                                if (k < i) sv[h] -= Ud[k * (TS_NB + 1) + i] * xi;
                            }
                        }
                    }
                } else {
                    // Ud holds (U^T)_II, unit lower: forward
#pragma unroll
                    for (int i = 0; i < TS_NB - 1; i++) {
                        if (i < ib) {
                            const double xi = row_bcast(sv[i >> 4], i);
#pragma unroll
                            for (int h = 0; h < 2; h++) {
                                const int k = 16 * h + rl;
                                if (k > i) sv[h] -= Ud[k * (TS_NB + 1) + i] * xi;
                            }
                        }
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; h++)
                    if (16 * h + rl < ib && c < TS_NB) col[16 * h + rl] = sv[h];
            }
        }
        __syncthreads();
    }
    for (int r = tid; r < n; r += blockDim.x) {
#pragma unroll
        for (int c = 0; c < TS_NB; c++)
            if (c < cw) X[r + (long long)(c0 + c) * n] = Xs[c * ldx + r];
    }
}

