/*
 * oracle.h -- CPU restatement of the reference DQMC Green's-function path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (libdqmc_gpu.so and the host driver) never links,
 * loads or calls it.
 *
 * Plain C, column-major n x n doubles (A[i + j*n], src/linalg.h:4), its own
 * dense linear algebra (no LAPACK).  Each function cites the reference
 * file:line it restates.  How each function is pinned (golden fixtures,
 * the compiled reference, closed forms, invariants) is listed in
 * tests/test_oracle.py; every function below has at least one such check.
 */
#ifndef SAIQMC_ORACLE_H
#define SAIQMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers: xoshiro256** seeded by SplitMix64 ---- */
void orc_rng_seed(uint64_t s[4], uint64_t seed);        /* src/rng.c:16-21 */
double orc_rng_double(uint64_t s[4]);                   /* src/rng.c:23-37 */
uint64_t orc_replica_seed(uint64_t base, int beta_index,
                          int replica_id);              /* src/replica.c:17-34 */

/* ---- lattice / model ---- */
/* type 0 = chain (src/lattice.c:26-44), 1 = square (src/lattice.c:46-82).
   Returns is_bipartite. */
int orc_lattice(int type, int Lx, int Ly, int pbc, double t, double *thop,
                int *bipart);
/* exp(scale * A) for symmetric A (src/linalg.c:319-346 computes it by an
   eigendecomposition; the oracle uses scaling-and-squaring of a Taylor
   series in long double, an independent method). */
void orc_expm_sym(int n, const double *A, double scale, double *out);
/* model_init with half filling (src/model.c:8-31) + lambda
   (src/field.c:9). */
void orc_model(int n, const double *thop, double U, double dtau,
               double *expK, double *expKinv, double *lambda);

/* ---- dense linear algebra (LAPACK conventions) ---- */
void orc_matmul(int n, const double *A, const double *B, double *C);
void orc_geqrf(int n, double *A, double *tau);          /* dgeqrf */
void orc_orgqr(int n, double *A, const double *tau);    /* dorgqr */
int orc_getrf(int n, double *A, int *piv);              /* dgetrf, 0-based */
int orc_inverse(int n, const double *A, double *Ainv);  /* la_inverse */
void orc_trtri_unit_upper(int n, double *T);            /* dtrtri U,U */

/* ---- UDV chain (src/linalg.c) ; return 0 ok, 1 failed ---- */
int orc_udv_lmul(int n, double *U, double *D, double *T, const double *B);
int orc_udv_rmul(int n, double *U, double *D, double *T, const double *B);
int orc_udv_combine(int n, const double *Ul, const double *Dl,
                    const double *Tl, const double *Ur, const double *Dr,
                    const double *Tr, double *Uo, double *Do, double *To);
int orc_udv_inv_one_plus(int n, const double *U, const double *D,
                         const double *T, double *g, int *sign);
int orc_udv_inv_one_plus_off(int n, const double *U, const double *D,
                             const double *T, double off, double *g,
                             int *sign);           /* src/linalg.c:911-1115 */
int orc_udv_inv_two_sided(int n, const double *Ul, const double *Dl,
                          const double *Tl, double offl, const double *Ur,
                          const double *Dr, const double *Tr, double offr,
                          double *g, int *sign);   /* src/linalg.c:1117-1389 */
int orc_udv_recenter(int n, double *D, double *off); /* src/linalg.c:453-513 */
int orc_udv_lmul_centered(int n, double *U, double *D, double *T, double *off,
                          const double *B);        /* src/linalg.c:515-653 */
int orc_udv_rmul_centered(int n, double *U, double *D, double *T, double *off,
                          const double *B);        /* src/linalg.c:662-794 */

/* ---- B matrices and wrap (src/green.c) ---- */
/* expv[0] for s=-1, expv[1] for s=+1 (src/green.c:54-67, sigma=+1). */
void orc_expv(double lambda, double U, double dtau, double sigma,
              double expv[2]);
void orc_build_B(int n, const double *expK, const double expv[2],
                 const signed char *s_l, double *out);
void orc_build_Binv(int n, const double *expKinv, const double expv[2],
                    const signed char *s_l, double *out);
void orc_build_Bblock(int n, const double *expK, const double expv[2],
                      const signed char *s, int l_begin, int len,
                      double *out);
void orc_wrap(int n, const double *expK, const double *expKinv,
              const double expv[2], const signed char *s_l, double *g);
void orc_wrap_backward(int n, const double *expK, const double *expKinv,
                       const double expv[2], const signed char *s_l,
                       double *g);                        /* src/green.c:529-546 */
void orc_ph_down(int n, const int *bipart, const double *gu, double *gd);

/* ---- measurement ---- */
/* out5 = {E, ekin, eint, ntot, doublon} (src/measure.c:6-39). */
void orc_measure_sample(int n, const double *thop, double U,
                        const double *gu, const double *gd, double *out5);
/* phase_cos[disp + q*n] (src/structure_factor.c:185-203). */
void orc_phase_cos(int Lx, int Ly, int nq, const int *mx, const int *my,
                   double *phase_cos);
/* measure_szz_sperp_sample (src/structure_factor.c:328-400). */
int orc_szz_sperp(int Lx, int Ly, int nq, const double *phase_cos,
                  const double *gu, const double *gd, double *szz,
                  double *sperp);
void orc_jackknife(const double *x, int N, double *mean, double *err);

/* ---- exact G of a field (exact.c): (1 + B_{tau-1}..B_0 B_{L-1}..B_tau)^-1
        in 80-bit long double, slices grouped by blk; s holds L*n values
        +-1 (slice-major).  Not a reference function: the quantity the FP64
        paths approximate.  Returns 0 ok. ---- */
int orc_green_exact(int n, int L, int tau, int blk, const double *expK,
                    const double expv[2], const signed char *s, double *g);

/* ---- full replica engine, PH sweeps: forward (src/dqmc.c:609-813) and
        alternating forward/backward (src/dqmc.c:815-963) ---- */
typedef struct orc_replica orc_replica;
orc_replica *orc_replica_new(int n, int L, int stab, const double *thop,
                             const int *bipart, double U, double dtau,
                             uint64_t seed, int delay_cap, int drift);
int orc_replica_status(const orc_replica *R);
int orc_replica_sweep(orc_replica *R);
void orc_replica_set_alternating(orc_replica *R, int on);
int orc_replica_set_rebuild(orc_replica *R, int mode);
int orc_replica_reason(const orc_replica *R);
void orc_replica_green(const orc_replica *R, double *gu, double *gd);
void orc_replica_field(const orc_replica *R, signed char *s);
void orc_replica_set_field(orc_replica *R, const signed char *s);
void orc_replica_stats(const orc_replica *R, unsigned long long *acc,
                       unsigned long long *att, double *sign, int *det_sign);
void orc_replica_rng_state(const orc_replica *R, uint64_t *s4);
void orc_replica_model(const orc_replica *R, double *expK, double *expKinv,
                       double *lambda);
void orc_replica_drift(const orc_replica *R, double *max_inf,
                       double *mean_inf, unsigned long long *samples);
/* Re-run the initial rebuild from the current field (src/dqmc.c:466). */
int orc_replica_rebuild(orc_replica *R);
/* Ratio log: after orc_replica_log_ratios, every proposal's R_up R_dn
   (src/green.c:401-416) is kept per slice and site; orc_replica_ratios
   copies the L * n values of the latest visit of each slice. */
int orc_replica_log_ratios(orc_replica *R);
void orc_replica_ratios(const orc_replica *R, double *out);
void orc_replica_free(orc_replica *R);

#ifdef __cplusplus
}
#endif

#endif
