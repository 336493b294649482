/*
 * ref_trace.c -- test infrastructure, not product code.
 *
 * A flat C API over the reference SAI-QMC engine, compiled together with the
 * reference's own sources (src/*.c except src/main.c, built by
 * oracle/Makefile into oracle/_ref/).  Tests and the golden-fixture script
 * load it through ctypes to read the reference's exact Monte Carlo
 * trajectory (fields, accept counts, Green's functions) and its per-kernel
 * linear-algebra results.  Nothing in the product links or loads it.
 *
 * Every function only sequences calls the reference itself makes; the order
 * follows dqmc_run_replica (src/replica_run.c:89-477).
 */
#include "dqmc.h"
#include "field.h"
#include "green.h"
#include "lattice.h"
#include "linalg.h"
#include "measure.h"
#include "model.h"
#include "replica.h"
#include "rng.h"
#include "structure_factor.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    Lattice lat;
    Model m;
    Rng r;
    Field f;
    Dqmc D;
    int init_rc;
} RefReplica;

/* lattice_type: 0 = chain (src/lattice.c:26-44), 1 = square (:46-82).
   alternating selects DQMC_SWEEP_ALTERNATING (src/dqmc.h:45-48),
   rebuild the GreenRebuildMode (src/green.h:10-14: 0 combine, 1 two_sided,
   2 centered). */
void *ref_replica_new_mode(int lattice_type, int Lx, int Ly, int pbc,
                           double thop, double U, double dtau, int Ltr,
                           int stab, unsigned long long seed, int drift,
                           int alternating, int rebuild)
{
    RefReplica *R = calloc(1, sizeof(*R));
    if (R == NULL) return NULL;
    if (lattice_type == 0) {
        lattice_chain(&R->lat, Lx, thop, pbc);
    } else {
        lattice_square(&R->lat, Lx, Ly, thop, pbc);
    }
    /* src/replica_run.c:106-126 */
    model_init(&R->m, &R->lat, U, dtau, 1, 0.0);
    rng_seed(&R->r, seed);
    field_init(&R->f, R->lat.n, Ltr, U, dtau, &R->r);
    R->init_rc = dqmc_init_modes(
        &R->D, &R->m, &R->f, &R->r, stab, NULL,
        alternating ? DQMC_SWEEP_ALTERNATING : DQMC_SWEEP_FORWARD,
        (GreenRebuildMode)rebuild);
    if (drift) {
        dqmc_enable_stab_drift(&R->D, 1);
    }
    return R;
}

void *ref_replica_new(int lattice_type, int Lx, int Ly, int pbc, double thop,
                      double U, double dtau, int Ltr, int stab,
                      unsigned long long seed, int drift)
{
    return ref_replica_new_mode(lattice_type, Lx, Ly, pbc, thop, U, dtau, Ltr,
                                stab, seed, drift, 0, 0);
}

int ref_replica_init_rc(void *p) { return ((RefReplica *)p)->init_rc; }

int ref_replica_sweep(void *p)
{
    RefReplica *R = p;
    dqmc_sweep(&R->D);
    return R->D.status;
}

void ref_replica_green(void *p, double *gu, double *gd)
{
    RefReplica *R = p;
    const size_t nn = (size_t)R->lat.n * (size_t)R->lat.n;
    if (gu) memcpy(gu, R->D.Gu.g, nn * sizeof(double));
    if (gd) memcpy(gd, R->D.Gd.g, nn * sizeof(double));
}

void ref_replica_field(void *p, signed char *s)
{
    RefReplica *R = p;
    memcpy(s, R->f.s, (size_t)R->f.L * (size_t)R->f.n);
}

void ref_replica_set_field(void *p, const signed char *s)
{
    RefReplica *R = p;
    memcpy(R->f.s, s, (size_t)R->f.L * (size_t)R->f.n);
}

void ref_replica_stats(void *p, unsigned long long *acc,
                       unsigned long long *att, double *sign, int *status,
                       int *det_sign)
{
    RefReplica *R = p;
    *acc = R->D.accept_accepted;
    *att = R->D.accept_attempts;
    *sign = R->D.sign;
    *status = R->D.status;
    *det_sign = R->D.Gu.det_sign;
}

void ref_replica_drift(void *p, double *max_inf, double *mean_inf,
                       unsigned long long *samples)
{
    RefReplica *R = p;
    *max_inf = R->D.stab_drift.max_inf;
    *samples = R->D.stab_drift.samples;
    *mean_inf = R->D.stab_drift.samples
                    ? R->D.stab_drift.sum_inf / (double)R->D.stab_drift.samples
                    : 0.0;
}

void ref_replica_model(void *p, double *expK, double *expKinv, double *thop,
                       int *bipart, double *lambda)
{
    RefReplica *R = p;
    const int n = R->lat.n;
    const size_t nn = (size_t)n * (size_t)n;
    if (expK) memcpy(expK, R->m.expK, nn * sizeof(double));
    if (expKinv) memcpy(expKinv, R->m.expKinv, nn * sizeof(double));
    if (thop) memcpy(thop, R->lat.t, nn * sizeof(double));
    if (bipart) memcpy(bipart, R->lat.bipart, (size_t)n * sizeof(int));
    if (lambda) *lambda = R->f.lambda;
}

void ref_replica_rng_state(void *p, unsigned long long *s4)
{
    RefReplica *R = p;
    for (int i = 0; i < 4; i++) s4[i] = R->r.s[i];
}

/* measure_sample (src/measure.c:6-39) on the current G, packed as
   {E, ekin, eint, ntot, doublon}. */
void ref_replica_measure(void *p, double *out5)
{
    RefReplica *R = p;
    MeasSample s = measure_sample(R->lat.n, R->lat.t, R->m.U, R->D.Gu.g,
                                  R->D.Gd.g);
    out5[0] = s.E;
    out5[1] = s.ekin;
    out5[2] = s.eint;
    out5[3] = s.ntot;
    out5[4] = s.doublon;
}

/* measure_szz_sperp_sample with selector "all"
   (src/structure_factor.c:107-206, 328-400). */
int ref_replica_szz_sperp(void *p, double *szz, double *sperp)
{
    RefReplica *R = p;
    StructureFactorPlan plan;
    StructureFactorWorkspace wz = {0}, wp = {0};
    if (structure_factor_plan_init(&plan, &R->lat, "all") != 0) return 1;
    structure_factor_workspace_init(&wz, &plan);
    structure_factor_workspace_init(&wp, &plan);
    const int rc = measure_szz_sperp_sample(&plan, &wz, szz, &plan, &wp,
                                            sperp, R->D.Gu.g, R->D.Gd.g);
    structure_factor_workspace_free(&wz);
    structure_factor_workspace_free(&wp);
    structure_factor_plan_free(&plan);
    return rc;
}

void ref_replica_free(void *p)
{
    RefReplica *R = p;
    if (R == NULL) return;
    dqmc_free(&R->D);
    field_free(&R->f);
    model_free(&R->m);
    lattice_free(&R->lat);
    free(R);
}

/* ---- per-kernel entry points on caller-provided UDV factors ---- */

static void udv_load(UDV *s, int n, const double *U, const double *D,
                     const double *T)
{
    udv_init(s, n);
    memcpy(s->U, U, sizeof(double) * (size_t)n * (size_t)n);
    memcpy(s->D, D, sizeof(double) * (size_t)n);
    memcpy(s->T, T, sizeof(double) * (size_t)n * (size_t)n);
}

static void udv_store(const UDV *s, double *U, double *D, double *T)
{
    const int n = s->n;
    memcpy(U, s->U, sizeof(double) * (size_t)n * (size_t)n);
    memcpy(D, s->D, sizeof(double) * (size_t)n);
    memcpy(T, s->T, sizeof(double) * (size_t)n * (size_t)n);
}

/* udv_lmul_work (src/linalg.c:515-642), in place. Returns work.failed. */
int ref_udv_lmul(int n, double *U, double *D, double *T, const double *B)
{
    UDV s;
    LinalgWork w;
    udv_load(&s, n, U, D, T);
    linalg_work_init(&w, n);
    udv_lmul_work(&s, B, &w);
    const int failed = w.failed;
    udv_store(&s, U, D, T);
    linalg_work_free(&w);
    udv_free(&s);
    return failed;
}

/* udv_rmul (src/linalg.c:662-784), in place. */
int ref_udv_rmul(int n, double *U, double *D, double *T, const double *B)
{
    UDV s;
    LinalgWork w;
    udv_load(&s, n, U, D, T);
    linalg_work_init(&w, n);
    udv_rmul(&s, B, &w);
    const int failed = w.failed;
    udv_store(&s, U, D, T);
    linalg_work_free(&w);
    udv_free(&s);
    return failed;
}

/* udv_combine (src/linalg.c:796-909). */
int ref_udv_combine(int n, const double *Ul, const double *Dl,
                    const double *Tl, const double *Ur, const double *Dr,
                    const double *Tr, double *Uo, double *Do, double *To)
{
    UDV l, r, o;
    LinalgWork w;
    udv_load(&l, n, Ul, Dl, Tl);
    udv_load(&r, n, Ur, Dr, Tr);
    udv_init(&o, n);
    linalg_work_init(&w, n);
    udv_combine(&l, &r, &o, &w);
    const int failed = w.failed;
    udv_store(&o, Uo, Do, To);
    linalg_work_free(&w);
    udv_free(&l);
    udv_free(&r);
    udv_free(&o);
    return failed;
}

/* udv_inv_one_plus_work (src/linalg.c:911-1115). */
int ref_udv_inv_one_plus(int n, const double *U, const double *D,
                         const double *T, double *g, int *sign)
{
    UDV s;
    LinalgWork w;
    udv_load(&s, n, U, D, T);
    linalg_work_init(&w, n);
    const int rc = udv_inv_one_plus_work(&s, g, sign, &w);
    linalg_work_free(&w);
    udv_free(&s);
    return rc;
}

unsigned long long ref_replica_seed(unsigned long long base, int b, int r)
{
    return replica_seed(base, b, r);
}

void ref_expm_sym(int n, const double *A, double scale, double *out)
{
    la_expm_sym(n, A, scale, out);
}
