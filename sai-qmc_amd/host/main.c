/*
 * main.c -- the host program of the GPU engine: reads the reference's
 * key=value input, runs every beta of beta_list on one MI355X through
 * include/dqmc_gpu.h and writes observables.dat, szz.dat and sperp.dat in
 * the reference's formats (src/main.c:221-317, 880-990, 1035-2086).
 *
 *     dqmc_gpu_main <input file>
 *
 * Per beta (src/main.c:1035-1080): L = lround(beta/dtau) slices, replica r
 * gets replica_seed(seed, beta_index, r) (src/replica.c:17-34), the engine
 * draws its fields exactly as field_init does, runs nwarm sweeps and nbin
 * bins of nmeas/nbin measured sweeps (src/replica_run.c:180-455), and the
 * host turns the bins into per-bin ratios (replica_bin_values,
 * src/replica.c:88-120; szz_bin_ratio, src/structure_factor.c:423-440),
 * checks the Szz/Sperp sum rules on every bin (src/structure_factor.c:446-
 * 510) and jackknifes over nrep*nbin bins (src/measure.c:48-69).
 */
#include "host.h"
#include "dqmc_gpu.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int enabled, is_all, nq;
    int *mx, *my;
    int *col; /* column of this momentum in the engine's q list */
} Plan;

static void plan_free(Plan *p)
{
    free(p->mx);
    free(p->my);
    free(p->col);
    memset(p, 0, sizeof(*p));
}

/* structure_factor_plan_init selectors (src/structure_factor.c:44-206):
   none, af, all (mx fastest) or an explicit "mx:my,mx:my,..." list. */
static int plan_init(Plan *p, const Lattice *L, const char *sel)
{
    memset(p, 0, sizeof(*p));
    if (!sel[0] || strpbrk(sel, " \t\r\n")) return 1;
    if (!strcmp(sel, "none")) return 0;
    p->enabled = 1;
    if (!strcmp(sel, "af")) {
        if ((L->Lx > 1 && L->Lx % 2) || (L->Ly > 1 && L->Ly % 2)) return 1;
        p->nq = 1;
        p->mx = malloc(sizeof(int));
        p->my = malloc(sizeof(int));
        p->mx[0] = L->Lx > 1 ? L->Lx / 2 : 0;
        p->my[0] = L->Ly > 1 ? L->Ly / 2 : 0;
    } else if (!strcmp(sel, "all")) {
        p->is_all = 1;
        p->nq = L->n;
        p->mx = malloc(sizeof(int) * L->n);
        p->my = malloc(sizeof(int) * L->n);
        int q = 0;
        for (int my = 0; my < L->Ly; my++)
            for (int mx = 0; mx < L->Lx; mx++) {
                p->mx[q] = mx;
                p->my[q] = my;
                q++;
            }
    } else {
        int nq = 1;
        for (const char *c = sel; *c; c++) nq += (*c == ',');
        p->mx = malloc(sizeof(int) * nq);
        p->my = malloc(sizeof(int) * nq);
        const char *c = sel;
        for (int q = 0; q < nq; q++) {
            char *end;
            const long a = strtol(c, &end, 10);
            if (end == c || a < 0 || *end != ':') return 1;
            c = end + 1;
            const long b = strtol(c, &end, 10);
            if (end == c || b < 0) return 1;
            if ((q + 1 < nq && *end != ',') || (q + 1 == nq && *end != '\0')) return 1;
            if (a >= L->Lx || b >= L->Ly || (L->is_chain && b != 0)) return 1;
            for (int k = 0; k < q; k++)
                if (p->mx[k] == a && p->my[k] == b) return 1;
            p->mx[q] = (int)a;
            p->my[q] = (int)b;
            c = end + (q + 1 < nq ? 1 : 0);
        }
        p->nq = nq;
    }
    p->col = malloc(sizeof(int) * p->nq);
    return p->col ? 0 : 1;
}

/* the engine measures one q list per model: the union of both plans */
static int union_plans(const Lattice *L, Plan *a, Plan *b, double **phase, int *nq_out)
{
    const int cap = (a->enabled ? a->nq : 0) + (b->enabled ? b->nq : 0);
    int *ux = malloc(sizeof(int) * (cap + 1)), *uy = malloc(sizeof(int) * (cap + 1));
    int nq = 0;
    Plan *ps[2] = {a, b};
    for (int k = 0; k < 2; k++) {
        if (!ps[k]->enabled) continue;
        for (int q = 0; q < ps[k]->nq; q++) {
            int f = -1;
            for (int u = 0; u < nq; u++)
                if (ux[u] == ps[k]->mx[q] && uy[u] == ps[k]->my[q]) f = u;
            if (f < 0) {
                ux[nq] = ps[k]->mx[q];
                uy[nq] = ps[k]->my[q];
                f = nq++;
            }
            ps[k]->col[q] = f;
        }
    }
    if (nq == 0) { /* nothing requested: measure q = 0 so the scalar bins still run */
        ux[0] = 0;
        uy[0] = 0;
        nq = 1;
    }
    double *ph = malloc(sizeof(double) * (size_t)L->n * nq);
    const double two_pi = 2.0 * acos(-1.0);
    for (int q = 0; q < nq; q++)
        for (int d = 0; d < L->n; d++) {
            const int dx = d % L->Lx, dy = d / L->Lx;
            ph[d + (size_t)q * L->n] =
                cos(two_pi * ((double)ux[q] * (double)dx / (double)L->Lx + (double)uy[q] * (double)dy / (double)L->Ly));
        }
    free(ux);
    free(uy);
    *phase = ph;
    *nq_out = nq;
    return 0;
}

static int sum_rule_ok(const Plan *p, const Lattice *L, const double *v, double ntot, double doublon, double pref)
{
    if (!p->is_all) return 1;
    const double rhs = pref * (ntot - 2.0 * (double)L->n * doublon);
    const double tol = 1e-12 + 1e-10 * fmax(1.0, fabs(rhs));
    double lhs = 0.0;
    for (int q = 0; q < p->nq; q++) lhs += v[q];
    return isfinite(lhs) && fabs(lhs - rhs) <= tol;
}

static void write_q(FILE *fp, const Plan *p, const Lattice *L, double beta_req, double beta, double T,
                    const double *qbins, int total)
{
    for (int q = 0; q < p->nq; q++) {
        double mean, err;
        jackknife(&qbins[(size_t)q * total], total, &mean, &err);
        const int fx = (L->Lx > 1 && p->mx[q] > L->Lx / 2) ? p->mx[q] - L->Lx : p->mx[q];
        const int fy = (L->Ly > 1 && p->my[q] > L->Ly / 2) ? p->my[q] - L->Ly : p->my[q];
        const double qx = L->Lx > 1 ? 2.0 * p->mx[q] / L->Lx : 0.0;
        const double qy = L->Ly > 1 ? 2.0 * p->my[q] / L->Ly : 0.0;
        const double qfx = L->Lx > 1 ? 2.0 * fx / L->Lx : 0.0;
        const double qfy = L->Ly > 1 ? 2.0 * fy / L->Ly : 0.0;
        fprintf(fp, "%.17g %.17g %.17g %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g\n", beta_req, beta, T, q,
                p->mx[q], p->my[q], qx, qy, qfx, qfy, mean, err);
    }
    fflush(fp);
}

static int fail(const char *what, dqmc_gpu_ctx *ctx)
{
    fprintf(stderr, "ERROR: %s: %s\n", what, dqmc_gpu_last_error());
    if (ctx) dqmc_gpu_destroy(ctx);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s <input file>\n", argv[0]);
        return 2;
    }
    Params p;
    if (params_read(&p, argv[1])) return 1;
    if (strcmp(p.parallel, "serial") && strcmp(p.parallel, "omp")) {
        fprintf(stderr, "ERROR: parallel=%s: the GPU driver runs all replicas on one device\n", p.parallel);
        return 1;
    }
    if (p.stab_drift_file[0] && strcmp(p.parallel, "serial")) {
        /* src/main.c:724-730 */
        fprintf(stderr, "ERROR: stab_drift_file currently supports parallel=serial only\n");
        return 1;
    }
    if (p.stab_drift_file[0]) {
        /* init_stab_drift_file (src/main.c:453-472) */
        FILE *fd = fopen(p.stab_drift_file, "w");
        if (!fd) {
            fprintf(stderr, "ERROR: failed to open stab_drift_file %s\n", p.stab_drift_file);
            return 1;
        }
        fprintf(fd, "# beta_index Ltr replica_id seed U dtau stab use_ph samples max_inf mean_inf max_tau max_sweep "
                    "failed\n");
        fclose(fd);
    }
    Lattice L;
    if (lattice_build(&L, &p)) {
        fprintf(stderr, "ERROR: invalid lattice\n");
        return 1;
    }
    if (!L.is_bipartite) {
        fprintf(stderr, "ERROR: the GPU engine runs the particle-hole symmetric path (bipartite lattice)\n");
        return 1;
    }
    {
        /* lattice_dump (src/lattice.c:184-201), written by every run
           (src/main.c:663-665) */
        FILE *fh = fopen("hopping_used.txt", "w");
        if (fh) {
            fprintf(fh, "%d\n", L.n);
            for (int i = 0; i < L.n; i++)
                for (int j = 0; j < L.n; j++)
                    fprintf(fh, "%.17g%c", L.t[i + (size_t)j * L.n], j == L.n - 1 ? '\n' : ' ');
            fclose(fh);
        } else {
            perror("hopping_used.txt");
        }
    }
    const int n = L.n;
    const double mu = p.U / 2.0;
    double *K = malloc(sizeof(double) * n * n), *eK = malloc(sizeof(double) * n * n),
           *eKi = malloc(sizeof(double) * n * n);
    memcpy(K, L.t, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) K[i + i * n] -= mu;
    if (expm_sym(n, K, -p.dtau, eK) || expm_sym(n, K, p.dtau, eKi)) {
        fprintf(stderr, "ERROR: eigendecomposition of K did not converge\n");
        return 1;
    }
    const double lambda = acosh(exp(p.dtau * p.U / 2.0));

    Plan szz, sperp;
    if (plan_init(&szz, &L, p.szz_q) || plan_init(&sperp, &L, p.sperp_q)) {
        fprintf(stderr, "ERROR: invalid szz_q / sperp_q selector\n");
        return 1;
    }
    /* DeltaSU2 = Sperp/2 - Szz needs the same momenta in both plans
       (plans_have_same_q, src/main.c:379-392, 807-826) */
    const int consistency = strcmp(p.spin_consistency_file, "none") != 0;
    if (consistency) {
        int same = szz.enabled && sperp.enabled && szz.nq == sperp.nq;
        for (int q = 0; same && q < szz.nq; q++) same = szz.mx[q] == sperp.mx[q] && szz.my[q] == sperp.my[q];
        if (!same) {
            fprintf(stderr,
                    "ERROR: invalid spin structure-factor setup szz_selector=%s szz_file=%s sperp_selector=%s "
                    "sperp_file=%s consistency_file=%s lattice=%s Lx=%d Ly=%d\n",
                    p.szz_q, p.szz_file, p.sperp_q, p.sperp_file, p.spin_consistency_file, p.lattice, L.Lx, L.Ly);
            return 1;
        }
    }
    double *phase;
    int nq;
    union_plans(&L, &szz, &sperp, &phase, &nq);

    int precision = DQMC_GPU_FP64;
    if (!strcmp(p.precision, "mixed")) precision = DQMC_GPU_MIXED;
    else if (!strcmp(p.precision, "mixed_fp32_chain")) precision = DQMC_GPU_MIXED_FP32_CHAIN;

    FILE *fo = fopen(p.output_file, "w");
    if (!fo) {
        perror(p.output_file);
        return 1;
    }
    fprintf(fo, "# lattice=%s n=%d U=%g mu=%g dtau=%g bipartite=%d green_rebuild=%s parallel=%s nrep=%d bins=%d",
            p.lattice, n, p.U, mu, p.dtau, L.is_bipartite, p.green_rebuild, p.parallel, p.nrep, p.nrep * p.nbin);
    if (szz.enabled) fprintf(fo, " szz_file=%s szz_nq=%d", p.szz_file, szz.nq);
    if (sperp.enabled) fprintf(fo, " sperp_file=%s sperp_nq=%d", p.sperp_file, sperp.nq);
    if (consistency) fprintf(fo, " spin_consistency_file=%s", p.spin_consistency_file);
    fprintf(fo, "\n# T  E_hub dE_hub  E_gc dE_gc  E_ph dE_ph  ntot dN  doublon dD  sign  acceptance dAcceptance\n");
    fflush(fo);
    FILE *fz = NULL, *fp = NULL;
    if (szz.enabled) {
        if (!(fz = fopen(p.szz_file, "w"))) {
            perror(p.szz_file);
            return 1;
        }
        fprintf(fz, "# definition=Szz(q)=N^-1 sum_ij exp[-iq.(ri-rj)] <Szi Szj>\n");
        fprintf(fz, "# spin_operator=Szi=(n_up-n_down)/2 factor3_applied=0\n");
        fprintf(fz, "# lattice=%s Lx=%d Ly=%d n=%d pbc=%d t=%.17g U=%.17g mu=%.17g dtau=%.17g\n", p.lattice, L.Lx,
                L.Ly, n, p.pbc, p.thop, p.U, mu, p.dtau);
        fprintf(fz, "# szz_q=%s seed=%llu parallel=%s nrep=%d bins=%d nbeta=%d\n", p.szz_q, p.seed, p.parallel,
                p.nrep, p.nrep * p.nbin, p.nbeta);
        fprintf(fz, "# beta_requested beta T q_index mx my qx_over_pi qy_over_pi qx_folded_over_pi "
                    "qy_folded_over_pi Szz dSzz\n");
    }
    if (sperp.enabled) {
        if (!(fp = fopen(p.sperp_file, "w"))) {
            perror(p.sperp_file);
            return 1;
        }
        fprintf(fp, "# definition=Sperp(q)=N^-1 sum_ij exp[-iq.(ri-rj)] <Sxi Sxj + Syi Syj>\n");
        fprintf(fp, "# ladder_definition=Sperp(q)=0.5*(S+-(q)+S-+(q)); Splus=Sx+iSy\n");
        fprintf(fp, "# su2_relation=ensemble Sperp(q)=2*Szz(q) at any dtau for the current SU(2)-invariant model\n");
        fprintf(fp, "# hs_caveat=individual spin-channel-HS samples select the z axis; equality is restored by "
                    "ensemble averaging\n");
        fprintf(fp, "# lattice=%s Lx=%d Ly=%d n=%d pbc=%d t=%.17g U=%.17g mu=%.17g dtau=%.17g\n", p.lattice, L.Lx,
                L.Ly, n, p.pbc, p.thop, p.U, mu, p.dtau);
        fprintf(fp, "# sperp_q=%s seed=%llu parallel=%s nrep=%d bins=%d nbeta=%d\n", p.sperp_q, p.seed,
                p.parallel, p.nrep, p.nrep * p.nbin, p.nbeta);
        fprintf(fp, "# beta_requested beta T q_index mx my qx_over_pi qy_over_pi qx_folded_over_pi "
                    "qy_folded_over_pi Sperp dSperp\n");
    }

    /* replica log (src/main.c:829-834, 917-933): replicas.dat by default
       when nrep > 1, replica_log=<path> always, replica_log=none never */
    FILE *fr = NULL;
    if (strcmp(p.replica_log, "none") != 0 && (p.replica_log[0] || p.nrep > 1)) {
        const char *path = p.replica_log[0] ? p.replica_log : "replicas.dat";
        if (!(fr = fopen(path, "w"))) {
            fprintf(stderr, "ERROR: failed to open replica_log %s\n", path);
            return 1;
        }
        fprintf(fr, "# beta T replica_id seed nwarm nmeas nbin status\n");
    }
    FILE *fc = NULL;
    if (consistency) {
        /* src/main.c:996-1021 */
        if (!(fc = fopen(p.spin_consistency_file, "w"))) {
            fprintf(stderr, "ERROR: failed to open spin_consistency_file %s\n", p.spin_consistency_file);
            return 1;
        }
        fprintf(fc, "# definition=DeltaSU2(q)=0.5*Sperp(q)-Szz(q)\n");
        fprintf(fc, "# error=paired jackknife over the same sign-reweighted replica/bin samples\n");
        fprintf(fc, "# lattice=%s Lx=%d Ly=%d n=%d pbc=%d t=%.17g U=%.17g mu=%.17g dtau=%.17g\n", p.lattice, L.Lx,
                L.Ly, n, p.pbc, p.thop, p.U, mu, p.dtau);
        fprintf(fc, "# szz_q=%s sperp_q=%s seed=%llu parallel=%s nrep=%d bins=%d nbeta=%d\n", p.szz_q, p.sperp_q,
                p.seed, p.parallel, p.nrep, p.nrep * p.nbin, p.nbeta);
        fprintf(fc, "# beta_requested beta T q_index mx my qx_over_pi qy_over_pi qx_folded_over_pi "
                    "qy_folded_over_pi DeltaSU2 dDeltaSU2\n");
    }

    const int nbin = p.nbin, nrep = p.nrep, total = nrep * nbin, per = p.nmeas / p.nbin;
    double *E[7];
    for (int k = 0; k < 7; k++) E[k] = malloc(sizeof(double) * total);
    double *zq = malloc(sizeof(double) * (size_t)(szz.enabled ? szz.nq : 1) * total);
    double *pq = malloc(sizeof(double) * (size_t)(sperp.enabled ? sperp.nq : 1) * total);
    double *b8 = malloc(sizeof(double) * nbin * 8), *bz = malloc(sizeof(double) * (size_t)nbin * nq),
           *bp = malloc(sizeof(double) * (size_t)nbin * nq), *vz = malloc(sizeof(double) * nq),
           *vp = malloc(sizeof(double) * nq);
    int *cnt = malloc(sizeof(int) * nbin);
    unsigned long long *seeds = malloc(sizeof(unsigned long long) * nrep);
    int *st = malloc(sizeof(int) * nrep), *why = malloc(sizeof(int) * nrep);

    for (int b = 0; b < p.nbeta; b++) {
        const double beta = p.beta_list[b];
        const int Ltr = (int)lround(beta / p.dtau);
        if (fabs(beta / p.dtau - (double)Ltr) > 1e-9) {
            fprintf(stderr, "ERROR: beta=%g not a multiple of dtau=%g (beta/dtau=%g)\n", beta, p.dtau, beta / p.dtau);
            return 1;
        }
        const double beta_eff = Ltr * p.dtau, T = 1.0 / beta_eff;
        /* stab_drift_file: the reference's drift (against a from-scratch G
           before each rebuild, dqmc_record_stab_drift, src/dqmc.c:64-98) */
        const double guard = p.drift_guard >= 0.0 ? p.drift_guard : (precision != DQMC_GPU_FP64 ? 1e-6 : 0.0);
        dqmc_gpu_opts o = {precision, 1, p.stab_drift_file[0] ? DQMC_GPU_DRIFT_SCRATCH : DQMC_GPU_DRIFT_OFF,
                           p.autotune, guard};
        dqmc_gpu_ctx *ctx = NULL;
        if (dqmc_gpu_create(&ctx, p.device, &o)) return fail("dqmc_gpu_create", NULL);
        int mid, first;
        if (dqmc_gpu_add_model(ctx, n, Ltr, p.stab, p.delay, eK, eKi, L.t, L.bipart, p.U, p.dtau, mu, lambda, L.Lx,
                               L.Ly, &mid))
            return fail("dqmc_gpu_add_model", ctx);
        for (int r = 0; r < nrep; r++) seeds[r] = replica_seed(p.seed, b, r);
        if (dqmc_gpu_add_replicas(ctx, mid, nrep, seeds, NULL, &first)) return fail("dqmc_gpu_add_replicas", ctx);
        if (dqmc_gpu_set_measure(ctx, mid, nq, phase, nbin)) return fail("dqmc_gpu_set_measure", ctx);
        if (dqmc_gpu_set_sweep_order(ctx, mid,
                                     strcmp(p.sweep_order, "alternating") ? DQMC_GPU_SWEEP_FORWARD
                                                                          : DQMC_GPU_SWEEP_ALTERNATING))
            return fail("dqmc_gpu_set_sweep_order", ctx);
        const int rebuild = !strcmp(p.green_rebuild, "centered")    ? DQMC_GPU_REBUILD_CENTERED
                            : !strcmp(p.green_rebuild, "two_sided") ? DQMC_GPU_REBUILD_TWO_SIDED
                                                                    : DQMC_GPU_REBUILD_COMBINE;
        if (dqmc_gpu_set_rebuild(ctx, mid, rebuild)) return fail("dqmc_gpu_set_rebuild", ctx);
        if (dqmc_gpu_init(ctx)) return fail("dqmc_gpu_init", ctx);
        if (dqmc_gpu_run(ctx, p.nwarm, per, nbin)) return fail("dqmc_gpu_run", ctx);
        if (dqmc_gpu_read_status(ctx, st, why, NULL, NULL, NULL)) return fail("dqmc_gpu_read_status", ctx);
        for (int r = 0; r < nrep; r++)
            if (st[r]) {
                fprintf(stderr, "ERROR: replica %d (seed %llu) failed at beta=%.17g, failure_reason=%d\n", r,
                        seeds[r], beta, why[r]);
                dqmc_gpu_destroy(ctx);
                return 1;
            }
        for (int r = 0; r < nrep; r++) {
            if (dqmc_gpu_read_bins(ctx, first + r, b8, cnt, bz, bp)) return fail("dqmc_gpu_read_bins", ctx);
            for (int bi = 0; bi < nbin; bi++) {
                const double *v = b8 + 8 * bi;
                const int out = r * nbin + bi;
                if (cnt[bi] <= 0 || v[5] == 0.0 || !isfinite(v[5])) {
                    fprintf(stderr, "ERROR: invalid bin at beta=%.17g replica=%d bin=%d count=%d sum_sign=%.17g\n",
                            beta, r, bi, cnt[bi], v[5]);
                    dqmc_gpu_destroy(ctx);
                    return 1;
                }
                for (int k = 0; k < 5; k++) E[k][out] = v[k] / v[5];
                E[5][out] = v[5] / (double)cnt[bi];
                E[6][out] = v[7] > 0.0 ? v[6] / v[7] : 0.0;
                for (int q = 0; q < nq; q++) {
                    vz[q] = bz[(size_t)bi * nq + q] / v[5];
                    vp[q] = bp[(size_t)bi * nq + q] / v[5];
                }
                if (szz.enabled) {
                    double *sv = malloc(sizeof(double) * szz.nq);
                    for (int q = 0; q < szz.nq; q++) sv[q] = vz[szz.col[q]];
                    if (!sum_rule_ok(&szz, &L, sv, E[3][out], E[4][out], 0.25)) {
                        fprintf(stderr, "ERROR: Szz sum rule failed at beta=%.17g replica=%d bin=%d\n", beta, r, bi);
                        dqmc_gpu_destroy(ctx);
                        return 1;
                    }
                    for (int q = 0; q < szz.nq; q++) zq[(size_t)q * total + out] = sv[q];
                    free(sv);
                }
                if (sperp.enabled) {
                    double *sv = malloc(sizeof(double) * sperp.nq);
                    for (int q = 0; q < sperp.nq; q++) sv[q] = vp[sperp.col[q]];
                    if (!sum_rule_ok(&sperp, &L, sv, E[3][out], E[4][out], 0.5)) {
                        fprintf(stderr, "ERROR: Sperp sum rule failed at beta=%.17g replica=%d bin=%d\n", beta, r,
                                bi);
                        dqmc_gpu_destroy(ctx);
                        return 1;
                    }
                    for (int q = 0; q < sperp.nq; q++) pq[(size_t)q * total + out] = sv[q];
                    free(sv);
                }
            }
        }
        if (p.stab_drift_file[0]) {
            /* append_stab_drift_row (src/replica_run.c:58-88), one row per
               replica */
            double *dmx = malloc(sizeof(double) * nrep), *dmn = malloc(sizeof(double) * nrep);
            unsigned long long *dct = malloc(sizeof(unsigned long long) * nrep),
                               *dsw = malloc(sizeof(unsigned long long) * nrep);
            int *dtau = malloc(sizeof(int) * nrep);
            FILE *fd = fopen(p.stab_drift_file, "a");
            if (!fd || dqmc_gpu_read_drift(ctx, dmx, dmn, dct) || dqmc_gpu_read_drift_at(ctx, dtau, dsw)) {
                fprintf(stderr, "ERROR: failed to write stab_drift_file %s\n", p.stab_drift_file);
                if (fd) fclose(fd);
                dqmc_gpu_destroy(ctx);
                return 1;
            }
            for (int r = 0; r < nrep; r++)
                fprintf(fd, "%d %d %d %llu %.17g %.17g %d %d %llu %.17g %.17g %d %llu %d\n", b, Ltr, r, seeds[r], p.U,
                        p.dtau, p.stab, 1, dct[first + r], dmx[first + r], dmn[first + r], dtau[first + r],
                        dsw[first + r], 0);
            fclose(fd);
            free(dmx); free(dmn); free(dct); free(dsw); free(dtau);
        }
        dqmc_gpu_destroy(ctx);
        double m[7], e[7];
        for (int k = 0; k < 7; k++) {
            jackknife(E[k], total, &m[k], &e[k]);
            if (!isfinite(m[k]) || !isfinite(e[k])) {
                fprintf(stderr, "ERROR: non-finite jackknife output at T=%.17g\n", T);
                return 1;
            }
        }
        fprintf(fo, "%.6g %.8g %.3g %.8g %.3g %.8g %.3g %.6g %.2g %.8g %.2g %.6g %.8g %.3g\n", T, m[0], e[0], m[1],
                e[1], m[2], e[2], m[3], e[3], m[4], e[4], m[5], m[6], e[6]);
        fflush(fo);
        if (szz.enabled) write_q(fz, &szz, &L, beta, beta_eff, T, zq, total);
        if (sperp.enabled) write_q(fp, &sperp, &L, beta, beta_eff, T, pq, total);
        if (fc) {
            /* write_spin_consistency_beta (src/main.c:319-377): paired
               jackknife of Sperp/2 - Szz over the same bins */
            double *dq = malloc(sizeof(double) * (size_t)szz.nq * total);
            for (size_t k = 0; k < (size_t)szz.nq * total; k++) dq[k] = 0.5 * pq[k] - zq[k];
            write_q(fc, &szz, &L, beta, beta_eff, T, dq, total);
            free(dq);
        }
        if (fr) {
            /* src/main.c:1706-1713 */
            for (int r = 0; r < nrep; r++)
                fprintf(fr, "%.17g %.17g %d %llu %d %d %d %s\n", beta, T, r, seeds[r], p.nwarm, p.nmeas, p.nbin,
                        st[r] ? "error" : "ok");
            fflush(fr);
        }
    }
    fclose(fo);
    if (fz) fclose(fz);
    if (fp) fclose(fp);
    if (fc) fclose(fc);
    if (fr) fclose(fr);
    plan_free(&szz);
    plan_free(&sperp);
    lattice_release(&L);
    return 0;
}
