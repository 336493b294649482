/*
 * dqmc_gpu.h -- C ABI of the MI355X DQMC Green's-function engine.
 *
 * The engine advances many independent replicas of the half-filled Hubbard
 * model (forward BSS sweep, particle-hole symmetric path: only G_up is
 * propagated, G_dn = I - P G_up^T P) on one AMD Instinct MI355X (gfx950).
 * It replaces, for a batch of replicas, the block of dqmc_run_replica
 * (src/replica_run.c:124-426) that calls dqmc_init_modes, dqmc_sweep,
 * measure_sample and measure_szz_sperp_sample.  Host code keeps the input
 * parser, seeds, field initialisation order, jackknife and writers.
 *
 * Conventions (kept from the reference):
 *   - every function returns int 0 = ok, nonzero = failure
 *     (src/linalg.h:48-53, src/dqmc.h:79-81); dqmc_gpu_last_error() gives
 *     a message for the calling thread;
 *   - matrices are column-major n x n doubles, A[i + j*n] (src/linalg.h:4);
 *   - caller-owned host memory is copied, device memory is owned by the
 *     context;
 *   - per-replica status is sticky: after a numerical failure the replica's
 *     sweeps are no-ops (src/dqmc.c:611-614) and the failure reason uses
 *     the LINALG_FAILURE_* numbering (src/linalg.h:48-53) extended below;
 *   - a context is not thread-safe: one host thread per device.
 */
#ifndef DQMC_GPU_H
#define DQMC_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Precision modes. */
#define DQMC_GPU_FP64 0  /* every stage in FP64 */
#define DQMC_GPU_MIXED 1 /* wrap and delayed updates in FP32 on MFMA;
                            B-chains, QR, rebuild, sign, measurement FP64 */
#define DQMC_GPU_MIXED_FP32_CHAIN 2 /* as MIXED, and the B-chain products
                            B_{l+s-1}...B_l also in FP32 */

/* Failure reasons: 0..3 as LINALG_FAILURE_* (src/linalg.h:48-53). */
#define DQMC_GPU_FAIL_NONE 0
/* LINALG_FAILURE_* of the centered rebuild (src/linalg.h:48-53) */
#define DQMC_GPU_FAIL_CENTERED_RADIUS 1        /* log|D| range too wide to recentre */
#define DQMC_GPU_FAIL_CENTERED_UPDATE_MARGIN 2 /* the update would leave the double range */
#define DQMC_GPU_FAIL_EFFECTIVE_LOG_RANGE 3    /* e^-|log|D| + offset| underflows */
#define DQMC_GPU_FAIL_NONFINITE 16 /* non-finite matrix at a stage */
#define DQMC_GPU_FAIL_ZERO_D 17    /* zero diagonal in a QR (D_i == 0) */
#define DQMC_GPU_FAIL_SIGN 18      /* det sign evaluated to 0 */
#define DQMC_GPU_FAIL_DRIFT 19     /* wrap-vs-rebuild drift above opts.drift_guard */

/* Steps for dqmc_gpu_step (per-kernel golden tests). */
#define DQMC_GPU_STEP_PROPOSE_SLICE 0 /* Metropolis over all sites of slice arg */
#define DQMC_GPU_STEP_WRAP 1          /* G <- B_arg G B_arg^-1 */
#define DQMC_GPU_STEP_REBUILD 2       /* G(tau=arg) from scratch (L lmuls) */

typedef struct dqmc_gpu_ctx dqmc_gpu_ctx;

/* Stabilisation drift recording (dqmc_gpu_opts.drift). */
#define DQMC_GPU_DRIFT_OFF 0
#define DQMC_GPU_DRIFT_SCRATCH 1 /* max|G_wrap - G_scratch(tau)| against a
                                    from-scratch G of the current fields before
                                    each rebuild, as dqmc_record_stab_drift
                                    (src/dqmc.c:55-98): L/stab more UDV updates
                                    per rebuild */
#define DQMC_GPU_DRIFT_REBUILD 2 /* max|G_wrap - G_rebuilt| against the G the
                                    rebuild itself forms (no counterpart in the
                                    reference; no extra work) */

typedef struct {
    int precision;   /* DQMC_GPU_FP64, DQMC_GPU_MIXED or DQMC_GPU_MIXED_FP32_CHAIN */
    int use_graph;   /* capture one sweep per model as a hipGraph */
    int drift;       /* DQMC_GPU_DRIFT_* */
    int autotune;    /* time the Metropolis site segments and delay window of
                        each model on its first sweeps and keep the fastest */
    double drift_guard; /* > 0: a replica whose max|G_wrap - G_rebuilt| at a
                        rebuild exceeds it fails with DQMC_GPU_FAIL_DRIFT
                        (sticky); required (> 0) by the mixed modes */
} dqmc_gpu_opts;

/* Create a context on HIP device `device`.  `o` may be NULL (FP64, graphs
   on, drift off, no guard).  The mixed modes are refused without a drift
   guard: their FP32 wrap and delayed updates are not accurate enough for
   every workload, and the guard turns a replica whose G has drifted into a
   sticky failure instead of a silently wrong Green's function. */
int dqmc_gpu_create(dqmc_gpu_ctx **out, int device, const dqmc_gpu_opts *o);

/* Register one model (lattice, U, dtau, L = beta/dtau slices).  expK and
   expKinv are exp(-dtau K) and exp(+dtau K) with K = t - mu I, as
   model_init computes them (src/model.c:8-31); thop is the hopping matrix
   (src/lattice.c) used by the energy, bipart the sublattice signs,
   lambda = acosh(exp(dtau U / 2)) (src/field.c:9).  stab is the
   stabilisation interval (src/green.c:168-184), delay_k the delayed-update
   depth (the reference's GREEN_DELAY_CAPACITY, src/green.c:10); delay_k <= 0
   lets the engine choose.  Lx, Ly give the geometry for the structure
   factors (Ly = 1 for a chain).  Replaces model_init + green_alloc +
   green_stack_alloc (src/model.c:8-31, src/green.c:54-86, 168-184). */
int dqmc_gpu_add_model(dqmc_gpu_ctx *ctx, int n, int L, int stab,
                       int delay_k, const double *expK, const double *expKinv,
                       const double *thop, const int *bipart, double U,
                       double dtau, double mu, double lambda, int Lx, int Ly,
                       int *model_id);

/* Add nrep replicas of a model.  With fields == NULL each replica's field
   is drawn from its seed exactly as rng_seed + field_init do
   (src/rng.c:16-37, src/field.c:20-22: L*n draws, s = draw < 0.5 ? -1 : 1);
   otherwise fields holds nrep * L * n values of +-1 and the RNG stream
   starts right after the L*n field draws of each seed.  Replica ids are
   consecutive from *first_replica_id. */
int dqmc_gpu_add_replicas(dqmc_gpu_ctx *ctx, int model_id, int nrep,
                          const unsigned long long *seeds,
                          const signed char *fields, int *first_replica_id);

/* Momenta and bins for the measurement of a model: phase_cos[disp + q*n]
   as structure_factor_plan_init builds it (src/structure_factor.c:185-203);
   nq = 0 disables Szz/Sperp (scalars are always measured). */
int dqmc_gpu_set_measure(dqmc_gpu_ctx *ctx, int model_id, int nq,
                         const double *phase_cos, int nbin);

/* Sweep order of a model (the sweep_order key, src/io.c:284-286, 374-380;
   dqmc_init_modes' mode, src/dqmc.c:419-462): DQMC_GPU_SWEEP_FORWARD (the
   default) or DQMC_GPU_SWEEP_ALTERNATING, forward and backward sweeps in
   turn, each reusing the stack the previous one stored
   (dqmc_sweep_forward with carry_prefix / dqmc_sweep_backward,
   src/dqmc.c:609-963).  Setting it restarts the alternation with a forward
   sweep that rebuilds both stacks. */
#define DQMC_GPU_SWEEP_FORWARD 0
#define DQMC_GPU_SWEEP_ALTERNATING 1
int dqmc_gpu_set_sweep_order(dqmc_gpu_ctx *ctx, int model_id, int order);

/* Boundary rebuild of a model (the green_rebuild key, src/io.c:287-289;
   GreenRebuildMode, src/green.h:10-14), before dqmc_gpu_init.  COMBINE and
   TWO_SIDED both run the one-QR two-sided inverse (the reference's
   two_sided formula, src/linalg.c:1117-1389); CENTERED in addition keeps
   every UDV's D recentred around a per-replica log offset
   (udv_lmul_centered_work / udv_rmul_centered / udv_recenter,
   src/linalg.c:453-794), so that D may exceed the double range at low
   temperature, and fails a replica with DQMC_GPU_FAIL_CENTERED_* or
   _EFFECTIVE_LOG_RANGE as the reference does. */
#define DQMC_GPU_REBUILD_COMBINE 0
#define DQMC_GPU_REBUILD_TWO_SIDED 1
#define DQMC_GPU_REBUILD_CENTERED 2
int dqmc_gpu_set_rebuild(dqmc_gpu_ctx *ctx, int model_id, int mode);

/* Initial stabilised Green's function at tau = 0 for every replica
   (dqmc_init_modes -> green_from_scratch(Gu, 0), src/dqmc.c:418-483). */
int dqmc_gpu_init(dqmc_gpu_ctx *ctx);

/* nsweep sweeps of every replica in each model's sweep order (dqmc_sweep,
   src/dqmc.c:952-963).
   If bin >= 0, each sweep is followed by the tau = 0 measurement
   (measure_sample + measure_szz_sperp_sample) added with the replica's sign
   into bin `bin`, plus the sweep's acceptance deltas
   (src/replica_run.c:288-451). */
int dqmc_gpu_sweep(dqmc_gpu_ctx *ctx, int nsweep, int bin);

/* nwarm sweeps, then nbin * nmeas_per_bin measured sweeps
   (src/replica_run.c:184-186, 284-451). */
int dqmc_gpu_run(dqmc_gpu_ctx *ctx, int nwarm, int nmeas_per_bin, int nbin);

/* Bins of one replica.  bins8[nbin*8] packs each ReplicaBin as
   replica_mpi_pack_bins does (src/replica_mpi.c:102-126): sum sign*E_hub,
   sum sign*E_gc, sum sign*E_ph, sum sign*N, sum sign*D, sum sign,
   accepted, attempts; counts[nbin]; szz/sperp [nbin*nq] hold
   sum sign*S(q) (src/replica.c:240-269).  Any output pointer may be NULL. */
int dqmc_gpu_read_bins(dqmc_gpu_ctx *ctx, int replica, double *bins8,
                       int *counts, double *szz, double *sperp);

/* Per-replica status (0 ok), failure reason, accepted / attempted flips
   (src/dqmc.h:79-86) and sign.  Arrays hold one entry per replica. */
int dqmc_gpu_read_status(dqmc_gpu_ctx *ctx, int *status, int *failure_reason,
                         unsigned long long *accepted,
                         unsigned long long *attempts, double *sign);

/* Stabilisation drift per replica: maximum, mean and sample count of the
   samples recorded under opts.drift (dqmc_record_stab_drift,
   src/dqmc.c:64-98; DQMC_GPU_DRIFT_SCRATCH is the reference's quantity). */
int dqmc_gpu_read_drift(dqmc_gpu_ctx *ctx, double *max_inf, double *mean_inf,
                        unsigned long long *samples);
/* Where each replica's largest drift sample was taken: the slice tau of
   the rebuild and the number of sweeps completed before it (max_tau,
   max_sweep of DqmcStabDrift, src/dqmc.h and src/dqmc.c:90-95; the
   stab_drift_file columns, src/replica_run.c:58-88). */
int dqmc_gpu_read_drift_at(dqmc_gpu_ctx *ctx, int *max_tau,
                           unsigned long long *max_sweep);

/* Launch shape of a model: Metropolis delay window (sites; 0 for the FP64
   submatrix path, which decides each segment on its submatrix of G), site
   segments per slice (each ending in a rank-P flush) and whether the
   B-matrix products run as separable 16 x 16 transforms (1) or dense GEMMs
   (0): the engine's defaults, or the autotuner's choice after
   dqmc_gpu_init with opts.autotune (no counterpart in the reference, whose
   delay capacity is the fixed GREEN_DELAY_CAPACITY, src/green.c:10). */
int dqmc_gpu_read_tuning(dqmc_gpu_ctx *ctx, int model_id, int *window,
                         int *segments, int *separable);

/* Ratio R_up R_dn = (1 + (1 - G_ii) N_up)(1 + G_ii N_dn) of every site at
   the latest Metropolis slice of a replica (n values in site order), as
   the proposal loop of src/dqmc.c:656-679 evaluates it
   (green_delay_ratio, src/green.c:401-416; the reference keeps no log of
   it).  A parity hook: the model must have been added with
   DQMC_GPU_RATIO_LOG=1 in the environment, otherwise it returns an error. */
int dqmc_gpu_read_ratios(dqmc_gpu_ctx *ctx, int replica, double *ratios);

/* Parity and debug hooks. */
int dqmc_gpu_read_green(dqmc_gpu_ctx *ctx, int replica, double *g);
int dqmc_gpu_write_green(dqmc_gpu_ctx *ctx, int replica, const double *g);
int dqmc_gpu_read_field(dqmc_gpu_ctx *ctx, int replica, signed char *s);
int dqmc_gpu_write_field(dqmc_gpu_ctx *ctx, int replica, const signed char *s);
int dqmc_gpu_read_rng(dqmc_gpu_ctx *ctx, int replica, unsigned long long *s4);
/* One stage for every replica of every model (DQMC_GPU_STEP_*). */
int dqmc_gpu_step(dqmc_gpu_ctx *ctx, int step, int arg);

/* Wall time of the last dqmc_gpu_sweep / dqmc_gpu_run call measured with
   HIP events on the engine's streams, milliseconds. */
int dqmc_gpu_last_elapsed_ms(dqmc_gpu_ctx *ctx, double *ms);

/* Stage classes of dqmc_gpu_profile_sweep. */
#define DQMC_GPU_PROF_METRO 0   /* Metropolis + delayed updates, one slice */
#define DQMC_GPU_PROF_WRAP 1    /* G <- B G B^-1 */
#define DQMC_GPU_PROF_STACK 2   /* one suffix-stack entry (B-chain + rmul) */
#define DQMC_GPU_PROF_LMUL 3    /* left B-chain + lmul */
#define DQMC_GPU_PROF_COMBINE 4 /* udv_combine */
#define DQMC_GPU_PROF_INV 5     /* udv_inv_one_plus */
#define DQMC_GPU_PROF_ACCEPT 6  /* drift bookkeeping and copy of the rebuilt G */
#define DQMC_GPU_PROF_MEASURE 7 /* tau = 0 measurement into bin 0 */
#define DQMC_GPU_PROF_QR 8      /* every QR (also counted in its stage) */
#define DQMC_GPU_PROF_TRSM 9    /* every triangular solve (also in INV) */
#define DQMC_GPU_PROF_GEMM 10   /* every n x n x n GEMM launch (also in its stage) */
#define DQMC_GPU_PROF_FLUSH 11  /* every delayed-update flush launch (also in METRO) */
#define DQMC_GPU_PROF_PROPOSE 12 /* every Metropolis proposal launch (also in METRO) */
#define DQMC_GPU_PROF_PERSIST 13 /* one launch of the persistent sweep kernel (whole sweeps, n <= 64) */
#define DQMC_GPU_PROF_N 14

/* Profiling aid without a counterpart in the reference (which uses its
   PROF_BEGIN/PROF_END region timers, src/profiler.h:93-94): runs nsweep
   sweeps without graph capture (plus a measurement into bin 0 when bins are
   configured), timing every stage with HIP events on its stream.
   ms[DQMC_GPU_PROF_N] gets the summed milliseconds per class and
   launches[DQMC_GPU_PROF_N] the number of timed stage invocations. */
int dqmc_gpu_profile_sweep(dqmc_gpu_ctx *ctx, int nsweep, double *ms, int *launches);

void dqmc_gpu_destroy(dqmc_gpu_ctx *ctx);

/* ---- batched linear-algebra entry points on host arrays ----
   Each takes `batch` independent problems laid out back to back
   (stride n*n for matrices, n for vectors) and runs them on `device`,
   1 <= n <= 512.  They replace the reference's linalg.c calls one for one. */

/* udv_lmul_work (src/linalg.c:515-642), in place. */
int dqmc_gpu_udv_lmul(int device, int n, int batch, double *U, double *D,
                      double *T, const double *B);
/* udv_rmul (src/linalg.c:662-784), in place. */
int dqmc_gpu_udv_rmul(int device, int n, int batch, double *U, double *D,
                      double *T, const double *B);
/* udv_combine (src/linalg.c:796-909). */
int dqmc_gpu_udv_combine(int device, int n, int batch, const double *Ul,
                         const double *Dl, const double *Tl, const double *Ur,
                         const double *Dr, const double *Tr, double *Uo,
                         double *Do, double *To);
/* udv_inv_one_plus_work (src/linalg.c:911-1115); sign[batch]. */
int dqmc_gpu_udv_inv_one_plus(int device, int n, int batch, const double *U,
                              const double *D, const double *T, double *g,
                              int *sign);
/* udv_inv_one_plus_work on UDVs that carry a log offset each (the centered
   rebuild's UDV.log_offset: the diagonal is D e^log_offset; src/linalg.c:
   911-1115 with the offset-aware split of src/linalg.c:166-182, 963-986).
   log_offset[batch]; reason[batch] (may be NULL) gets each entry's failure
   reason (0 = ok; DQMC_GPU_FAIL_EFFECTIVE_LOG_RANGE when e^-|log|D_i| + off|
   underflows, tests/test_udv.c:199-209).  Returns nonzero if any entry
   failed; g and sign are then not written. */
int dqmc_gpu_udv_inv_one_plus_off(int device, int n, int batch, const double *U,
                                  const double *D, const double *T,
                                  const double *log_offset, double *g,
                                  int *sign, int *reason);
/* measure_szz_sperp_sample + measure_sample on G_up (G_dn by the PH map),
   (src/structure_factor.c:328-400, src/measure.c:6-39).  scal5[batch*5] =
   {E, ekin, eint, ntot, doublon}; szz, sperp [batch*nq]. */
int dqmc_gpu_measure(int device, int Lx, int Ly, int batch, const double *gu,
                     const double *thop, const int *bipart, double U, int nq,
                     const double *phase_cos, double *scal5, double *szz,
                     double *sperp);

/* No counterpart in the reference: dqmc_gpu_udv_inv_one_plus for callers
   that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc).  U, D, T (inputs), g
   (output, batch * n * n), sign (output, batch ints: the det sign, 0 for an
   entry that failed) and reason (output, batch ints or NULL: the failure
   reason, 0 = ok) are DEVICE pointers owned by the caller; the work is
   enqueued on `stream` (a hipStream_t, NULL for the null stream) and the
   call returns without waiting for it.  Nothing is allocated: the workspace
   is a static device array of 64 MiB per device, held until the process
   ends, which bounds the problem (about 4 n^2 + 6 n + 8192 doubles per
   batch entry: batch <= 29 at n = 256, <= 460 at n = 64); larger calls are
   refused with an error.  All calls on one device share that workspace:
   calls that overlap on one device race. */
int dqmc_gpu_udv_inv_one_plus_noalloc(int n, int batch, const double *U,
                                      const double *D, const double *T,
                                      double *g, int *sign, int *reason,
                                      void *stream);

const char *dqmc_gpu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
