/*
 * host.h -- C host layer of the GPU DQMC driver: input parameters,
 * lattice and model set-up, replica seeds, bins, jackknife.
 *
 * This is the part of the program that stays on the CPU: it reads the same
 * key=value input as the reference (src/io.c), builds the same lattice and
 * model (src/lattice.c, src/model.c), derives the same per-replica seeds
 * (src/replica.c:17-34) and writes the same observables.dat / szz.dat /
 * sperp.dat (src/main.c:221-317), while every sweep and measurement runs in
 * libdqmc_gpu through include/dqmc_gpu.h.
 */
#ifndef SAIQMC_GPU_HOST_H
#define SAIQMC_GPU_HOST_H

#include <stdint.h>

typedef struct {
    char lattice[16];
    int Lx, Ly, pbc;
    double thop, U, dtau;
    double beta_list[64];
    int nbeta;
    int nwarm, nmeas, nbin, stab;
    char output_file[256];
    char parallel[16];
    int nrep;
    char szz_q[256], szz_file[256], sperp_q[256], sperp_file[256];
    char stab_drift_file[256];
    unsigned long long seed;
    char replica_log[256];           /* "" = replicas.dat when nrep > 1, "none" = off */
    char spin_consistency_file[256]; /* "none" = off */
    char sweep_order[16];   /* forward | alternating */
    char green_rebuild[16]; /* combine | two_sided (centered is not run) */
    /* engine keys without a counterpart in the reference */
    char precision[24]; /* fp64 | mixed | mixed_fp32_chain */
    int delay;          /* delayed-update depth, 0 = engine default */
    int device;
    double drift_guard; /* fail a replica whose G drifts further from the
                           rebuilt one; < 0 = engine default (1e-6 in the
                           mixed modes, off in fp64) */
    int autotune;       /* 1 = time the Metropolis launch shapes per model */
} Params;

/* key=value parser with the reference's defaults and checks
   (src/io.c:71-103, 156-418). Returns 0 on success. */
int params_read(Params *p, const char *path);

typedef struct {
    int n, Lx, Ly, is_bipartite, is_chain;
    double *t;   /* n x n hopping, column-major */
    int *bipart; /* sublattice signs */
} Lattice;

/* lattice_chain / lattice_square (src/lattice.c:26-82). */
int lattice_build(Lattice *L, const Params *p);
void lattice_release(Lattice *L);

/* exp(scale * A) of a symmetric matrix by Taylor series with scaling and
   squaring (the reference uses LAPACK dsyev in la_expm_sym,
   src/linalg.c:319-346). Returns 0 on success. */
int expm_sym(int n, const double *A, double scale, double *out);

/* replica_seed (src/replica.c:17-34). */
unsigned long long replica_seed(unsigned long long base, int beta_index,
                                int replica_id);

/* jackknife (src/measure.c:48-69). */
void jackknife(const double *x, int N, double *mean, double *err);

#endif
