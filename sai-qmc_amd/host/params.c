/*
 * params.c -- key=value input for the GPU driver.
 *
 * Same file format, keys, defaults and validation as the reference's
 * params_read (src/io.c:71-103 defaults, 156-418 parsing and checks):
 * '#' starts a comment, one key=value per line, unknown keys are errors.
 * Keys that select reference code paths the GPU engine does not implement
 * (lattice=file, profile=1, the UDV diagnostics files) are accepted only
 * at their default values.
 * Additional keys: precision=fp64|mixed|mixed_fp32_chain, delay=<k>, device=<id>,
 * drift_guard=<bound> (a replica whose G drifts further from the rebuilt G
 * fails; default 1e-6 in the mixed modes, which require it, off in fp64),
 * autotune=0|1 (time the Metropolis launch shapes of each model).
 */
#include "host.h"

#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void defaults(Params *p)
{
    memset(p, 0, sizeof(*p));
    strcpy(p->lattice, "chain");
    p->Lx = 4;
    p->Ly = 1;
    p->pbc = 1;
    p->thop = -1.0;
    p->U = 4.0;
    p->dtau = 0.1;
    p->nwarm = 200;
    p->nmeas = 2000;
    p->nbin = 20;
    p->stab = 8;
    strcpy(p->output_file, "observables.dat");
    strcpy(p->parallel, "serial");
    p->nrep = 1;
    strcpy(p->szz_q, "none");
    strcpy(p->szz_file, "szz.dat");
    strcpy(p->sperp_q, "none");
    strcpy(p->sperp_file, "sperp.dat");
    p->seed = 12345ULL;
    strcpy(p->precision, "fp64");
    strcpy(p->spin_consistency_file, "none"); /* src/io.c:96-101 */
    strcpy(p->sweep_order, "forward");  /* src/io.c:92-93 */
    strcpy(p->green_rebuild, "combine");
    p->delay = 0;
    p->device = 0;
    p->drift_guard = -1.0;
    p->autotune = 0;
}

static char *trim(char *s)
{
    while (isspace((unsigned char)*s)) s++;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char)e[-1])) *--e = '\0';
    return s;
}

static int to_int(const char *v, int *out)
{
    char *end = NULL;
    errno = 0;
    const long x = strtol(v, &end, 10);
    if (errno || end == v || *end != '\0' || x < INT_MIN || x > INT_MAX) return 1;
    *out = (int)x;
    return 0;
}

static int to_double(const char *v, double *out)
{
    char *end = NULL;
    errno = 0;
    const double x = strtod(v, &end);
    if (errno || end == v || *end != '\0') return 1;
    *out = x;
    return 0;
}

static int to_ull(const char *v, unsigned long long *out)
{
    char *end = NULL;
    errno = 0;
    if (v[0] == '-') return 1;
    const unsigned long long x = strtoull(v, &end, 10);
    if (errno || end == v || *end != '\0') return 1;
    *out = x;
    return 0;
}

static int copy_str(char *dst, size_t cap, const char *v)
{
    if (strlen(v) == 0 || strlen(v) >= cap) return 1;
    strcpy(dst, v);
    return 0;
}

static int beta_list(Params *p, char *v)
{
    p->nbeta = 0;
    for (char *tok = strtok(v, ","); tok; tok = strtok(NULL, ",")) {
        if (p->nbeta >= 64) return 2;
        double b;
        if (to_double(trim(tok), &b) || !(b > 0.0)) return 1;
        p->beta_list[p->nbeta++] = b;
    }
    return p->nbeta > 0 ? 0 : 1;
}

/* keys of reference code paths this engine does not run: only the default */
static int default_only(const char *key, const char *val)
{
    static const char *fixed[][2] = {
        {"profile", "0"},
    };
    for (size_t i = 0; i < sizeof fixed / sizeof fixed[0]; i++)
        if (strcmp(key, fixed[i][0]) == 0) {
            if (strcmp(val, fixed[i][1]) != 0) {
                fprintf(stderr, "ERROR: %s=%s is not supported by the GPU engine (only %s)\n", key, val,
                        fixed[i][1]);
                return -1;
            }
            return 1;
        }
    static const char *unsupported[] = {"latfile", "profile_file", "udv_scale_file", "udv_scale_diagnostics_file",
                                        "udv_centered_file", "udv_centered_diagnostics_file"};
    for (size_t i = 0; i < sizeof unsupported / sizeof unsupported[0]; i++)
        if (strcmp(key, unsupported[i]) == 0) {
            fprintf(stderr, "ERROR: %s is not supported by the GPU engine\n", key);
            return -1;
        }
    return 0;
}

int params_read(Params *p, const char *path)
{
    defaults(p);
    FILE *fp = fopen(path, "r");
    if (!fp) {
        perror(path);
        return 1;
    }
    char line[512];
    int ln = 0, rc = 0;
    while (!rc && fgets(line, sizeof line, fp)) {
        ln++;
        if (!strchr(line, '\n') && !feof(fp)) {
            fprintf(stderr, "ERROR: input line %d exceeds %zu bytes\n", ln, sizeof line - 1);
            rc = 1;
            break;
        }
        char *h = strchr(line, '#');
        if (h) *h = '\0';
        char *eq = strchr(line, '=');
        if (!eq) {
            if (*trim(line)) {
                fprintf(stderr, "ERROR: line %d is not key=value\n", ln);
                rc = 1;
            }
            continue;
        }
        *eq = '\0';
        char *key = trim(line), *val = trim(eq + 1);
        if (!*key || !*val) {
            fprintf(stderr, "ERROR: invalid %s value on line %d\n", key, ln);
            rc = 1;
            break;
        }
        const int d = default_only(key, val);
        if (d < 0) {
            rc = 1;
            break;
        }
        if (d > 0) continue;
#define BAD(...)                          \
    do {                                  \
        fprintf(stderr, __VA_ARGS__);     \
        rc = 1;                           \
    } while (0)
        if (!strcmp(key, "lattice")) {
            if (copy_str(p->lattice, sizeof p->lattice, val)) BAD("ERROR: invalid lattice\n");
        } else if (!strcmp(key, "Lx")) {
            if (to_int(val, &p->Lx)) BAD("ERROR: Lx must be integer (got %s)\n", val);
        } else if (!strcmp(key, "Ly")) {
            if (to_int(val, &p->Ly)) BAD("ERROR: Ly must be integer (got %s)\n", val);
        } else if (!strcmp(key, "pbc")) {
            if (to_int(val, &p->pbc)) BAD("ERROR: pbc must be 0 or 1 (got %s)\n", val);
        } else if (!strcmp(key, "bc")) {
            if (!strcmp(val, "periodic") || !strcmp(val, "pbc") || !strcmp(val, "1")) p->pbc = 1;
            else if (!strcmp(val, "open") || !strcmp(val, "obc") || !strcmp(val, "0")) p->pbc = 0;
            else BAD("ERROR: bc must be periodic/open/pbc/obc/1/0 (got %s)\n", val);
        } else if (!strcmp(key, "t")) {
            if (to_double(val, &p->thop)) BAD("ERROR: t must be numeric (got %s)\n", val);
        } else if (!strcmp(key, "U")) {
            if (to_double(val, &p->U)) BAD("ERROR: U must be numeric (got %s)\n", val);
        } else if (!strcmp(key, "dtau")) {
            if (to_double(val, &p->dtau)) BAD("ERROR: dtau must be numeric (got %s)\n", val);
        } else if (!strcmp(key, "nwarm")) {
            if (to_int(val, &p->nwarm)) BAD("ERROR: nwarm must be integer (got %s)\n", val);
        } else if (!strcmp(key, "nmeas")) {
            if (to_int(val, &p->nmeas)) BAD("ERROR: nmeas must be integer (got %s)\n", val);
        } else if (!strcmp(key, "nbin")) {
            if (to_int(val, &p->nbin)) BAD("ERROR: nbin must be integer (got %s)\n", val);
        } else if (!strcmp(key, "stab") || !strcmp(key, "stabilize_interval")) {
            if (to_int(val, &p->stab)) BAD("ERROR: stab must be integer (got %s)\n", val);
        } else if (!strcmp(key, "output_file")) {
            if (copy_str(p->output_file, sizeof p->output_file, val)) BAD("ERROR: invalid output_file\n");
        } else if (!strcmp(key, "stab_drift_file") || !strcmp(key, "stabilization_drift_file")) {
            if (copy_str(p->stab_drift_file, sizeof p->stab_drift_file, val)) BAD("ERROR: invalid %s\n", key);
        } else if (!strcmp(key, "parallel")) {
            if (copy_str(p->parallel, sizeof p->parallel, val)) BAD("ERROR: invalid parallel\n");
        } else if (!strcmp(key, "nrep")) {
            if (to_int(val, &p->nrep)) BAD("ERROR: nrep must be integer (got %s)\n", val);
        } else if (!strcmp(key, "szz_q")) {
            if (copy_str(p->szz_q, sizeof p->szz_q, val)) BAD("ERROR: invalid szz_q value on line %d\n", ln);
        } else if (!strcmp(key, "szz_file")) {
            if (copy_str(p->szz_file, sizeof p->szz_file, val)) BAD("ERROR: invalid szz_file\n");
        } else if (!strcmp(key, "sperp_q")) {
            if (copy_str(p->sperp_q, sizeof p->sperp_q, val)) BAD("ERROR: invalid sperp_q value on line %d\n", ln);
        } else if (!strcmp(key, "sperp_file")) {
            if (copy_str(p->sperp_file, sizeof p->sperp_file, val)) BAD("ERROR: invalid sperp_file\n");
        } else if (!strcmp(key, "seed")) {
            if (to_ull(val, &p->seed)) BAD("ERROR: seed must be unsigned integer (got %s)\n", val);
        } else if (!strcmp(key, "beta_list")) {
            const int r = beta_list(p, val);
            if (r == 2) BAD("ERROR: beta_list has too many entries (max 64)\n");
            else if (r) BAD("ERROR: beta_list must contain positive numeric values\n");
        } else if (!strcmp(key, "replica_log")) {
            if (copy_str(p->replica_log, sizeof p->replica_log, val)) BAD("ERROR: invalid replica_log\n");
        } else if (!strcmp(key, "spin_consistency_file")) {
            if (copy_str(p->spin_consistency_file, sizeof p->spin_consistency_file, val))
                BAD("ERROR: invalid spin_consistency_file\n");
        } else if (!strcmp(key, "sweep_order")) {
            if (copy_str(p->sweep_order, sizeof p->sweep_order, val)) BAD("ERROR: invalid sweep_order\n");
        } else if (!strcmp(key, "green_rebuild")) {
            if (copy_str(p->green_rebuild, sizeof p->green_rebuild, val)) BAD("ERROR: invalid green_rebuild\n");
        } else if (!strcmp(key, "precision")) {
            if (copy_str(p->precision, sizeof p->precision, val)) BAD("ERROR: invalid precision\n");
        } else if (!strcmp(key, "delay")) {
            if (to_int(val, &p->delay)) BAD("ERROR: delay must be integer (got %s)\n", val);
        } else if (!strcmp(key, "device")) {
            if (to_int(val, &p->device)) BAD("ERROR: device must be integer (got %s)\n", val);
        } else if (!strcmp(key, "drift_guard")) {
            if (to_double(val, &p->drift_guard)) BAD("ERROR: drift_guard must be numeric (got %s)\n", val);
        } else if (!strcmp(key, "autotune")) {
            if (to_int(val, &p->autotune)) BAD("ERROR: autotune must be integer (got %s)\n", val);
        } else {
            BAD("ERROR: unknown key %s\n", key);
        }
#undef BAD
    }
    fclose(fp);
    if (rc) return 1;
    if (p->nbeta == 0) {
        p->beta_list[0] = 2.0;
        p->nbeta = 1;
    }
    const char *err = NULL;
    if (strcmp(p->lattice, "chain") && strcmp(p->lattice, "square"))
        err = "lattice must be chain or square";
    else if (p->Lx <= 0) err = "Lx must be > 0";
    else if (p->Ly <= 0) err = "Ly must be > 0";
    else if (p->pbc != 0 && p->pbc != 1) err = "pbc must be 0 or 1";
    else if (p->U < 0.0) err = "U must be >= 0";
    else if (!(p->dtau > 0.0)) err = "dtau must be > 0";
    else if (p->nwarm < 0) err = "nwarm must be >= 0";
    else if (p->stab <= 0) err = "stab must be > 0";
    else if (strcmp(p->sweep_order, "forward") && strcmp(p->sweep_order, "alternating")) {
        fprintf(stderr, "ERROR: sweep_order must be forward or alternating (got %s)\n", p->sweep_order);
        return 1;
    } else if (strcmp(p->green_rebuild, "combine") && strcmp(p->green_rebuild, "two_sided") &&
               strcmp(p->green_rebuild, "centered")) {
        fprintf(stderr, "ERROR: green_rebuild must be combine, two_sided, or centered (got %s)\n",
                p->green_rebuild);
        return 1;
    } else if (!strcmp(p->green_rebuild, "centered") && !strcmp(p->precision, "mixed_fp32_chain"))
        err = "green_rebuild=centered needs precision=fp64 or mixed";
    else if (strcmp(p->parallel, "serial") && strcmp(p->parallel, "omp") && strcmp(p->parallel, "mpi") &&
             strcmp(p->parallel, "hybrid"))
        err = "parallel must be serial, omp, mpi, or hybrid";
    else if (p->nrep < 1) err = "nrep must be >= 1";
    else if (p->nbin < 2) err = "nbin must be >= 2";
    else if (p->nmeas < p->nbin) err = "nmeas must be >= nbin";
    else if (p->nmeas % p->nbin != 0) err = "nmeas must be divisible by nbin";
    else if (strcmp(p->precision, "fp64") && strcmp(p->precision, "mixed") && strcmp(p->precision, "mixed_fp32_chain"))
        err = "precision must be fp64, mixed or mixed_fp32_chain";
    else if (p->delay < 0) err = "delay must be >= 0";
    else if (!(p->drift_guard >= 0.0) && p->drift_guard != -1.0) err = "drift_guard must be >= 0";
    else if (p->autotune != 0 && p->autotune != 1) err = "autotune must be 0 or 1";
    if (err) {
        fprintf(stderr, "ERROR: %s\n", err);
        return 1;
    }
    return 0;
}
