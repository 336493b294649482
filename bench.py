"""Replica-sweeps/s of the DQMC Green's-function engine on one MI355X.

Workload (the metric's configuration): 16x16 periodic square lattice, t=-1,
U=4, mu=U/2, dtau=0.1, beta=8 (80 slices), stabilisation every 8 slices,
delayed-update depth 32, 128 independent replicas with seeds 1..128 whose
Ising fields come from field_init's RNG stream, tau=0 measurement with
Szz/Sperp at all 256 momenta after every timed sweep.

A step is one forward sweep of all 128 replicas plus its measurement.  The
mixed-precision engine (FP32 wrap / delayed updates / B-chains on MFMA, FP64
QR, rebuild and measurement) gives `value`; the FP64 engine is timed on the
same workload for the mixed/FP64 ratio.  Inputs are resident on the GPU
before the timed region, which is bracketed by HIP events.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sai-qmc_amd"))

METRIC = "replica-sweeps/s, 16x16 U=4 β=8: mixed-FP32 vs FP64 GPU at G drift ≤1e-6"
LX, LY, U, DTAU, BETA, STAB, DELAY, NREP = 16, 16, 4.0, 0.1, 8.0, 8, 32, 128
P32, P64 = 157.3e12, 78.6e12  # MI355X dense FP32 / FP64 matrix peaks


def stage_flops(n, L, s, acc):
    """Algorithmic flops per replica of one invocation of each stage class,
    split into (FP32-class in mixed, FP64-class) work.  Counts: GEMM 2n^3,
    Householder QR + explicit Q 8/3 n^3, unit-triangular inverse n^3/3."""
    n3 = float(n) ** 3
    return {
        "metro": (2.0 * acc * n3, 0.0),                 # rank-k flushes of a*n accepted flips
        "wrap": (4.0 * n3, 0.0),                        # B G B^-1
        "stack": (2.0 * n3 * (s - 1), 20.0 / 3 * n3),  # B-chain; T B, QR, U Q
        "lmul": (2.0 * n3 * (s - 1), 20.0 / 3 * n3),   # B-chain; B U, QR, R T
        "combine": (0.0, 26.0 / 3 * n3),               # Tl Ur, QR, Ul Q, R Tr
        "inv_one_plus": (0.0, 29.0 / 3 * n3),          # trtri, U^T T^-1, QR(M), trtri, 3 GEMMs
    }


def run_engine(dq, precision, steps, warmup, drift):
    eng = dq.Engine(precision, drift=drift)
    L = int(round(BETA / DTAU))
    mid = eng.add_model("square", LX, LY, U, DTAU, L, STAB, delay_k=DELAY)
    eng.add_replicas(mid, list(range(1, NREP + 1)))
    eng.set_measure(mid, nbin=1, q="all")
    eng.init()
    eng.sweep(warmup)
    eng.sweep(steps, 0)
    ms = eng.elapsed_ms()
    st = eng.status()
    if st["status"].any():
        raise RuntimeError(f"replica failure: reasons {sorted(set(st['reason'].tolist()))}")
    acc = float(st["accepted"].sum()) / float(st["attempts"].sum())
    out = dict(ms=ms, value=NREP * steps / (ms / 1e3), acc=acc)
    if drift:
        mx, mn, cnt = eng.drift()
        out["drift_max"] = float(mx.max())
        out["drift_mean"] = float(mn.mean())
    out["profile"] = eng.profile(1)
    eng.close()
    return out


def roofline(prof, acc, mixed):
    n, L, s = LX * LY, int(round(BETA / DTAU)), STAB
    fl = stage_flops(n, L, s, acc)
    best = None
    for k, (lo, hi) in fl.items():
        ms, cnt = prof[k]
        if cnt == 0:
            continue
        if best is None or ms > best[1]:
            best = (k, ms, cnt, lo, hi)
    k, ms, cnt, lo, hi = best
    t = ms / cnt / 1e3
    flops = (lo + hi) * NREP
    p_lo = P32 if mixed else P64
    t_roof = lo * NREP / p_lo + hi * NREP / P64
    peak = flops / t_roof
    ach = flops / t
    return {"bound": "mfma", "kernel": k, "achieved": round(ach / 1e12, 3), "peak": round(peak / 1e12, 3),
            "unit": "TFLOP/s", "frac": round(ach / peak, 4), "traffic": None,
            "avg_launch_ms": round(t * 1e3, 4), "launches_per_sweep": cnt}


def cpu_baseline():
    """The reference's own OpenMP replica driver (oracle/_ref/dqmc_omp) on
    the host cores, two runs differing only in nmeas so setup cancels."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dqmc_omp")
    cores = max(1, min(16, os.cpu_count() or 1))
    if not os.path.exists(exe):
        return None
    nrep = cores
    walls = []
    for nmeas in (2, 10):
        with tempfile.TemporaryDirectory() as d:
            inp = os.path.join(d, "in.txt")
            with open(inp, "w") as f:
                f.write(f"lattice=square\nLx={LX}\nLy={LY}\nU={U}\ndtau={DTAU}\nbeta_list={BETA}\nstab={STAB}\n"
                        f"nwarm=0\nnmeas={nmeas}\nnbin=2\nseed=1\nparallel=omp\nnrep={nrep}\n"
                        "szz_q=all\nsperp_q=all\n")
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OPENBLAS_NUM_THREADS="1")
            t0 = time.perf_counter()
            subprocess.run([exe, inp], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            walls.append(time.perf_counter() - t0)
    rate = nrep * (10 - 2) / (walls[1] - walls[0])
    return {"value": round(rate, 3), "unit": "replica-sweeps/s", "cores": cores, "kind": "reference",
            "sample": f"reference dqmc_omp, {nrep} replicas x 8 measured sweeps of the same 16x16 U=4 beta=8 "
                      f"stab=8 workload (difference of nmeas=10 and nmeas=2 runs, nwarm=0), "
                      f"OMP_NUM_THREADS={cores}, sequential OpenBLAS"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: replicas are independent Markov chains and this engine runs on a "
              "single MI355X; more than one GPU is not supported", file=sys.stderr)
        sys.exit(2)
    import dqmc_gpu as dq

    runs = {}
    for name, prec in (("mixed", dq.MIXED), ("fp64", dq.FP64), ("mixed_fp32_chain", dq.MIXED_FP32_CHAIN)):
        runs[name] = run_engine(dq, prec, a.steps, a.warmup, drift=True)
        r = runs[name]
        print(f"# {name}: {r['value']:.1f} replica-sweeps/s, {r['ms'] / a.steps:.2f} ms/sweep, "
              f"drift max {r['drift_max']:.3e} mean {r['drift_mean']:.3e}", file=sys.stderr, flush=True)
    mixed, fp64 = runs["mixed"], runs["fp64"]
    cpu = None if a.no_cpu else cpu_baseline()
    line = {
        "metric": METRIC,
        "value": round(mixed["value"], 2),
        "unit": "replica-sweeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(mixed["ms"] / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic: Ising fields drawn from seeds 1..128 by field_init's RNG stream",
        "config": {"workload": "16x16 periodic square, t=-1, U=4, dtau=0.1, beta=8 (80 slices), stab=8, "
                               "delay depth 32, 128 replicas, Szz/Sperp at all 256 q",
                   "replicas": NREP, "precision": "mixed FP32/FP64"},
        "fp64": {"value": round(fp64["value"], 2), "ms_per_step": round(fp64["ms"] / a.steps, 3),
                 "drift_max": fp64["drift_max"], "drift_mean": fp64["drift_mean"]},
        "mixed_over_fp64": round(mixed["value"] / fp64["value"], 3),
        "mixed_fp32_chain": {"value": round(runs["mixed_fp32_chain"]["value"], 2),
                             "drift_max": runs["mixed_fp32_chain"]["drift_max"],
                             "drift_mean": runs["mixed_fp32_chain"]["drift_mean"]},
        "drift_max": mixed["drift_max"],
        "drift_mean": mixed["drift_mean"],
        "acceptance": round(mixed["acc"], 4),
        "roofline": roofline(mixed["profile"], mixed["acc"], True),
        "roofline_fp64": roofline(fp64["profile"], fp64["acc"], False),
        "profile_ms_per_sweep": {k: round(v[0], 3) for k, v in mixed["profile"].items()},
        "profile_fp64_ms_per_sweep": {k: round(v[0], 3) for k, v in fp64["profile"].items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
