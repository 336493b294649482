"""Replica-sweeps/s of the DQMC Green's-function engine on one MI355X.

Workload (the metric's configuration): 16x16 periodic square lattice, t=-1,
U=4, mu=U/2, dtau=0.1, beta=8 (80 slices), stabilisation every 8 slices,
delayed-update depth 32, 128 independent replicas with seeds 1..128 whose
Ising fields come from field_init's RNG stream, tau=0 measurement with
Szz/Sperp at all 256 momenta after every timed sweep.

A step is one forward sweep of all 128 replicas plus its measurement.  The
FP64 engine gives `value`: it is the engine whose wrap-vs-rebuild drift
stays at the reference's own level on this workload.  The two mixed
FP32/FP64 engines (FP32 wrap / delayed updates / B-chains on MFMA, FP64 QR,
rebuild and measurement) are timed on the same workload and reported beside
it with their drift.  Inputs are resident on the GPU before the timed
region, which is bracketed by HIP events.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--precision all|fp64|mixed|mixed_fp32_chain] [--calibrate-copy]
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
HBM = 8.0e12  # MI355X HBM3E bytes/s


def run_engine(dq, precision, steps, warmup, drift, groups=1, delay=DELAY, order="forward"):
    """The 128 replicas as `groups` equal models: each model has its own
    stream and sweep graph, so the groups' sweeps run concurrently.
    order = "alternating" runs forward and backward sweeps in turn."""
    eng = dq.Engine(precision, drift=drift)
    L = int(round(BETA / DTAU))
    per = NREP // groups
    for g in range(groups):
        mid = eng.add_model("square", LX, LY, U, DTAU, L, STAB, delay_k=delay)
        eng.set_sweep_order(mid, order)
        eng.add_replicas(mid, list(range(1 + g * per, 1 + (g + 1) * per)))
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


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads
    for a device copy of known size)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_16x16_beta8_128rep_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch"):
            return float(line.split()[-1])
    return None


def roofline_metro_kernel(prof, acc, mixed, nrep_launch=NREP):
    """Roofline of the dominant kernel by total time, the Metropolis kernel
    (one launch per half-slice site segment).  Algorithmic bytes per replica
    and launch: the window columns and rows of G it corrects, n^2 values for
    the n/2 sites of a segment, plus the pending pairs it writes, a n/2
    pairs of 2n values (a = acceptance).  Launch duration from HIP events of
    a profiled sweep: the Metropolis stage minus its flush GEMMs, over the
    number of flushes (one per launch)."""
    n = LX * LY
    ms, _ = prof["metro"]
    fms, fcnt = prof["flush"]
    t = (ms - fms) / fcnt / 1e3
    w = 4 if mixed else 8
    byts = (n * n + acc * (n // 2) * 2 * n) * w * nrep_launch
    return {"bound": "hbm", "kernel": "k_metro_slice<float, 32>" if mixed else "k_metro_slice<double, 16>",
            "achieved": round(byts / t / 1e9, 1), "peak": round(HBM / 1e9, 1), "unit": "GB/s",
            "frac": round(byts / t / HBM, 4),
            "traffic": None if mixed else traffic_from_profile("k_metro_slice"),
            "avg_launch_ms": round(t * 1e3, 4), "launches_per_sweep": fcnt, "bytes_per_launch": round(byts)}


def roofline_gemm(prof, mixed, nrep_launch=NREP):
    """The batched n x n x n FP64 GEMM (k_gemm_dl; the few transposed ones
    run on k_gemm_big), 2 n^3 flops per replica per launch, average launch
    duration from HIP events around every such launch of a profiled sweep."""
    n = LX * LY
    ms, cnt = prof["gemm"]
    t = ms / cnt / 1e3
    flops = 2.0 * n ** 3 * nrep_launch
    peak = P64
    return {"bound": "mfma", "kernel": "k_gemm_dl", "achieved": round(flops / t / 1e12, 3),
            "peak": round(peak / 1e12, 3), "unit": "TFLOP/s", "frac": round(flops / t / peak, 4),
            "traffic": None if mixed else traffic_from_profile("k_gemm_dl"),
            "avg_launch_ms": round(t * 1e3, 4), "launches_per_sweep": cnt, "flops_per_launch": flops}


def roofline_metro(prof, acc, nrep_launch=NREP):
    """The Metropolis stage of one slice (proposal kernels + flush GEMMs)
    against the delayed-update work the reference does for it: 2 a n^3
    flops per replica (a = acceptance)."""
    n = LX * LY
    ms, cnt = prof["metro"]
    t = ms / cnt / 1e3
    flops = 2.0 * acc * n ** 3 * nrep_launch
    return {"bound": "mfma", "stage": "metro (k_metro_slice + flush k_gemm_big<double,0,1>)",
            "achieved": round(flops / t / 1e12, 3), "peak": round(P64 / 1e12, 3), "unit": "TFLOP/s",
            "frac": round(flops / t / P64, 4), "avg_ms_per_slice": round(t * 1e3, 4)}


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


def calibrate_copy():
    """A plain device copy of known size (256 MiB read, 256 MiB written) for
    scaling the PMC byte counters."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"# calibration copy: {a.numel() * 4} bytes read and written per dispatch", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="all", choices=["all", "fp64", "mixed", "mixed_fp32_chain"])
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--groups", type=int, default=1, help="replica groups swept concurrently")
    ap.add_argument("--delay", type=int, default=DELAY, help="delayed-update window (sites)")
    ap.add_argument("--no-alternating", action="store_true",
                    help="skip the FP64 sweep_order=alternating run reported beside the main line")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: replicas are independent Markov chains and this engine runs on a "
              "single MI355X; more than one GPU is not supported", file=sys.stderr)
        sys.exit(2)
    import dqmc_gpu as dq

    if a.calibrate_copy:
        calibrate_copy()
    modes = {"fp64": dq.FP64, "mixed": dq.MIXED, "mixed_fp32_chain": dq.MIXED_FP32_CHAIN}
    names = list(modes) if a.precision == "all" else [a.precision]
    runs = {}
    for name in names:
        runs[name] = run_engine(dq, modes[name], a.steps, a.warmup, drift=True, groups=a.groups, delay=a.delay)
        r = runs[name]
        print(f"# {name}: {r['value']:.1f} replica-sweeps/s, {r['ms'] / a.steps:.2f} ms/sweep, "
              f"drift max {r['drift_max']:.3e} mean {r['drift_mean']:.3e}", file=sys.stderr, flush=True)
    if not a.no_alternating and "fp64" in names:
        r = run_engine(dq, dq.FP64, a.steps, a.warmup, drift=True, groups=a.groups, delay=a.delay,
                       order="alternating")
        runs["fp64_alternating"] = r
        names.append("fp64_alternating")
        print(f"# fp64_alternating: {r['value']:.1f} replica-sweeps/s, {r['ms'] / a.steps:.2f} ms/sweep, "
              f"drift max {r['drift_max']:.3e} mean {r['drift_mean']:.3e}", file=sys.stderr, flush=True)
    main_name = names[0]
    main_run = runs[main_name]
    cpu = None if a.no_cpu else cpu_baseline()
    line = {
        "metric": METRIC,
        "value": round(main_run["value"], 2),
        "unit": "replica-sweeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(main_run["ms"] / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_name == "fp64" else "f32+f64",
        "data": "synthetic: Ising fields drawn from seeds 1..128 by field_init's RNG stream",
        "config": {"workload": "16x16 periodic square, t=-1, U=4, mu=U/2, dtau=0.1, beta=8 (80 slices), stab=8, "
                               "delay depth 32, 128 replicas, Szz/Sperp at all 256 q",
                   "replicas": NREP, "precision": main_name},
        "drift_max": main_run["drift_max"],
        "drift_mean": main_run["drift_mean"],
        "acceptance": round(main_run["acc"], 4),
        "roofline": roofline_metro_kernel(main_run["profile"], main_run["acc"], main_name != "fp64",
                                          NREP // a.groups),
        "roofline_gemm": roofline_gemm(main_run["profile"], main_name != "fp64", NREP // a.groups),
        "roofline_metro": roofline_metro(main_run["profile"], main_run["acc"], NREP // a.groups),
        "profile_ms_per_sweep": {k: round(v[0], 3) for k, v in main_run["profile"].items()},
    }
    for name in names[1:]:
        r = runs[name]
        line[name] = {"value": round(r["value"], 2), "ms_per_step": round(r["ms"] / a.steps, 3),
                      "over_fp64": round(r["value"] / main_run["value"], 3),
                      "drift_max": r["drift_max"], "drift_mean": r["drift_mean"],
                      "roofline_metro": roofline_metro(r["profile"], r["acc"], NREP // a.groups),
                      "profile_ms_per_sweep": {k: round(v[0], 3) for k, v in r["profile"].items()}}
    line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
