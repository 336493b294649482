"""Replica-sweeps/s of the DQMC Green's-function engine on one MI355X.

Workload (the metric's configuration): 16x16 periodic square lattice, t=-1,
U=4, mu=U/2, dtau=0.1, beta=8 (80 slices), stabilisation every 8 slices,
128 independent replicas with seeds 1..128 whose Ising fields come from
field_init's RNG stream, tau=0 measurement with Szz/Sperp at all 256
momenta after every timed sweep.

A step is one forward sweep of all 128 replicas plus its measurement.  The
FP64 engine gives `value`; inputs are resident on the GPU before the timed
region, which is bracketed by HIP events on the engine's stream.  Beside it:
the FP64 drift as the reference defines it (against a from-scratch G before
each rebuild) next to the reference's own value on this workload; the
alternating sweep order; the two mixed FP32/FP64 engines, which need a drift
guard and fail every replica at the metric's bound of 1e-6 (their speed is
reported with the guard out of reach); the compiled reference on the host
cores as the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--fp64-only]
                    [--calibrate-copy] [--groups G] [--delay K] [--no-small] [--only-small]

Beside the metric's line: the 8x8 workload (beta=4, 64 replicas) on the
persistent sweep kernel and on the batched path, and the uneven batch of 20
models (256 replicas) in one context.
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


def run_engine(dq, precision, steps, warmup, drift=False, guard=0.0, groups=1, delay=DELAY, order="forward",
               profile=True, allow_fail=False, autotune=True):
    """The 128 replicas as `groups` equal models: each model has its own
    stream and sweep graph, so the groups' sweeps run concurrently.
    order = "alternating" runs forward and backward sweeps in turn."""
    eng = dq.Engine(precision, drift=drift, drift_guard=guard, autotune=autotune)
    L = int(round(BETA / DTAU))
    per = NREP // groups
    for g in range(groups):
        mid = eng.add_model("square", LX, LY, U, DTAU, L, STAB, delay_k=delay)
        eng.set_sweep_order(mid, order)
        eng.add_replicas(mid, list(range(1 + g * per, 1 + (g + 1) * per)))
        eng.set_measure(mid, nbin=1, q="all")
    eng.init()
    tuning = eng.tuning(0)
    eng.sweep(warmup)
    eng.sweep(steps, 0)
    ms = eng.elapsed_ms()
    st = eng.status()
    nfail = int((st["status"] != 0).sum())
    if nfail and not allow_fail:
        raise RuntimeError(f"replica failure: reasons {sorted(set(st['reason'].tolist()))}")
    acc = float(st["accepted"].sum()) / max(1.0, float(st["attempts"].sum()))
    out = dict(ms=ms, value=NREP * steps / (ms / 1e3), acc=acc, failed=nfail, tuning=tuning,
               reasons=sorted(set(int(x) for x in st["reason"][st["status"] != 0])))
    if drift:
        mx, mn, cnt = eng.drift()
        ok = st["status"] == 0
        out["drift_max"] = float(mx[ok].max()) if ok.any() else None
        out["drift_mean"] = float(mn[ok].mean()) if ok.any() else None
    if profile:
        out["profile"] = eng.profile(1)
    eng.close()
    return out


def reference_drift():
    """The compiled reference's own stab_drift_file on this workload
    (tests/golden/drift_16x16_beta8.txt: 4 serial replicas, 6 sweeps)."""
    path = os.path.join(ROOT, "tests", "golden", "drift_16x16_beta8.txt")
    if not os.path.exists(path):
        return None
    rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
    mx = [float(r[9]) for r in rows]
    mn = [float(r[10]) for r in rows]
    return {"max": max(mx), "mean": sum(mn) / len(mn), "replicas": len(rows),
            "source": "reference dqmc_omp stab_drift_file, 4 replicas x (2 warm-up + 4 measured) sweeps"}


def traffic_from_profile(kernel, segments=None):
    """HBM bytes per launch of `kernel` (its name's prefix) from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, each scaled by what the
    counter reads for a device copy of known size); None when the summary
    was taken with another number of Metropolis site segments than this
    run's."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fp64_16x16_beta8_128rep_pmc.txt")
    if not os.path.exists(path):
        return None
    val, seg = None, None
    for line in open(path):
        if line.startswith(kernel) and val is None:
            val = float(line.split()[-1]) * 1e6
        if line.startswith("metro_segments"):
            seg = int(line.split()[-1])
    if segments is not None and seg != segments:
        return None
    return val


def roofline_metro_kernel(prof, acc, segments, window, nrep_launch=NREP):
    """Roofline of the dominant kernel by total time, the Metropolis kernel
    (one launch per site segment of n/S sites, S segments per slice).
    Algorithmic bytes per replica and launch: the window columns and rows of
    G it corrects, 2 n^2 / S values, plus the pending pairs it writes,
    a n / S pairs of 2n values (a = acceptance), 8 bytes each.  Launch
    duration from HIP events around every proposal launch of a profiled
    sweep."""
    n = LX * LY
    pms, fcnt = prof["propose"]
    t = pms / fcnt / 1e3
    byts = (2 * n * n / segments + acc * (n / segments) * 2 * n) * 8 * nrep_launch
    return {"bound": "hbm", "kernel": f"k_metro_slice<double, {window}>",
            "achieved": round(byts / t / 1e9, 1), "peak": round(HBM / 1e9, 1), "unit": "GB/s",
            "frac": round(byts / t / HBM, 4),
            "traffic": traffic_from_profile(f"void k_metro_slice<double, {window}>", segments),
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
            "traffic": None if mixed else traffic_from_profile("void k_gemm_dl<0>"),
            "avg_launch_ms": round(t * 1e3, 4), "launches_per_sweep": cnt, "flops_per_launch": flops}


def roofline_metro(prof, acc, nrep_launch=NREP):
    """The Metropolis stage of one slice (proposal kernels + rank-P flushes)
    against the delayed-update work the reference does for it: 2 a n^3
    flops per replica (a = acceptance)."""
    n = LX * LY
    ms, cnt = prof["metro"]
    t = ms / cnt / 1e3
    flops = 2.0 * acc * n ** 3 * nrep_launch
    flush = "k_flush64" if n % 64 == 0 else "k_gemm_dl<1> (per-replica K)"
    return {"bound": "mfma", "stage": f"metro (k_metro_slice + flush {flush})",
            "achieved": round(flops / t / 1e12, 3), "peak": round(P64 / 1e12, 3), "unit": "TFLOP/s",
            "frac": round(flops / t / P64, 4), "avg_ms_per_slice": round(t * 1e3, 4)}


def roofline_qr(prof, nrep_launch=NREP):
    """Every Householder QR of a sweep (k_qr_udv + the explicit Q of
    k_qr_formq_reg where formed): 4/3 n^3 + 4/3 n^3 flops per replica."""
    n = LX * LY
    ms, cnt = prof["qr"]
    t = ms / cnt / 1e3
    flops = (8.0 / 3.0) * n ** 3 * nrep_launch
    return {"bound": "mfma", "stage": "qr (k_qr_udv + k_qr_formq_reg)", "achieved": round(flops / t / 1e12, 3),
            "peak": round(P64 / 1e12, 3), "unit": "TFLOP/s", "frac": round(flops / t / P64, 4),
            "avg_ms_per_qr": round(t * 1e3, 4), "qr_per_sweep": cnt}


def measure_gbs(prof, nrep_launch=NREP):
    """The tau = 0 measurement (k_measure): bytes per sample 2 n^2 * 8 (G_up,
    G_dn) + n nq * 8 (phase table) + 2 nq * 8 (Szz, Sperp) per replica."""
    n = nq = LX * LY
    ms, cnt = prof["measure"]
    if cnt == 0:
        return None
    t = ms / cnt / 1e3
    byts = (2 * n * n * 8 + n * nq * 8 + 2 * nq * 8) * nrep_launch
    return {"bound": "hbm", "kernel": "k_measure", "achieved": round(byts / t / 1e9, 1),
            "peak": round(HBM / 1e9, 1), "unit": "GB/s", "frac": round(byts / t / HBM, 4),
            "avg_launch_ms": round(t * 1e3, 4), "bytes_per_launch": byts}


def work_per_replica_sweep(n, L, stab, acc):
    """The algorithmic flops of one forward replica-sweep (the reference's
    dense counts: wrap 4n^3 L, rank-k flushes 2a n^3 L, the B-block chains
    2n^3 (s-1)(2M-1); FP64 QR class (2M-1)(20/3)n^3 + (M-1)(26/3)n^3 +
    M (29/3)n^3)."""
    M = -(-L // stab)
    lo = 4 * n ** 3 * L + 2 * acc * n ** 3 * L + 2 * n ** 3 * (stab - 1) * (2 * M - 1)
    hi = (2 * M - 1) * (20 / 3) * n ** 3 + (M - 1) * (26 / 3) * n ** 3 + M * (29 / 3) * n ** 3
    return lo + hi


def small_lattice_lines(dq, steps, warmup, which="all"):
    """The 8x8 workload (U=4, dtau=0.1, beta=4: 40 slices, stab=10, 64
    replicas with seeds 1..64, Szz/Sperp at all 64 q after every sweep) on
    the persistent sweep kernel (one workgroup per replica, G in LDS) and on
    the batched path (DQMC_GPU_PERSIST=0); and the uneven batch of 20 models
    (4x4 .. 16x16 at beta 2, 4, 8, 12, stab=10, 256 replicas) in one
    context, the lattices of at most 64 sites in one persistent launch."""
    out = {}
    n, L, stab, R = 64, 40, 10, 64
    runs = {}
    if which == "c4":
        return dict(uneven_batch_line(dq, steps))
    for mode in (("1",) if which == "c2" else ("1", "0")):
        os.environ["DQMC_GPU_PERSIST"] = mode
        eng = dq.Engine(dq.FP64, autotune=True)
        mid = eng.add_model("square", 8, 8, U, DTAU, L, stab, delay_k=16)
        eng.add_replicas(mid, list(range(1, R + 1)))
        eng.set_measure(mid, nbin=1, q="all")
        eng.init()
        eng.sweep(warmup)
        eng.sweep(steps, 0)
        ms = eng.elapsed_ms()
        st = eng.status()
        acc = float(st["accepted"].sum()) / max(1.0, float(st["attempts"].sum()))
        prof = eng.profile(2) if mode == "1" else None
        eng.close()
        runs[mode] = (ms / steps, R * steps / (ms / 1e3), acc, prof)
    os.environ.pop("DQMC_GPU_PERSIST", None)
    ms_p, v_p, acc, prof = runs["1"]
    ms_b, v_b, _, _ = runs.get("0", (float("nan"), float("nan"), None, None))
    flops = work_per_replica_sweep(n, L, stab, acc)
    achieved = v_p * flops
    out["c2_8x8_beta4"] = {
        "workload": "8x8 periodic square, U=4, dtau=0.1, beta=4 (40 slices), stab=10, 64 replicas, Szz/Sperp at "
                    "all 64 q after every sweep",
        "persistent": {"value": round(v_p, 1), "ms_per_step": round(ms_p, 4),
                       "phases_ms_per_sweep": {k: round(prof[k][0] / 2, 4) for k in
                                               ("stack", "metro", "wrap", "lmul", "inv_one_plus", "measure")}},
        "batched": ({"value": round(v_b, 1), "ms_per_step": round(ms_b, 4)} if "0" in runs else None),
        "persistent_over_batched": round(v_p / v_b, 3) if "0" in runs else None,
        "unit": "replica-sweeps/s",
        "roofline": {"bound": "mfma", "flops_per_replica_sweep": round(flops), "achieved": round(achieved / 1e12, 3),
                     "peak_chip": round(P64 / 1e12, 1), "frac_chip": round(achieved / P64, 4),
                     "peak_of_cus_used": round(P64 * R / 256 / 1e12, 2),
                     "frac_of_cus_used": round(achieved / (P64 * R / 256), 4), "unit": "TFLOP/s",
                     "note": "one workgroup per replica: 64 of 256 CUs work; latency-bound (sequential proposals "
                             "and Householder column steps)"}}
    print(f"# c2 8x8: persistent {v_p:.1f}, batched {v_b:.1f} replica-sweeps/s", file=sys.stderr, flush=True)
    if which == "c2":
        return out
    out.update(uneven_batch_line(dq, steps))
    return out


def uneven_batch_line(dq, steps):
    """C4: 20 models (4x4 .. 16x16 at beta 2, 4, 8, 12), 256 replicas, one
    context."""
    out = {}
    eng = dq.Engine(dq.FP64, autotune=True)
    total, k = 256, 0
    for Lx in (4, 6, 8, 12, 16):
        for beta in (2.0, 4.0, 8.0, 12.0):
            Lb = int(round(beta / DTAU))
            nrep = total // 20 + (1 if k < total % 20 else 0)
            mid = eng.add_model("square", Lx, Lx, U, DTAU, Lb, 10, delay_k=0)
            eng.add_replicas(mid, [dq.replica_seed(1, k, r) for r in range(nrep)])
            eng.set_measure(mid, nbin=1, q="all")
            k += 1
    eng.init()
    eng.sweep(1)
    eng.sweep(max(2, steps // 4), 0)
    ms = eng.elapsed_ms() / max(2, steps // 4)
    st = eng.status()
    eng.close()
    if (st["status"] != 0).any():
        raise RuntimeError("C4 replica failure")
    out["c4_uneven_256"] = {
        "workload": "20 models: 4x4, 6x6, 8x8, 12x12, 16x16 square, each at beta 2, 4, 8, 12 (dtau=0.1), U=4, "
                    "stab=10, 256 replicas, Szz/Sperp at all q; 4x4 .. 8x8 in one persistent launch per sweep, "
                    "12x12 and 16x16 on the batched path",
        "value": round(256 / (ms / 1e3), 1), "ms_per_step": round(ms, 3), "unit": "replica-sweeps/s"}
    print(f"# c4: {256 / (ms / 1e3):.1f} replica-sweeps/s, {ms:.2f} ms per sweep of all 256", file=sys.stderr,
          flush=True)
    return out


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                return ln.split(":", 1)[1].strip()
    except (OSError, subprocess.CalledProcessError):
        pass
    return None


def cpu_baseline():
    """The reference's own OpenMP replica driver (oracle/_ref/dqmc_omp), 128
    replicas on every core this process may use, sequential OpenBLAS, two
    runs differing only in nmeas so set-up cancels."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dqmc_omp")
    if not os.path.exists(exe):
        return None
    cores = len(os.sched_getaffinity(0))
    walls = []
    lo, hi = 2, 6
    for nmeas in (lo, hi):
        with tempfile.TemporaryDirectory() as d:
            inp = os.path.join(d, "in.txt")
            with open(inp, "w") as f:
                f.write(f"lattice=square\nLx={LX}\nLy={LY}\nU={U}\ndtau={DTAU}\nbeta_list={BETA}\nstab={STAB}\n"
                        f"nwarm=0\nnmeas={nmeas}\nnbin=2\nseed=1\nparallel=omp\nnrep={NREP}\n"
                        "szz_q=all\nsperp_q=all\n")
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OPENBLAS_NUM_THREADS="1")
            print(f"# cpu baseline: {NREP} replicas, nmeas={nmeas}, {cores} threads", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            subprocess.run([exe, inp], cwd=d, env=env, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            walls.append(time.perf_counter() - t0)
    rate = NREP * (hi - lo) / (walls[1] - walls[0])
    return {"value": round(rate, 3), "unit": "replica-sweeps/s", "cores": cores, "kind": "reference",
            "cpu": cpu_model(),
            "sample": f"reference dqmc_omp, {NREP} replicas x {hi - lo} measured sweeps of the same 16x16 U=4 "
                      f"beta=8 stab=8 workload with Szz/Sperp at all q (difference of nmeas={hi} and nmeas={lo} "
                      f"runs, nwarm=0), OMP_NUM_THREADS={cores} (every core this process may use), "
                      "sequential OpenBLAS"}


def calibrate_copy():
    """A plain device copy of known size (256 MiB read, 256 MiB written,
    hipMemcpy device to device) for scaling the PMC byte counters."""
    import ctypes as C
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    nbytes = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(a), nbytes) or h.hipMalloc(C.byref(b), nbytes) or h.hipMemset(a, 1, nbytes):
        raise RuntimeError("calibration buffers")
    for _ in range(3):
        if h.hipMemcpy(b, a, nbytes, 3):
            raise RuntimeError("calibration copy")
    h.hipDeviceSynchronize()
    h.hipFree(a)
    h.hipFree(b)
    print(f"# calibration copy: {nbytes} bytes read and written per dispatch", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--groups", type=int, default=1, help="replica groups swept concurrently")
    ap.add_argument("--delay", type=int, default=DELAY, help="delayed-update window (sites)")
    ap.add_argument("--fp64-only", action="store_true", help="skip the mixed, drift and alternating runs")
    ap.add_argument("--no-small", action="store_true", help="skip the 8x8 and uneven-batch lines")
    ap.add_argument("--only-small", choices=("all", "c2", "c4"), default=None,
                    help="print only the 8x8 (c2) and/or uneven-batch (c4) lines (for profiling the persistent kernel)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the engine's default launch shape (DQMC_GPU_METRO_SPLIT sets the segments)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: replicas are independent Markov chains and this engine runs on a "
              "single MI355X; more than one GPU is not supported", file=sys.stderr)
        sys.exit(2)
    import dqmc_gpu as dq

    if a.only_small:
        print(json.dumps(small_lattice_lines(dq, a.steps, a.warmup, a.only_small)))
        return
    if a.calibrate_copy:
        calibrate_copy()
    main_run = run_engine(dq, dq.FP64, a.steps, a.warmup, groups=a.groups, delay=a.delay,
                          autotune=not a.no_autotune)
    print(f"# fp64: {main_run['value']:.1f} replica-sweeps/s, {main_run['ms'] / a.steps:.2f} ms/sweep",
          file=sys.stderr, flush=True)
    extra = {}
    if not a.fp64_only:
        # drift as the reference defines it (against a from-scratch G before
        # every rebuild), on a short separate run: the from-scratch G costs
        # L/stab more UDV updates per rebuild and is not part of the timed run
        d = run_engine(dq, dq.FP64, 2, 1, drift="scratch", profile=False)
        extra["drift"] = {"definition": "max|G_wrap - G_scratch(tau)| before each rebuild, inf-norm "
                                        "(dqmc_record_stab_drift)", "sweeps": 3, "max": d["drift_max"],
                          "mean": d["drift_mean"], "reference": reference_drift()}
        print(f"# fp64 drift (from scratch): max {d['drift_max']:.3e} mean {d['drift_mean']:.3e}", file=sys.stderr,
              flush=True)
        r = run_engine(dq, dq.FP64, a.steps, a.warmup, groups=a.groups, delay=a.delay, order="alternating",
                       drift="rebuild")
        extra["fp64_alternating"] = {"value": round(r["value"], 2), "ms_per_step": round(r["ms"] / a.steps, 3),
                                     "over_fp64": round(r["value"] / main_run["value"], 3),
                                     "drift_max_vs_rebuilt": r["drift_max"], "drift_mean_vs_rebuilt": r["drift_mean"]}
        print(f"# fp64_alternating: {r['value']:.1f} replica-sweeps/s", file=sys.stderr, flush=True)
        for name, prec in (("mixed", dq.MIXED), ("mixed_fp32_chain", dq.MIXED_FP32_CHAIN)):
            # with the metric's bound as the guard every replica fails at its
            # first rebuild; the speed is taken with the guard out of reach
            g = run_engine(dq, prec, 1, 0, guard=1e-6, profile=False, allow_fail=True)
            r = run_engine(dq, prec, a.steps, a.warmup, guard=1e300, drift="rebuild", groups=a.groups,
                           delay=a.delay, profile=False)
            extra[name] = {"replicas_failed_at_guard_1e-6": g["failed"], "fail_reasons": g["reasons"],
                           "value_unguarded": round(r["value"], 2), "over_fp64": round(r["value"] / main_run["value"], 3),
                           "drift_max_vs_rebuilt": r["drift_max"], "drift_mean_vs_rebuilt": r["drift_mean"]}
            print(f"# {name}: {g['failed']}/{NREP} replicas fail at drift_guard=1e-6; unguarded "
                  f"{r['value']:.1f} replica-sweeps/s, drift max {r['drift_max']:.3e}", file=sys.stderr, flush=True)
    if not a.no_small:
        extra.update(small_lattice_lines(dq, a.steps, a.warmup))
    cpu = None if a.no_cpu else cpu_baseline()
    n = LX * LY
    acc = main_run["acc"]
    tw, ts, tsep = main_run["tuning"]
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
        "dtype": "f64",
        "data": "synthetic: Ising fields drawn from seeds 1..128 by field_init's RNG stream",
        "config": {"workload": "16x16 periodic square, t=-1, U=4, mu=U/2, dtau=0.1, beta=8 (80 slices), stab=8, "
                               "128 replicas, Szz/Sperp at all 256 q after every sweep; delayed updates in "
                               f"{tw}-site windows, the pending pairs of each of the {ts} site segments of a slice "
                               f"(about a*n/{ts} = {round(acc * n / ts)} at acceptance a) applied as one rank-P "
                               "update (window and segments chosen by the autotuner), column pre-pivoted UDV "
                               f"chain, {'separable' if tsep else 'dense'} B-matrix products",
                   "replicas": NREP, "precision": "fp64", "metro_window": tw, "metro_segments": ts},
        "acceptance": round(acc, 4),
        "roofline": roofline_metro_kernel(main_run["profile"], acc, ts, tw, NREP // a.groups),
        "roofline_qr": roofline_qr(main_run["profile"], NREP // a.groups),
        "roofline_gemm": roofline_gemm(main_run["profile"], False, NREP // a.groups),
        "roofline_metro": roofline_metro(main_run["profile"], acc, NREP // a.groups),
        "measurement": measure_gbs(main_run["profile"], NREP // a.groups),
        "profile_ms_per_sweep": {k: round(v[0], 3) for k, v in main_run["profile"].items()},
    }
    line.update(extra)
    line["cpu_baseline"] = cpu
    if cpu:
        line["speedup_vs_cpu_baseline"] = round(main_run["value"] / cpu["value"], 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
