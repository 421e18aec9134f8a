"""Benchmark: Goldilocks forward NTT of 2^24 words on one MI355X.

A step is one forward NTT over 2^24 HBM-resident words (synthetic, seeded
uniform in [0, p)).  Prints one JSON line with the whole-job throughput
(Gelem/s), the roofline of the dominant kernel and a CPU baseline (the
oracle's port of the reference algorithm on this host's cores, bounded
sample).  Extra workloads: --workload batch (mixed-size batched NTTs,
batched vs one-launch-sequence-per-transform, with p50/p99 latency),
--workload commit (coset LDE + K12 Merkle commit), --workload serve (open-loop
Poisson NTT/commit jobs through the continuous-batching server), --workload
permute (the reference's transpose and bit_reverse on HBM-resident words),
--workload hostbatch (NTT batches on pageable host arrays, PCIe included).

The default line also carries the metric's second half, "mixed_batch": the
1024-transform mixed-size batch (n = 2^k * f, k in [10, 20], f in
{1, 3, 5, 15, 17, 257}) timed grouped, one launch sequence per transform and
as a captured hipGraph, with p50/p99 batch latency.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--n", type=int, default=0,
                    help="ntt: transform length, any divisor of p - 1 (overrides --log-n)")
    ap.add_argument("--workload", default="ntt", choices=["ntt", "batch", "commit", "serve", "permute", "hostbatch"])
    ap.add_argument("--qps", type=float, default=0.0, help="serve: arrival rate (0: 60%% of capacity)")
    ap.add_argument("--slots", type=int, default=8, help="serve: server slots (streams), 1-16")
    ap.add_argument("--no-host-buffer", action="store_true",
                    help="ntt: skip the host-array (PCIe) timing, e.g. under a profiler")
    ap.add_argument("--batch-k", type=int, nargs=2, default=[10, 20], metavar=("LO", "HI"),
                    help="batch: transform lengths 2^k * f with k uniform in [LO, HI]")
    ap.add_argument("--no-batch", action="store_true",
                    help="ntt: leave out the mixed-size batch sub-object")
    ap.add_argument("--no-other-sizes", action="store_true",
                    help="ntt: leave out the other lengths (2^16, 2^20, 2^27), e.g. under a profiler")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound on the CPU baseline sample (0 disables it)")
    return ap.parse_args()


def traffic_for(name, workload):
    """HBM bytes per launch of kernel `name` in `workload`, measured with
    rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE runs, each scaled by the same
    counter's reading for a plain copy of known size); recorded in
    profiles/traffic.json by goldilocks_amd.profiling.  None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload, {}).get(name)
    except (OSError, ValueError):
        return None
    return rec["bytes_per_launch"] if rec else None


def kernel_table(timings):
    """[(name, ms)] -> {name: {"launches", "total_ms", "avg_us"}}"""
    tab = {}
    for name, ms in timings:
        t = tab.setdefault(name, {"launches": 0, "total_ms": 0.0})
        t["launches"] += 1
        t["total_ms"] += ms
    for t in tab.values():
        t["avg_us"] = round(1e3 * t["total_ms"] / t["launches"], 2)
        t["total_ms"] = round(t["total_ms"], 4)
    return tab


def host_cpu():
    """(model name, logical CPUs this process may run on, threads used)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    # the host's CPU share for this GPU, where the environment states one
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(usable, share) if share > 0 else usable
    return model, usable, max(1, threads)


def cpu_baseline_ntt(n, seconds):
    """The oracle (C port of the reference's plan tree) on this host: the
    top Cooley-Tukey level split over every usable host core (as the
    reference's ntt-par splits it over rayon), then on one thread; whole 2^k
    transforms timed with the reference's rule (bench/src/main.rs:61-95:
    repeat until the accumulated time passes a budget, report the median),
    each budget about seconds/2 of wall time."""
    if seconds <= 0:
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    x = oracle.fill_random(n, 1)
    model, usable, threads = host_cpu()

    def timed(nthreads, budget):
        oracle.lib().or_set_threads(nthreads)
        times = []
        total = 0.0
        while total < budget and len(times) < 50:
            t0 = time.perf_counter()
            oracle.ntt(x)
            times.append(time.perf_counter() - t0)
            total += times[-1]
        oracle.lib().or_set_threads(1)
        return n / statistics.median(times) / 1e9, len(times)

    par, k_par = timed(threads, seconds / 2)
    one, k_one = timed(1, seconds / 2)
    return {"value": round(par, 5), "unit": "Gelem/s", "cores": threads, "kind": "port",
            "sample": f"forward NTTs of 2^{n.bit_length() - 1} words, median of {k_par} "
                      f"({threads} threads) and of {k_one} (1 thread)",
            "value_1core": round(one, 5), "cpu_model": model, "host_logical_cpus": usable,
            "note": "C port of the reference plan tree; reference serial speed at 2^24 is "
                    "~0.010 Gelem/s on EPYC 7R13 and ~0.031 on M1 Max (bench/analysis.ipynb plots)"}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def _reference_rule(fn, budget, max_reps=50):
    """Median seconds of fn over repeats until the accumulated time passes
    budget (bench/src/main.rs:61-95); returns (median, repeats)."""
    times = []
    while not times or (sum(times) < budget and len(times) < max_reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), len(times)


def _cpu_fields(threads, usable, model):
    return {"cores": threads, "kind": "port", "cpu_model": model, "host_logical_cpus": usable}


def cpu_baseline_commit(seconds):
    """The oracle's commit (the reference's intt / ntt plan tree per column
    and its merkle_tree) on every usable host core, on a bounded sample of
    the commit workload: 2^16 of its 2^20 rows, same 64 columns and blowup."""
    if seconds <= 0:
        return None
    oracle = _oracle()
    model, usable, threads = host_cpu()
    rows, cols, blowup = 1 << 16, 64, 2
    m = oracle.fill_random(rows * cols, 5).reshape(rows, cols)
    oracle.set_threads(threads)
    med, k = _reference_rule(lambda: oracle.commit(m, blowup), seconds)
    oracle.set_threads(1)
    return {"value": round(rows * blowup * cols / med / 1e9, 5), "unit": "Gelem/s",
            "sample": f"commit of {rows} x {cols} words, blowup 2 (1/16 of the workload's rows), "
                      f"median of {k}", **_cpu_fields(threads, usable, model)}


def cpu_baseline_batch(seconds, batch_k=(10, 20), count=1024, seed=7, every=32):
    """The oracle's forward NTTs on every usable host core over a bounded
    sample of the mixed batch: every 32nd of its 1024 lengths."""
    if seconds <= 0:
        return None
    oracle = _oracle()
    from goldilocks_amd.benchmarks import mixed_sizes
    model, usable, threads = host_cpu()
    sizes = mixed_sizes(count, seed, *batch_k)[::every]
    xs = [oracle.fill_random(n, 1000 + i) for i, n in enumerate(sizes)]
    oracle.set_threads(threads)

    def run():
        for x in xs:
            oracle.ntt(x)
    med, k = _reference_rule(run, seconds, max_reps=5)
    oracle.set_threads(1)
    return {"value": round(sum(sizes) / med / 1e9, 5), "unit": "Gelem/s",
            "sample": f"{len(sizes)} of the batch's lengths (one in every {every}, {sum(sizes)} words), median of {k}",
            **_cpu_fields(threads, usable, model)}


def cpu_baseline_permute(seconds):
    """The oracle's transpose (the reference's transpose_copy loop) of a
    4096 x 4096 matrix on one host core (the reference's copy loop is
    serial), 16 bytes per word."""
    if seconds <= 0:
        return None
    oracle = _oracle()
    model, usable, _ = host_cpu()
    n = 4096
    m = oracle.fill_random(n * n, 3)
    med, k = _reference_rule(lambda: oracle.transpose(m, n, n), seconds, max_reps=20)
    return {"value": round(16 * n * n / med / 1e9, 3), "unit": "GB/s",
            "sample": f"transpose of {n} x {n} words, median of {k}",
            **_cpu_fields(1, usable, model)}


def warm_up(fn, steps, torch, min_seconds=1.0, max_seconds=4.0):
    """`steps` untimed calls, then more until at least min_seconds of wall
    time have passed (GPU clocks settle under load), never past
    max_seconds; synchronized at the end."""
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    while time.perf_counter() - t0 < min_seconds:
        for _ in range(20):
            fn()
        torch.cuda.synchronize()
        if time.perf_counter() - t0 > max_seconds:
            break


def other_sizes(torch, gl, log_sizes=(16, 20, 27), seconds=0.3):
    """The other lengths of the plain-NTT workloads (2^16 forward + inverse;
    2^20 and 2^27 forward then inverse, HBM-resident): ms per call over
    back-to-back calls after a warm-up, with the transform's 16 n bytes
    against the 8 TB/s peak.  A roundtrip check guards each size."""
    stream = torch.cuda.current_stream()
    out = {}
    for lg in log_sizes:
        n = 1 << lg
        x = torch.empty(n, dtype=torch.int64, device="cuda")
        gl.fill_random(x, 40 + lg)
        ref = x.clone()
        res = {}
        for name, fn in (("forward", gl.ntt), ("inverse", gl.intt)):
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < seconds:
                fn(x)
                torch.cuda.synchronize()
            reps = max(5, min(2000, int(2e-3 / (1e-9 * n) + 5)))  # ~2 ms of calls at 1 ns/word
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn(x)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            res[name] = {"ms": round(ms, 5), "gelem_s": round(n / (ms * 1e-3) / 1e9, 2),
                         "hbm_frac": round(16 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        x.copy_(ref)
        gl.ntt(x)
        gl.intt(x)
        torch.cuda.synchronize()
        res["roundtrip_exact"] = bool(torch.equal(x, ref))  # intt(ntt(x)) == x
        out[f"2^{lg}"] = res
        del x, ref
        torch.cuda.empty_cache()
    return out


def size_label(n):
    """2^k, or f*2^k for a length with an odd factor f."""
    k = (n & -n).bit_length() - 1
    return f"2^{k}" if n == 1 << k else f"{n >> k}*2^{k}"


def bench_ntt(args, torch, gl):
    n = args.n or 1 << args.log_n
    label = size_label(n)
    x = torch.empty(n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 1)
    # A plain copy of the same n words (8n bytes read, 8n written): the
    # calibration dispatch for FETCH_SIZE / WRITE_SIZE when this command runs
    # under rocprofv3 --pmc.
    y = x.clone()
    del y
    stream = torch.cuda.current_stream()
    warm_up(lambda: gl.ntt(x), args.warmup, torch)
    # Timed region: exactly K steps, synchronized on both sides, bracketed by
    # two events on the launch stream.
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        gl.ntt(x)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # Per-step latency: the same steps with an event after each one.
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record(stream)
    for i in range(args.steps):
        gl.ntt(x)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
    per_sorted = sorted(per)
    p50 = per_sorted[len(per) // 2]
    p99 = per_sorted[min(len(per) - 1, int(0.99 * len(per)))]
    # The inverse transform on the same words, timed the same way (reported
    # beside the forward value).
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        gl.intt(x)
    e1.record(stream)
    torch.cuda.synchronize()
    inv_ms = e0.elapsed_time(e1) / args.steps
    # Kernel durations: the same steps again with every pass launch bracketed
    # by HIP events on the launch stream (kept out of the region above so the
    # extra event records do not perturb the step time).
    gl.timing_collect()
    gl.timing_enable(True)
    for _ in range(args.steps):
        gl.ntt(x)
    torch.cuda.synchronize()
    gl.timing_enable(False)
    kernels = kernel_table(gl.timing_collect())
    # The reference-shaped entry point on a host array (pageable memory):
    # host-to-device copy, transform, device-to-host copy.  Reported beside
    # the HBM-resident value, never as it.
    host = None
    if not args.no_host_buffer:
        import numpy as np
        h = x.cpu().numpy().view(np.uint64).copy()
        gl.ntt(h)
        host_times = []
        for _ in range(3):
            t0 = time.perf_counter()
            gl.ntt(h)
            host_times.append(time.perf_counter() - t0)
        host_s = statistics.median(host_times)
        host = {"gelem_s": round(n / host_s / 1e9, 3), "ms": round(host_s * 1e3, 3),
                "path": "gl_ntt on a pageable host array (PCIe copies included)"}
        # a host batch (strategy(n).ntt on many transforms): copies in,
        # transforms and copies out overlap in chunks
        bn, bc = 1 << 20, 64
        hb = np.concatenate([h[:bn]] * bc)
        s = gl.strategy(bn)
        s.ntt(hb)
        batch_times = []
        for _ in range(3):
            t0 = time.perf_counter()
            s.ntt(hb)
            batch_times.append(time.perf_counter() - t0)
        del hb
        bs = statistics.median(batch_times)
        host["batch"] = {"n": bn, "count": bc, "ms": round(bs * 1e3, 3),
                         "gelem_s": round(bn * bc / bs / 1e9, 3),
                         "path": "strategy(n).ntt on a pageable host array of count transforms"}
    # Achievable streaming rate on this device: a plain device copy of the
    # same n words (8n bytes read, 8n written), timed like the steps.
    buf = torch.empty_like(x)
    for _ in range(3):
        buf.copy_(x)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    c0.record(stream)
    for _ in range(20):
        buf.copy_(x)
    c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 16.0 * n / (c0.elapsed_time(c1) / 20 * 1e-3) / 1e9
    del buf
    plan = gl.plan_info(n)
    # Dominant kernel: largest share of the timed region.  The algorithmic
    # bytes of a whole transform are 16 n (read and write each word once);
    # a schedule of P pass launches is charged 16 n / P per launch.  The
    # bytes a launch actually moves (16 n: every pass reads and writes every
    # word) are reported beside it as launch_hbm_frac.
    kname = max(kernels, key=lambda k: kernels[k]["total_ms"]) if kernels else "none"
    avg_ms = kernels[kname]["total_ms"] / kernels[kname]["launches"] if kernels else ms
    npass = len(plan)
    alg_per_launch = 16.0 * n / npass
    achieved = alg_per_launch / (avg_ms * 1e-3) / 1e9
    moved = 16.0 * n / (avg_ms * 1e-3) / 1e9
    gelem = n / (ms * 1e-3) / 1e9
    return {
        "metric": f"Gelem/s, {label}-point forward Goldilocks NTT (HBM-resident)",
        "value": round(gelem, 3),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(gelem / 0.14, 1),
        "vs_baseline_ref": "M1 Max CPU ntt-par at n~2^24, ~1.4e8 elem/s (bench/analysis.ipynb plot)",
        "dtype": "u64",
        "data": "synthetic: seeded splitmix64 words uniform in [0,p)",
        "config": {"workload": f"forward NTT, n={label}, in place, words resident in HBM",
                   "n": n, "passes": [p["kernel"] for p in plan]},
        "latency_ms": {"p50": round(p50, 5), "p99": round(p99, 5)},
        "inverse": {"ms_per_step": round(inv_ms, 5), "gelem_s": round(n / (inv_ms * 1e-3) / 1e9, 3)},
        "transform_hbm_frac": round(16.0 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": kernels,
        "host_buffer": host,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_for(kname, f"ntt_{label}"),
                     "avg_launch_us": round(avg_ms * 1e3, 2), "launches_per_transform": npass,
                     "algorithmic_bytes_per_launch": int(alg_per_launch),
                     "launch_bytes_moved": 16 * n, "launch_hbm_gbs": round(moved, 1),
                     "launch_hbm_frac": round(moved / HBM_PEAK_GBS, 4),
                     "copy_gbs": round(copy_gbs, 1), "launch_frac_of_copy": round(moved / copy_gbs, 4)},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the library "
              "runs the transforms and the commit on a single MI355X", file=sys.stderr)
        sys.exit(2)
    import torch
    import goldilocks_amd as gl
    torch.cuda.set_device(0)
    rc = gl.lib().gl_init(0)
    if rc:
        raise SystemExit(gl.lib().gl_last_error().decode())
    if args.workload == "ntt":
        out = bench_ntt(args, torch, gl)
        torch.cuda.empty_cache()
        # the plain-NTT workloads' other lengths (2^16, 2^20, 2^27)
        if not args.no_other_sizes:
            out["other_sizes"] = other_sizes(torch, gl)
        if not args.no_batch:
            # the metric's second half: p50/p99 of the mixed-size batch
            from goldilocks_amd.benchmarks import batch_summary
            out["mixed_batch"] = batch_summary(torch, gl, args.batch_k)
            torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline_ntt(args.n or 1 << args.log_n, args.cpu_seconds)
    elif args.workload == "batch":
        from goldilocks_amd.benchmarks import bench_batch
        out = bench_batch(args, torch, gl)
        out["cpu_baseline"] = cpu_baseline_batch(args.cpu_seconds, args.batch_k)
    elif args.workload == "commit":
        from goldilocks_amd.benchmarks import bench_commit
        out = bench_commit(args, torch, gl)
        out["cpu_baseline"] = cpu_baseline_commit(args.cpu_seconds)
    elif args.workload == "permute":
        from goldilocks_amd.benchmarks import bench_permute
        out = bench_permute(args, torch, gl)
        out["cpu_baseline"] = cpu_baseline_permute(min(args.cpu_seconds, 4.0))
    elif args.workload == "hostbatch":
        from goldilocks_amd.benchmarks import bench_hostbatch
        out = bench_hostbatch(args, torch, gl)
    else:
        from goldilocks_amd.benchmarks import bench_serve
        out = bench_serve(args, torch, gl)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
