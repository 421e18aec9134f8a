"""Secondary bench workloads (bench.py --workload batch|commit)."""
import random
import time


def mixed_sizes(count, seed, lo=10, hi=20, factors=(1, 3, 5, 15, 17, 257)):
    """Lengths 2^k * f, k uniform in [lo, hi], f from `factors`."""
    rng = random.Random(seed)
    return [(1 << rng.randint(lo, hi)) * rng.choice(factors) for _ in range(count)]


def _timed_reps(torch, stream, fn, min_total_ms=1000.0, min_reps=5, max_reps=40):
    """Per-repetition milliseconds (HIP events on `stream`) after one
    untimed call, repeated until the accumulated time passes min_total_ms
    (the reference's rule, bench/src/main.rs:61-95), within the rep bounds."""
    fn()
    torch.cuda.synchronize()
    times = []
    while len(times) < max_reps and (len(times) < min_reps or sum(times) < min_total_ms):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return times


def _pct(times, q):
    t = sorted(times)
    return t[min(len(t) - 1, int(q * len(t)))]


def batch_summary(torch, gl, batch_k=(10, 20), count=1024, seed=7):
    """The mixed-size batch (count forward transforms, n = 2^k * f, k
    uniform in batch_k, f in {1, 3, 5, 15, 17, 257}, device-resident), timed
    three ways: grouped launches (gl_ntt_mixed_dev), one launch sequence per
    transform, and the grouped schedule captured once as a hipGraph and
    replayed.  Each: p50/p99 batch latency over the repetitions, Gelem/s at
    p50."""
    lo, hi = batch_k
    sizes = mixed_sizes(count, seed, lo, hi)
    total = sum(sizes)
    bufs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, b in enumerate(bufs):
        gl.fill_random(b, i)
    stream = torch.cuda.current_stream()
    res = {}
    for key, fn in (("batched", lambda: gl.ntt_batch_mixed(bufs)),
                    ("unbatched", lambda: gl.ntt_batch_mixed(bufs, batched=False))):
        times = _timed_reps(torch, stream, fn)
        res[key] = {"p50_ms": round(_pct(times, 0.5), 3), "p99_ms": round(_pct(times, 0.99), 3),
                    "reps": len(times), "gelem_s": round(total / (_pct(times, 0.5) * 1e-3) / 1e9, 3)}
    plan = gl.BatchPlan(bufs)
    times = _timed_reps(torch, stream, lambda: plan.launch(stream))
    plan.close()
    res["graph"] = {"p50_ms": round(_pct(times, 0.5), 3), "p99_ms": round(_pct(times, 0.99), 3),
                    "reps": len(times), "gelem_s": round(total / (_pct(times, 0.5) * 1e-3) / 1e9, 3)}
    # Summed kernel time of one more repetition of each eager mode (every pass
    # launch bracketed by HIP events): the floor no launch scheme can beat
    # with these kernels.  wall / kernel is the most any batching of the
    # same passes could gain over that mode.
    for key, fn in (("batched", lambda: gl.ntt_batch_mixed(bufs)),
                    ("unbatched", lambda: gl.ntt_batch_mixed(bufs, batched=False))):
        gl.timing_collect()
        gl.timing_enable(True)
        fn()
        torch.cuda.synchronize()
        gl.timing_enable(False)
        recs = gl.timing_collect()
        kern = sum(ms for _, ms in recs)
        res[key]["kernel_ms"] = round(kern, 3)
        res[key]["launches"] = len(recs)
        res[key]["wall_over_kernel"] = round(res[key]["p50_ms"] / kern, 3) if kern else None
    del bufs
    return {
        "workload": f"{count} forward NTTs, n = 2^k * f, k uniform in [{lo},{hi}], f in {{1,3,5,15,17,257}}, "
                    f"seed {seed}, device-resident",
        "elements": total,
        "latency_ms": {"p50": res["batched"]["p50_ms"], "p99": res["batched"]["p99_ms"]},
        "gelem_s": res["batched"]["gelem_s"],
        "hbm_frac": round(16.0 * total / (res["batched"]["p50_ms"] * 1e-3) / 8e12, 4),
        "batched": res["batched"], "unbatched": res["unbatched"], "graph": res["graph"],
        "speedup_vs_unbatched": round(res["unbatched"]["p50_ms"] / res["batched"]["p50_ms"], 3),
        "graph_speedup_vs_unbatched": round(res["unbatched"]["p50_ms"] / res["graph"]["p50_ms"], 3),
    }


def bench_batch(args, torch, gl):
    """The mixed-size batch as its own line (batch_summary), reported value
    = grouped launches."""
    sm = batch_summary(torch, gl, args.batch_k)
    return {
        "metric": "Gelem/s, mixed-size batched NTT (1024 transforms); p50/p99 batch latency",
        "value": sm["gelem_s"],
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": sm["batched"]["reps"],
        "ms_per_step": sm["batched"]["p50_ms"],
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic: seeded uniform [0,p) words, device-resident",
        "config": {"workload": sm["workload"], "elements": sm["elements"], "bytes": 8 * sm["elements"]},
        **{k: sm[k] for k in ("latency_ms", "hbm_frac", "batched", "unbatched", "graph", "speedup_vs_unbatched",
                              "graph_speedup_vs_unbatched")},
    }


def bench_commit(args, torch, gl):
    rows, cols, blowup = 1 << 20, 64, 2
    m = torch.empty((rows, cols), dtype=torch.int64, device="cuda")
    gl.fill_random(m, 5)
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        times = []
        for _ in range(max(1, min(args.steps, 10))):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        times.sort()
        return times[len(times) // 2], times[-1]

    ms, ms_max = timed(lambda: gl.commit(m, blowup))
    rows_ms, rows_max = timed(lambda: gl.commit_rows(m, blowup))
    lde, _ = gl.commit(m, blowup)
    tree_ms, _ = timed(lambda: gl.merkle_tree(lde))
    row_tree_ms, _ = timed(lambda: gl.row_tree(lde))
    elems = rows * blowup * cols
    # Algorithmic bytes of a commit: read the input matrix, write the LDE,
    # write every tree layer once (hashing assumed fused with the last pass).
    algo = 8 * rows * cols + 8 * blowup * rows * cols + (gl.merkle_tree_size(8 * blowup * rows * cols))
    achieved = algo / (ms * 1e-3) / 1e9
    return {
        "metric": "Gelem/s, Merkle commit 2^20 x 64, blowup 2 (LDE elements)",
        "value": round(elems / (ms * 1e-3) / 1e9, 3),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "column coset LDE (2^21-point NTTs) + K12 Merkle tree over 1 GiB",
                   "rows": rows, "cols": cols, "blowup": blowup},
        "latency_ms": {"p50": round(ms, 3), "max": round(ms_max, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "algorithmic_bytes": algo},
        "tree_ms": round(tree_ms, 3),
        # the same extension with the binary tree of row hashes (2^21 leaves
        # of 512 bytes, four permutations each, then 2^21 - 1 parents)
        "row_tree": {"commit_ms": round(rows_ms, 3), "commit_max_ms": round(rows_max, 3),
                     "tree_ms": round(row_tree_ms, 3),
                     "value": round(elems / (rows_ms * 1e-3) / 1e9, 3), "unit": "Gelem/s"},
    }


SERVE_FACTORS = (1, 3, 5, 15, 17, 257)


def serve_jobs(count, seed, max_elems=1 << 24):
    """Seeded job list: ("ntt", n, cols) or ("commit", rows, cols).
    n = 2^k * f with k in [12, 22], f from SERVE_FACTORS; 1-64 columns;
    n * cols capped at max_elems by lowering k, then cols."""
    rng = random.Random(seed)
    jobs = []
    for _ in range(count):
        cols = rng.randint(1, 64)
        if rng.random() < 0.25:
            k = rng.randint(12, 18)
            while (1 << k) * 2 * cols > max_elems:
                cols = max(1, cols // 2)
            jobs.append(("commit", 1 << k, cols))
        else:
            f = rng.choice(SERVE_FACTORS)
            k = rng.randint(12, 22)
            while k > 12 and (1 << k) * f > max_elems:
                k -= 1
            n = (1 << k) * f
            while n * cols > max_elems and cols > 1:
                cols //= 2
            jobs.append(("ntt", n, cols))
    return jobs


def bench_serve(args, torch, gl):
    """Open-loop Poisson arrivals of NTT and commit jobs into the
    continuous-batching server; p50/p99 latency, throughput, admissions."""
    count = 1000
    jobs = serve_jobs(count, 11)
    elems = [n * c if kind == "ntt" else 2 * n * c for kind, n, c in jobs]
    # Offered load: about 60% of a 60 Gelem/s device, as a fixed QPS.
    qps = args.qps or 0.6 * 60e9 / (sum(elems) / count)
    rng = random.Random(13)
    t, arrivals = 0.0, []
    for _ in jobs:
        t += rng.expovariate(qps)
        arrivals.append(t)
    bufs = []
    for i, (kind, n, c) in enumerate(jobs):
        if kind == "ntt":
            x = torch.empty(n * c, dtype=torch.int64, device="cuda")
            gl.fill_random(x, 100 + i)
            bufs.append((x,))
        else:
            m = torch.empty((n, c), dtype=torch.int64, device="cuda")
            gl.fill_random(m, 100 + i)
            lde = torch.empty((2 * n, c), dtype=torch.int64, device="cuda")
            layers = torch.empty(gl.merkle_tree_size(2 * n * c * 8), dtype=torch.uint8, device="cuda")
            bufs.append((m, lde, layers))
    torch.cuda.synchronize()
    budget = 32 << 30
    slots = getattr(args, "slots", 8) or 8
    with gl.Server(hbm_budget_bytes=budget, slots=slots) as srv:
        # warm the plans and twiddle tables of every shape once
        seen = set()
        for (kind, n, c), b in zip(jobs, bufs):
            if (kind, n) in seen:
                continue
            seen.add((kind, n))
            tk = srv.submit_ntt(b[0], n) if kind == "ntt" else srv.submit_commit(b[0], 2, b[1], b[2])
            srv.wait(tk)
        base = srv.stats()
        tickets = []
        t0 = time.perf_counter()
        for (kind, n, c), b, a in zip(jobs, bufs, arrivals):
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit_ntt(b[0], n) if kind == "ntt" else srv.submit_commit(b[0], 2, b[1], b[2]))
        lat = [srv.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat.sort()
    total = sum(elems)
    # The plain path for the same jobs: one job at a time through the
    # ordinary entry points on one stream, waiting for each (what a client
    # without the server does).
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for (kind, n, c), b in zip(jobs, bufs):
        if kind == "ntt":
            gl.strategy(n).ntt(b[0])
        else:
            gl.commit(b[0], 2)
        torch.cuda.synchronize()
    serial_s = time.perf_counter() - t1
    # The plain one-launch-per-transform path: each transform of an NTT job
    # on its own (one launch sequence per transform), a commit as one call,
    # waiting for each job.
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for (kind, n, c), b in zip(jobs, bufs):
        if kind == "ntt":
            for j in range(c):
                gl.ntt(b[0][j * n:(j + 1) * n])
        else:
            gl.commit(b[0], 2)
        torch.cuda.synchronize()
    per_transform_s = time.perf_counter() - t2
    return {
        "metric": "Gelem/s and p50/p99 latency, served mixed NTT/commit jobs (open-loop Poisson)",
        "value": round(total / wall / 1e9, 3),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic: seeded uniform [0,p) words, device-resident buffers",
        "config": {"workload": f"{count} jobs, 75% NTT (n = 2^k*f, k in [12,22], f in {list(SERVE_FACTORS)}, "
                               "1-64 transforms) and 25% commits (2^12..2^18 rows x 1-64 cols, blowup 2), "
                               "n*cols <= 2^24", "qps": round(qps, 1), "slots": slots,
                   "hbm_budget_bytes": budget},
        "latency_ms": {"p50": round(lat[len(lat) // 2], 4), "p99": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 4),
                       "max": round(lat[-1], 4)},
        "wall_s": round(wall, 4),
        "serial": {"wall_s": round(serial_s, 4), "gelem_s": round(total / serial_s / 1e9, 3),
                   "path": "the same jobs one at a time through gl.ntt / gl.commit, synchronized per job"},
        "speedup_vs_serial": round(serial_s / wall, 3),
        "per_transform": {"wall_s": round(per_transform_s, 4), "gelem_s": round(total / per_transform_s / 1e9, 3),
                          "path": "one launch sequence per transform (gl.ntt on each transform of an NTT job; "
                                  "gl.commit per commit job), synchronized per job"},
        "speedup_vs_per_transform": round(per_transform_s / wall, 3),
        "elements": total,
        "admitted": int(st["completed"] - base["completed"]),
        "rejected": int(st["rejected"] - base["rejected"]),
        "peak_bytes": int(st["peak_bytes"]),
    }


def bench_permute(args, torch, gl):
    """The reference's public permutations on device-resident words
    (permute::transpose, permute::bit_reverse): median time per call (over
    repetitions of 20 back-to-back calls) and the rate against the 16 bytes
    per word they must move (read + write)."""
    stream = torch.cuda.current_stream()
    cases = []

    calls = 20  # back-to-back calls per timed repetition (a call is ~40-90 us)

    def repeated(fn):
        def run():
            for _ in range(calls):
                fn()
        return run

    def case(name, words, fn, note):
        times = _timed_reps(torch, stream, repeated(fn), min_total_ms=300.0, min_reps=10, max_reps=200)
        ms = sorted(times)[len(times) // 2] / calls
        gbs = 16 * words / (ms * 1e-3) / 1e9
        cases.append({"case": name, "words": words, "ms_p50": round(ms, 4), "algorithmic_gbs": round(gbs, 1),
                      "hbm_frac": round(gbs / 8000.0, 4), "path": note})

    n = 4096
    sq = torch.empty(n * n, dtype=torch.int64, device="cuda")
    gl.fill_random(sq, 3)
    case(f"transpose {n}x{n} in place", n * n, lambda: gl.transpose(sq, n, n),
         "tile pairs swapped through LDS, one read and one write per word")
    r, c = 8192, 2048
    src = torch.empty(r * c, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    gl.fill_random(src, 4)
    case(f"transpose {r}x{c} out of place", r * c, lambda: gl.transpose(src, r, c, out=dst),
         "32x32 tiles through LDS")
    case(f"transpose {r}x{c} in place", r * c, lambda: gl.transpose(src, r, c),
         "out of place into pool scratch, then a device copy back (32 B per word moved)")
    k = 24
    v = torch.empty(1 << k, dtype=torch.int64, device="cuda")
    gl.fill_random(v, 5)
    case(f"bit_reverse 2^{k}", 1 << k, lambda: gl.bit_reverse(v),
         "tile pairs (middle bits b, rev(b)) swapped through LDS")
    copy_times = _timed_reps(torch, stream, repeated(lambda: dst.copy_(src)), min_total_ms=300.0, min_reps=10,
                             max_reps=200)
    copy_ms = sorted(copy_times)[len(copy_times) // 2] / calls
    head = cases[0]
    return {
        "metric": "GB/s of the reference's permutations on HBM-resident u64 words (16 B per word)",
        "value": head["algorithmic_gbs"], "unit": "GB/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "u64", "data": "synthetic: seeded uniform [0,p) words, device-resident",
        "config": {"workload": head["case"]}, "cases": cases,
        "copy_gbs": round(16 * r * c / (copy_ms * 1e-3) / 1e9, 1),
    }


def bench_hostbatch(args, torch, gl):
    """The host-slice batch path (strategy(n).ntt on a pageable numpy array
    of count transforms, the reference's own calling convention): one call
    per batch, whose copies in, transforms and copies out overlap in chunks,
    against one call per transform on the same words (each a copy in, the
    transform and a copy out, one after another).  Median of 5 wall times
    each; rates count the words moved both ways over PCIe."""
    import numpy as np
    rng = np.random.default_rng(11)
    cases = []
    for n, count in ((1 << 20, 64), (1 << 22, 16), (5 << 12, 4000), (1 << 24, 8)):
        x = rng.integers(0, (1 << 64) - (1 << 32), size=n * count, dtype=np.uint64)
        s = gl.strategy(n)

        def timed(fn):
            fn()
            ts = []
            for _ in range(5):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return sorted(ts)[2]

        def per_transform():
            for j in range(count):
                s.ntt(x[j * n:(j + 1) * n])

        tb = timed(lambda: s.ntt(x))
        tp = timed(per_transform)
        cases.append({"n": n, "count": count, "mib": round(n * count * 8 / 2**20),
                      "batch_ms": round(tb * 1e3, 3), "batch_gelem_s": round(n * count / tb / 1e9, 3),
                      "batch_pcie_gbs": round(16 * n * count / tb / 1e9, 1),
                      "per_transform_ms": round(tp * 1e3, 3),
                      "per_transform_gelem_s": round(n * count / tp / 1e9, 3),
                      "speedup": round(tp / tb, 3)})
        del x
    head = cases[0]
    return {
        "metric": "Gelem/s of NTT batches on pageable host arrays (PCIe copies included)",
        "value": head["batch_gelem_s"], "unit": "Gelem/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "u64", "data": "synthetic: seeded uniform [0,p) words in host memory",
        "config": {"workload": f"strategy({head['n']}).ntt on {head['count']} transforms in one host array"},
        "cases": cases,
    }
