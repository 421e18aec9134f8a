"""Secondary bench workloads (bench.py --workload batch|commit)."""
import random
import time


def mixed_sizes(count, seed, lo=10, hi=20, factors=(1, 3, 5, 15, 17)):
    """Lengths 2^k * f, k uniform in [lo, hi], f from `factors`."""
    rng = random.Random(seed)
    return [(1 << rng.randint(lo, hi)) * rng.choice(factors) for _ in range(count)]


def bench_batch(args, torch, gl):
    sizes = mixed_sizes(1024, 7)
    total = sum(sizes)
    bufs = [torch.empty(n, dtype=torch.int64, device="cuda") for n in sizes]
    for i, b in enumerate(bufs):
        gl.fill_random(b, i)
    stream = torch.cuda.current_stream()

    def run(batched):
        gl.ntt_batch_mixed(bufs, batched=batched)

    res = {}
    for batched in (True, False):
        run(batched)
        torch.cuda.synchronize()
        reps = max(1, min(args.steps, 5))
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run(batched)
            e1.record(stream)
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        times.sort()
        res["batched" if batched else "unbatched"] = {
            "ms": times[len(times) // 2], "gelem_s": total / (times[len(times) // 2] * 1e-3) / 1e9}
    return {
        "metric": "Gelem/s, mixed-size batched NTT (1024 transforms)",
        "value": round(res["batched"]["gelem_s"], 3),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "ms_per_step": round(res["batched"]["ms"], 3),
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "1024 forward NTTs, n = 2^k * f, k in [10,20], f in {1,3,5,15,17}",
                   "elements": total},
        "unbatched": res["unbatched"],
        "speedup_vs_unbatched": round(res["unbatched"]["ms"] / res["batched"]["ms"], 3),
    }


def bench_commit(args, torch, gl):
    rows, cols, blowup = 1 << 20, 64, 2
    m = torch.empty((rows, cols), dtype=torch.int64, device="cuda")
    gl.fill_random(m, 5)
    stream = torch.cuda.current_stream()
    gl.commit(m, blowup)
    torch.cuda.synchronize()
    times = []
    for _ in range(max(1, min(args.steps, 5))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        gl.commit(m, blowup)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    ms = times[len(times) // 2]
    elems = rows * blowup * cols
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
    }
