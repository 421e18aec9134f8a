"""Autotuner for NTT schedules.

For a length n it enumerates pass lists built from the compiled kernel
variants (radix splits and instances-per-workgroup), times each on the
device with HIP events, checks that every candidate reproduces the default
schedule's output bit for bit, installs the fastest with gl_plan_set and
records it in a JSON cache (GOLDILOCKS_TUNE_CACHE or
~/.cache/goldilocks_amd/tune.json) that load_cache() re-installs.

Usage: python -m goldilocks_amd.autotune 20 24 27   (log2 lengths)
       python -m goldilocks_amd.autotune --cols 64 20 21 --write   (column transforms)
"""
import itertools
import json
import os
import sys

import goldilocks_amd as gl

CACHE = os.environ.get("GOLDILOCKS_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "goldilocks_amd", "tune.json"))


def candidate_plans(n, max_extra_passes=1, min_radix=16):
    vs = [v for v in gl.variants() if v["R"] >= min_radix or v["R"] == n]
    by_r = {}
    for v in vs:
        by_r.setdefault(v["R"], []).append(v["id"])
    radices = sorted(r for r in by_r if n % r == 0)
    out = []
    best_p = None
    for p in range(1, 5):
        found = []
        for combo in itertools.product(radices, repeat=p):
            prod = 1
            for r in combo:
                prod *= r
            if prod != n:
                continue
            for ids in itertools.product(*[by_r[r] for r in combo]):
                found.append(list(ids))
        if found and best_p is None:
            best_p = p
        if best_p is not None:
            out.extend(found)
            if p >= best_p + max_extra_passes:
                break
    return out


def time_plan(torch, x, n, ids, reps=20, cols=0):
    """Time the schedule `ids` (None: the installed one) on x: a length-n
    vector, or with cols > 0 an (n, cols) matrix transformed by columns."""
    if ids is not None:
        gl.plan_set(n, ids, cols=cols > 0)
    run = (lambda t: gl.ntt_cols(t)) if cols else (lambda t: gl.ntt(t))
    y = x.clone()
    run(y)  # warm (and plan install check)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        run(y)
    e1.record(stream)
    torch.cuda.synchronize()
    z = x.clone()
    run(z)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, z


def warm_device(torch, x, cols=0, seconds=1.0):
    """Run the installed schedule for about `seconds` before any timing:
    the engine clock rises under load, and a cold first timing would make
    every later candidate look faster than the schedule timed first."""
    import time
    y = x.clone()
    run = (lambda t: gl.ntt_cols(t)) if cols else (lambda t: gl.ntt(t))
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            run(y)
        torch.cuda.synchronize()


def tune(n, verbose=True, max_candidates=400, cols=0):
    """Time every candidate schedule for length n (contiguous transforms, or
    with cols > 0 the columns of an (n, cols) matrix) and install the fastest."""
    import torch
    x = torch.empty((n, cols) if cols else n, dtype=torch.int64, device="cuda")
    gl.fill_random(x, 12345)
    gl.plan_set(n, [], cols=cols > 0)
    words = n * max(cols, 1)
    reps = 20 if words <= (1 << 22) else 5
    warm_device(torch, x, cols)
    base_ms, ref = time_plan(torch, x, n, None, reps, cols)
    vs_all = {v["id"]: v for v in gl.variants()}
    default = [vs_all[i]["name"] for i in gl.plan_ids(n, cols=cols > 0)]
    results = [(base_ms, None)]
    cands = candidate_plans(n)[:max_candidates]
    for c, ids in enumerate(cands):
        ms, z = time_plan(torch, x, n, ids, reps, cols)
        if not torch.equal(z, ref):
            raise RuntimeError(f"schedule {ids} for n={n} is not bit-exact")
        results.append((ms, ids))
        if verbose and (c + 1) % 100 == 0:
            print(f"  n={n}: {c + 1}/{len(cands)} candidates timed", flush=True)
    gl.plan_set(n, [], cols=cols > 0)
    # the default again, after the candidates: the better of its two timings
    # stands, so a clock still rising during the first one does not favour
    # the candidates
    base2, _ = time_plan(torch, x, n, None, reps, cols)
    base_ms = min(base_ms, base2)
    results[0] = (base_ms, None)
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    if best is not None:
        gl.plan_set(n, best, cols=cols > 0)
    if verbose:
        vs = {v["id"]: v for v in gl.variants()}
        def desc(ids):
            return "default " + str(default) if ids is None else \
                " + ".join(tune_key(vs[i]) for i in ids)
        what = f"n={n}" + (f" x {cols} columns" if cols else "")
        print(f"{what}: {len(cands)} candidates; default {base_ms * 1e3:.1f} us", flush=True)
        for ms, ids in results[:8]:
            print(f"   {ms * 1e3:9.1f} us  {desc(ids)}", flush=True)
    return best, best_ms, base_ms


TUNE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_gfx950.txt")


def tune_key(v):
    """Variant as written in the tune file: R1xR2/T."""
    return f"{v['R1']}x{v['R2']}/{v['T']}"


def write_tune_file(entries, path=TUNE_FILE):
    """Write schedules in the text format the library loads at init:
    "n R1xR2/T R1xR2/T ..." per line, "cn ..." for column transforms
    (merged with existing lines).  R2 is the variant's R2 field (R2*R3 for
    three-stage kernels); (R1, R2, T) names one compiled variant.  entries
    maps n (or the string "c<n>") to a list of variant ids (None: remove)."""
    vs = {v["id"]: v for v in gl.variants()}
    keys = [tune_key(v) for v in vs.values()]
    assert len(keys) == len(set(keys)), "tune-file keys must name one variant each"

    def order(key):
        return (key.startswith("c"), int(key.lstrip("c")))

    lines = {}
    if os.path.exists(path):
        for line in open(path):
            if line.strip() and not line.startswith("#"):
                lines[line.split()[0]] = line.strip()
    for n, ids in entries.items():
        key = str(n)
        if ids:
            lines[key] = f"{key} " + " ".join(tune_key(vs[i]) for i in ids)
        else:
            lines.pop(key, None)
    with open(path, "w") as f:
        f.write("# autotuned NTT schedules for gfx950 (python -m goldilocks_amd.autotune --write ...)\n")
        for key in sorted(lines, key=order):
            f.write(lines[key] + "\n")


def save(entries):
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    data = {}
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            data = json.load(f)
    data.update({str(k): v for k, v in entries.items()})
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1)


def load_cache(path=CACHE):
    """Install every tuned schedule from the cache file; returns how many."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        data = json.load(f)
    k = 0
    for n, ids in data.items():
        if ids:
            gl.plan_set(int(n), ids)
            k += 1
    return k


if __name__ == "__main__":
    import torch  # noqa: F401
    args = sys.argv[1:]
    write = "--write" in args
    out = None
    if "--out" in args:
        out = args[args.index("--out") + 1]
        args = [a for a in args if a not in ("--out", out)]
    cols = 0
    if "--cols" in args:  # tune the column transforms of (n, cols) matrices
        cols = int(args[args.index("--cols") + 1])
        args = [a for i, a in enumerate(args) if a != "--cols" and (i == 0 or args[i - 1] != "--cols")]
    args = [a for a in args if a != "--write"]
    gl.lib().gl_init(0)
    for arg in args:
        n = int(arg[1:]) if arg.startswith("=") else 1 << int(arg)
        best, best_ms, base_ms = tune(n, cols=cols)
        # keep the default when the win is within timing noise
        ids = best if best is not None and best_ms < 0.97 * base_ms else None
        key = f"c{n}" if cols else n
        # write after every length so a cut-off run keeps what it measured
        if not cols:
            save({n: ids})
        if write or out:
            write_tune_file({key: ids}, out or TUNE_FILE)
