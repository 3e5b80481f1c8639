#!/usr/bin/env python3
"""Benchmark of the MI355X SVDAG editing back-end (one GPU).

Headline line: the slab-hash microbenchmark -- 2^24 DAG interior-node items
(2..9 words, child mask uniform over the 255 non-empty masks, child words
uniform u32) bulk-inserted into per-size tables at 96 items per bucket, then
2^24 finds with a 50% hit rate.  One step = one find pass over the 2^24
needles already resident in HBM.  The wave64-group and per-lane variants are
all timed during warm-up; the fastest is used for the timed steps.

roofline: algorithmic bytes per find = 4*s (needle) + 4 (result) + 64 (bucket
probe) + 64 * hit (matched item), summed over the launch, divided by the find
kernel's average duration from HIP events on the library's stream; peak is
the MI355X HBM3E 8 TB/s.  cpu_baseline: the oracle's restatement of the
reference's CompactAccelerationHashTable host find, on host threads, over the
same find set.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-svdag-editing_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "voxel edits/s & p99 edit latency (depth 15); slab-hash find Mops/s vs HBM"
HBM_PEAK_GBS = 8000.0
VARIANTS = [0, 4, 8, 16, 32, 204, 404, 208, 408]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def variant_name(v):
    if v == 0:
        return "per-lane"
    if v >= 100:
        return f"wave64 group of {v % 100} lanes, {v // 100} finds in flight per group"
    return f"wave64 group of {v} lanes"


def read_traffic(path):
    """HBM bytes per find launch from a committed rocprofv3 --pmc summary, if any."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def slab_hash_bench(args):
    import svdag_py
    import workloads

    n_items = 1 << args.log2_items
    log(f"[bench] generating {n_items} items / finds")
    items = workloads.slab_hash_items(n_items)
    finds = workloads.slab_hash_finds(items, n_items, hit_rate=0.5)
    buckets = workloads.buckets_for(items, 96)
    sizes = list(range(2, 10))
    groups = sorted(items.items())
    fgroups = [(s, finds[s]) for s, _ in groups]

    words, segs, n = svdag_py.make_segments(groups)
    fwords, fsegs, fn = svdag_py.make_segments(fgroups)
    wdev = svdag_py.DeviceBuffer.from_array(words)
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    odev = svdag_py.DeviceBuffer(4 * n)
    rdev = svdag_py.DeviceBuffer(4 * fn)

    log("[bench] bulk insert")
    hs = svdag_py.HashSet(2, 9, [buckets.get(s, 1) for s in sizes], [len(items.get(s, [])) for s in sizes])
    add_ms = hs.add_dev(wdev, segs, odev, 16)

    if args.pmc_mode:
        # Counter runs: one device copy of a known size (calibrates FETCH_SIZE /
        # WRITE_SIZE) followed by a few launches of one find variant only.
        nbytes = 256 << 20
        a = svdag_py.DeviceBuffer(nbytes)
        b = svdag_py.DeviceBuffer(nbytes)
        svdag_py.check(svdag_py.lib().svdag_memcpy_d2d(b.ptr, a.ptr, svdag_py.ctypes.c_size_t(nbytes)))
        for _ in range(args.steps):
            hs.find_dev(fdev, fsegs, rdev, args.variant)
        log(f"[bench] pmc mode: copied {nbytes} bytes, {args.steps} find launches of variant {args.variant}")
        return None

    # Algorithmic bytes of one find launch.
    res = None
    per_variant = {}
    for v in ([args.variant] if args.variant is not None else VARIANTS):
        ms = [hs.find_dev(fdev, fsegs, rdev, v) for _ in range(max(2, args.warmup))]
        per_variant[v] = min(ms)
        log(f"[bench] variant {v}: {min(ms):.3f} ms")
    best = min(per_variant, key=per_variant.get)
    res = rdev.to_array(np.uint32, fn)
    hits = (res != svdag_py.NOT_FOUND)
    s_per = np.concatenate([np.full(len(a), s) for s, a in fgroups]).astype(np.float64)
    alg_bytes = float(np.sum(4 * s_per + 4 + 64 + 64 * hits))

    log(f"[bench] timed: {args.steps} steps, variant {best}")
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(hs.find_dev(fdev, fsegs, rdev, best))
    svdag_py.lib().svdag_device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = fn * args.steps / wall / 1e6
    avg_kernel_ms = float(np.mean(kernel_ms))
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(items, finds, buckets, args)

    hs.close()
    traffic = read_traffic(os.path.join(ROOT, "profiles", "slab_hash_find_traffic.txt"))
    return {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mfinds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"slab-hash find: 2^{args.log2_items} DAG node items of 2-9 u32 in per-size tables "
                        f"at 96 items/bucket, 2^{args.log2_items} finds, 50% hits",
            "variant": variant_name(best),
            "find_kernel_ms_by_variant": {str(k): round(v, 4) for k, v in per_variant.items()},
            "bulk_insert_ms": round(add_ms, 3),
            "hit_rate": round(float(hits.mean()), 4),
            "kernel_ms": round(avg_kernel_ms, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }


def cpu_baseline(items, finds, buckets, args):
    import oracle_py

    threads = min(args.cpu_threads, os.cpu_count() or 1)
    log(f"[bench] cpu baseline: oracle tables, {threads} threads")
    tables = {}
    for s, a in items.items():
        t = oracle_py.OracleTable(s, buckets[s])
        t.add(a)
        tables[s] = t
    total = 0
    t0 = time.perf_counter()
    reps = 0
    while True:
        for s, a in finds.items():
            tables[s].find(a, threads=threads)
            total += len(a)
        reps += 1
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    dt = time.perf_counter() - t0
    return {
        "value": round(total / dt / 1e6, 2),
        "unit": "Mfinds/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} pass(es) over the same {sum(len(a) for a in finds.values())} finds "
                  f"(oracle restatement of CompactAccelerationHashTable::find)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-items", type=int, default=24)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=None, help="skip the variant sweep and use this one")
    ap.add_argument("--pmc-mode", action="store_true", help="calibration copy + find launches only (for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this back-end has no multi-GPU path "
              "(see the sharding notes in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    out = slab_hash_bench(args)
    if out is not None:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
