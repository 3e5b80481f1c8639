#!/usr/bin/env python3
"""Benchmark of the MI355X SVDAG editing back-end (one GPU).

Headline line: the slab-hash microbenchmark -- 2^24 DAG interior-node items
(2..9 words, child mask uniform over the 255 non-empty masks, child words
uniform u32) bulk-inserted into per-size tables whose head lines are 75% full, then
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
VARIANTS = [0, 1, 2, 3, 4, 8, 16, 32, 204, 404, 208, 408, 1004, 1008, 1204, 1208]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def variant_name(v):
    if v == 0:
        return "per-lane"
    if v in (1, 2, 3):
        return f"per-lane, blocks assigned to item-size segments, {1 << (v - 1)} finds in flight per lane"
    if v >= 1000:
        n = max(1, (v % 1000) // 100)
        return (f"XCD-binned (bin pass + find pass, each XCD probes its own bucket range), wave64 group of "
                f"{v % 100} lanes, {n} finds in flight per group")
    if v >= 100:
        return f"wave64 group of {v % 100} lanes, {v // 100} finds in flight per group"
    return f"wave64 group of {v} lanes"


def read_traffic(path, variant, fill):
    """HBM bytes per find launch of `variant` at head-line fill `fill` from the
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated
    against a 256 MiB copy), if it holds that case."""
    key = f"traffic_bytes_per_launch[{variant}@{fill:.2f}]:"
    try:
        for line in open(path):
            if line.startswith(key):
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
    sizes = list(range(2, 10))
    groups = sorted(items.items())
    fgroups = [(s, finds[s]) for s, _ in groups]

    words, segs, n = svdag_py.make_segments(groups)
    fwords, fsegs, fn = svdag_py.make_segments(fgroups)
    wdev = svdag_py.DeviceBuffer.from_array(words)
    fdev = svdag_py.DeviceBuffer.from_array(fwords)
    odev = svdag_py.DeviceBuffer(4 * n)
    rdev = svdag_py.DeviceBuffer(4 * fn)

    # Bucket count: the head-line fill each table is sized for, picked by
    # timing the find (variant 1, best of three) at each candidate; within 2%
    # of the fastest the fuller tables win (less memory).
    fill_sweep = None
    fill = args.fill if args.fill is not None else (0.75 if args.pmc_mode else None)
    if fill is None:
        fill_sweep = {}
        for f in (0.5, 0.6, 0.75, 0.9):
            bf = workloads.buckets_for(items, f)
            t = svdag_py.HashSet(2, 9, [bf.get(s, 1) for s in sizes], [len(items.get(s, [])) for s in sizes])
            t.add_dev(wdev, segs, odev, 16)
            ms = min(t.find_dev(fdev, fsegs, rdev, 1) for _ in range(3))
            lines_f = sum(t.usage(s)[2] for s in sizes if s in items)
            t.close()
            fill_sweep[f] = {"find_ms": round(ms, 4), "bytes_per_stored_item": round(lines_f * 128.0 / n_items, 1)}
        fastest = min(v["find_ms"] for v in fill_sweep.values())
        fill = max(f for f, v in fill_sweep.items() if v["find_ms"] <= 1.02 * fastest)
        log(f"[bench] head-line fill sweep: {fill_sweep}, picked {fill}")
    buckets = workloads.buckets_for(items, fill)

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
    # The steps are queued back to back on the set's stream (a find launch
    # returns once queued); HIP events on that stream around them give the
    # launches' average device time, gaps between launches included.
    timed_stream = best < 1000
    svdag_py.lib().svdag_device_synchronize()
    t0 = time.perf_counter()
    if timed_stream:
        hs.timer(False)
        for _ in range(args.steps):
            hs.find_dev(fdev, fsegs, rdev, best, timed=False)
        region_ms = hs.timer(True)
        avg_kernel_ms = region_ms / args.steps
    else:
        kernel_ms = [hs.find_dev(fdev, fsegs, rdev, best) for _ in range(args.steps)]
        avg_kernel_ms = float(np.mean(kernel_ms))
    svdag_py.lib().svdag_device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = fn * args.steps / wall / 1e6
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9

    # Practical ceilings on this GPU: a 4 GiB stream read, and one random
    # line-head read (a line's first 32 bytes: the tags and the link, all a
    # probe of a chainless bucket touches) per find over a buffer the size
    # of the tables.
    probe = None
    if not args.no_probe:
        table_bytes = sum(hs.usage(s)[1] for s in sizes if s in items) * 128
        stream_bytes = 4 << 30
        stream_ms = svdag_py.hbm_probe(svdag_py.PROBE_STREAM, stream_bytes)
        line_ms = svdag_py.hbm_probe(svdag_py.PROBE_RANDOM_LINES, table_bytes, fn)
        probe = {
            "stream_read_GBs": round(stream_bytes / (stream_ms * 1e-3) / 1e9, 1),
            "random_line_buffer_MB": round(table_bytes / 2**20, 1),
            "random_lines": fn,
            "random_lines_ms": round(line_ms, 4),
            "random_line_head_reads_G_per_s": round(fn / (line_ms * 1e-3) / 1e9, 2),
            "find_kernel_over_random_lines": round(avg_kernel_ms / line_ms, 3),
        }
        log(f"[bench] probes: {probe}")

    # Bytes per stored item of this layout (every line handed out, head and
    # overflow, over the items stored).
    lines_used = sum(hs.usage(s)[2] for s in sizes if s in items)
    bytes_per_item = lines_used * 128.0 / n_items
    hs.close()

    # The slab-hash workload's sizing as written: ceil(n_s / 96) buckets per size
    # (the reference's 96 items per 32-slot bucket).  Here a bucket is one
    # 128-byte line plus its chain, so buckets hold 96 items in long chains.
    as_written = None
    if not args.no_as_written:
        b96 = {s: max(1, -(-len(a) // 96)) for s, a in items.items()}
        hs96 = svdag_py.HashSet(2, 9, [b96.get(s, 1) for s in sizes], [len(items.get(s, [])) for s in sizes])
        hs96.add_dev(wdev, segs, odev, 16)
        ms96 = {v: min(hs96.find_dev(fdev, fsegs, rdev, v) for _ in range(3)) for v in (1, 16, 208)}
        res96 = rdev.to_array(np.uint32, fn)
        if not np.array_equal(res96 != svdag_py.NOT_FOUND, hits):
            raise RuntimeError("96-per-bucket tables disagree with the 75%-fill tables on which finds hit")
        v96 = min(ms96, key=ms96.get)
        lines96 = sum(hs96.usage(s)[2] for s in sizes if s in items)
        hs96.close()
        as_written = {"buckets": "ceil(n_s / 96) per item size", "find_ms_by_variant": {str(k): round(v, 4)
                      for k, v in ms96.items()}, "G_finds_per_s": round(fn / (ms96[v96] * 1e-3) / 1e9, 2),
                      "roofline_frac": round(alg_bytes / (ms96[v96] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "bytes_per_stored_item": round(lines96 * 128.0 / n_items, 1)}
        log(f"[bench] 96 items per bucket: {as_written}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(items, finds, buckets, args)

    traffic = read_traffic(os.path.join(ROOT, "profiles", "slab_hash_find_traffic.txt"), best, fill)
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
                        f"with head lines {round(100 * fill)}% full, 2^{args.log2_items} finds, 50% hits",
            "head_line_fill": fill,
            "head_line_fill_sweep": {str(k): v for k, v in fill_sweep.items()} if fill_sweep else None,
            "variant": variant_name(best),
            "find_kernel_ms_by_variant": {str(k): round(v, 4) for k, v in per_variant.items()},
            "bulk_insert_ms": round(add_ms, 3),
            "hit_rate": round(float(hits.mean()), 4),
            "kernel_ms": round(avg_kernel_ms, 4),
            "bytes_per_stored_item": round(bytes_per_item, 1),
            "as_written_96_items_per_bucket": as_written,
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
        "hbm_probe": probe,
    }


def shading_bench(g, cam):
    """Frame passes at 1920x1080 on the DAG of g, each run three times (the
    best kept): primary paths, shadows and ambient occlusion (3 samples, rays
    of 2^(D-10) voxels: the reference's defaults, engine.h:202-213), the
    occlusion blur, colours (16 one-colour materials), lighting and one
    path-traced sample of depth 3."""
    D = g.levels
    pos, rot, fov, bmin, bmax = cam
    tr = [g.trace(pos, rot, fov, bmin, bmax, 1920, 1080)[1] for _ in range(3)]
    sh = [g.trace_shadows(1.0)[1] for _ in range(3)]
    ao = [g.trace_ambient_occlusion(3, 1.0, float(1 << (D - 10)), 123 + i)[1] for i in range(3)]
    blur = [g.blur_ambient_occlusion()[1] for _ in range(3)]
    pal = np.random.default_rng(3).uniform(0.2, 0.9, (16, 3)).astype(np.float32)
    col = [g.trace_colors(pal)[1] for _ in range(3)]
    lit = [g.trace_lighting(True, True, 0.0, 0.0)[1] for _ in range(3)]
    pt = [g.trace_path_tracing(pal, env=1.0, samples=1, depth=3, integrate=False, seed=1 + i,
                               accumulate=False)[2] for i in range(3)]
    shading = {"width": 1920, "height": 1080, "primary_ms": round(min(tr), 3), "shadows_ms": round(min(sh), 3),
               "ambient_occlusion_3spp_ms": round(min(ao), 3), "ao_blur_ms": round(min(blur), 3),
               "colors_ms": round(min(col), 3), "lighting_ms": round(min(lit), 3),
               "path_tracing_1spp_depth3_ms": round(min(pt), 3)}
    log(f"[bench] shading: {shading}")
    return shading


def shading_only(args):
    """The frame passes on a fresh depth-D terrain (for a kernel trace of them alone)."""
    import svdag_py
    import workloads

    D = args.edit_levels
    expected = int(64e6 * 4 ** (D - 15)) if D >= 12 else 0
    g = svdag_py.MyGPUHashDAG(D, 4, expected_items=expected)
    g.build_terrain(**workloads.terrain_params(D))
    c = workloads.camera_path(D, 64)[0]
    out = shading_bench(g, (c[0], svdag_py.look_at(c[0], c[1]), c[2], c[3], c[4]))
    g.close()
    return out


def editing_bench(args):
    """Depth-D editing on the fBm terrain (base 2^(D-1), amplitude 6000 * 2^D / 32768, 8 octaves).

    Every leg starts from a freshly built terrain (0.2 s at depth 15).
    plain:    the 256-edit batched-editing stream (40% sphere-add, 40%
              sphere-remove, 20% box fill, r log-uniform in [8, 512], centres
              within +-1024 of the surface), one svdag_edit call per edit (the
              reference's per-edit form); per-edit wall latency.
    batched:  the same stream as 16 fused batches of 16, followed by one
              garbage collection, all inside the timed region; an edit's
              latency is its batch's wall time.  `gc_each_batch` repeats it
              with a collection after every batch.
    serving:  open-loop Poisson arrivals at each rate (the same mix plus a 15%
              sphere paint) for `serve_seconds`, with a 1920x1080 primary trace
              every 16.7 ms; p50/p99 of submit -> root visible.  The
              saturation row submits its requests all at once (offered load far
              above capacity) and measures the sustained service rate.
    voxel edits/s counts, per edit, the voxels whose unit box meets the tool's
    overlap predicate (exact per-row interval counting)."""
    import svdag_py
    import workloads

    D = args.edit_levels
    tp = workloads.terrain_params(D)
    expected = int(64e6 * 4 ** (D - 15)) if D >= 12 else 0

    def fresh():
        t0 = time.perf_counter()
        g = svdag_py.MyGPUHashDAG(D, 4, expected_items=expected,
                                  workspace_items=min(1 << 28, int(2 * 4 ** max(D - 10, 0) * 400000)))
        g.build_terrain(**tp)
        st = g.garbage_collect()
        return g, time.perf_counter() - t0, st

    g, build_s, gc0 = fresh()
    log(f"[bench] depth {D} terrain: {build_s:.2f} s, {gc0.items_live} live items")
    terrain_dump = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        terrain_hash, _, terrain_dump = g.canonical(dump=True)
        log(f"[bench] terrain canonical dump: {terrain_dump.size} words, {time.perf_counter() - t:.1f} s")

    def h(x, z):
        return int(g.terrain_heights(x, z, 1, 1)[0, 0])

    edits = workloads.config3_edits(h, D, n=256, seed=6)
    counts = np.array([workloads.voxel_edit_count(k, c, r, D) for k, m, c, r in edits], np.float64)
    creq = {}
    cams = []
    for c in workloads.camera_path(D, 64):
        cams.append((c[0], svdag_py.look_at(c[0], c[1]), c[2], c[3], c[4]))
    for rate in args.serve_rates:
        creq[rate] = workloads.config4_requests(h, D, rate, args.serve_seconds, seed=7)
    sat_reqs, _ = workloads.config4_requests(h, D, 2000.0, 1.0, seed=9)
    # Frame passes on the fresh terrain at 1920x1080: primary paths, then
    # shadows and ambient occlusion (3 samples, rays of 2^(D-10) voxels), the
    # reference's defaults (engine.h:202-213).
    shading = None if args.no_shading else shading_bench(g, cams[0])
    g.close()

    # Every leg applies the same 256-edit stream to the same terrain, so every
    # leg must end on the same canonical DAG (hash over the renumbered dump);
    # any difference fails the run.
    hashes = {}

    def record_hash(leg, g):
        t = time.perf_counter()
        hashes[leg] = g.canonical()[0]
        log(f"[bench] canonical after {leg}: {hashes[leg]:#018x} ({time.perf_counter() - t:.1f} s)")

    # plain per-edit form
    g, _, _ = fresh()
    lat = []
    t1 = time.perf_counter()
    for e in edits:
        s0 = time.perf_counter()
        g.edit(*e)
        lat.append(time.perf_counter() - s0)
    plain_s = time.perf_counter() - t1
    plain = {"edits": len(edits), "voxel_edits_per_s": float(counts.sum() / plain_s),
             "edits_per_s": len(edits) / plain_s,
             "latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
             "latency_ms_p99": float(np.percentile(lat, 99) * 1e3)}
    log(f"[bench] plain: {plain}")
    record_hash("plain", g)
    g.close()

    def batched_leg(gc_each, mode, name, stream):
        """stream=True: the 256 edits in one svdag_edit_stream call (16-edit
        batches enqueued back to back, each followed by its collection when
        gc_each, one wait at the end).  stream=False: one svdag_edit_batch
        call per batch, and svdag_gc after it when gc_each, each waited for."""
        g, _, _ = fresh()
        g.set_batch_mode(mode)
        tune = g.autotune_all(edits[:16])
        if args.tune_load_factor:
            t_lf = time.perf_counter()
            tune.update(g.autotune_load_factor(edits[:16], (50, 75, 95)))
            tune["load_factor_tuning_s"] = round(time.perf_counter() - t_lf, 1)
            log(f"[bench] load factor: {tune['load_factor']} {tune['ms_by_load_factor']}")
        g.garbage_collect()
        blat = []
        gpu_ms = 0.0
        gc_ms = 0.0
        t2 = time.perf_counter()
        if stream:
            _, gpu_ms = g.edit_stream(edits, batch=16, collect=gc_each)
        else:
            for b in range(0, len(edits), 16):
                s0 = time.perf_counter()
                _, ms = g.edit_batch(edits[b:b + 16])
                gpu_ms += ms
                if gc_each:
                    gc_ms += g.garbage_collect().gpu_ms
                blat += [time.perf_counter() - s0] * len(edits[b:b + 16])
        if not gc_each:
            gc_ms += g.garbage_collect().gpu_ms
        dt = time.perf_counter() - t2
        record_hash(name, g)
        g.close()
        out = {"form": "one edit stream" if stream else "a call per batch",
               "voxel_edits_per_s": float(counts.sum() / dt), "edits_per_s": len(edits) / dt}
        if stream:
            out["gpu_ms_total"] = round(gpu_ms + gc_ms, 3)
        else:
            out.update({"edit_gpu_ms_total": round(gpu_ms, 3), "gc_gpu_ms_total": round(gc_ms, 3),
                        "latency_ms_p50": float(np.percentile(blat, 50) * 1e3),
                        "latency_ms_p99": float(np.percentile(blat, 99) * 1e3)})
        return out, tune

    batched, tune = batched_leg(False, 1, "fused stream, one collection", True)
    batched.update({"edits": len(edits), "batch": 16, "gc": "once after the stream",
                    "autotune": {"lanes_per_item": tune["lanes_per_item"], "max_blocks": tune["max_blocks"],
                                 "ms_by_lanes": {str(k): v for k, v in tune["ms_by_lanes"].items()},
                                 "ms_by_max_blocks": {str(k): v for k, v in tune["ms_by_max_blocks"].items()},
                                 **({"load_factor": tune["load_factor"],
                                     "ms_by_load_factor": {str(k): v for k, v in tune["ms_by_load_factor"].items()},
                                     "load_factor_tuning_s": tune["load_factor_tuning_s"]}
                                    if "load_factor" in tune else {})}})
    batched["speedup_vs_plain"] = batched["voxel_edits_per_s"] / plain["voxel_edits_per_s"]
    calls, _ = batched_leg(False, 1, "fused, a call per batch, one collection", False)
    calls["speedup_vs_plain"] = calls["voxel_edits_per_s"] / plain["voxel_edits_per_s"]
    batched["per_batch_calls"] = calls
    each, _ = batched_leg(True, 1, "fused stream, collection after each batch", True)
    each["speedup_vs_plain"] = each["voxel_edits_per_s"] / plain["voxel_edits_per_s"]
    batched["gc_each_batch"] = each
    waited, _ = batched_leg(True, 1, "fused, a call per batch, collection after each, each waited for", False)
    waited["speedup_vs_plain"] = waited["voxel_edits_per_s"] / plain["voxel_edits_per_s"]
    batched["gc_each_batch_per_batch_calls"] = waited
    graph, _ = batched_leg(True, 2, "graph stream, collection after each batch", True)
    graph["speedup_vs_plain"] = graph["voxel_edits_per_s"] / plain["voxel_edits_per_s"]
    batched["graph_gc_each_batch"] = graph
    log(f"[bench] batched: {batched}")

    # The same stream through the serving loop, all requests at t = 0.
    g, _, _ = fresh()
    g.autotune_all(edits[:16])
    g.garbage_collect()
    info, _ = g.serve_replay(edits, np.zeros(len(edits)), max_batch=64, allow_gc=True)
    if not (info["status"] == 0).all():
        raise RuntimeError("serving replay of the batched-editing stream: a request failed")
    record_hash("serving replay", g)
    g.close()
    if len(set(hashes.values())) != 1:
        raise RuntimeError(f"legs disagree on the final DAG: { {k: hex(v) for k, v in hashes.items()} }")

    serving = []
    g, _, _ = fresh()
    g.autotune_all(edits[:16])

    def serve_row(name, reqs, arr, max_batch):
        g = ctx[0]
        vox = np.array([workloads.voxel_edit_count(k, c, r, D) for k, m, c, r in reqs], np.float64)
        info, st = g.serve_replay(reqs, arr, max_batch=max_batch, frame=(1920, 1080), cameras=cams, allow_gc=True)
        latency = info["done_s"] - info["submit_s"]
        span = float(info["done_s"].max() - info["submit_s"].min())
        row = {"rate_per_s": name, "requests": len(reqs), "ok": int((info["status"] == 0).sum()),
               "latency_ms_p50": float(np.percentile(latency, 50) * 1e3),
               "latency_ms_p99": float(np.percentile(latency, 99) * 1e3),
               "throughput_req_per_s": len(reqs) / span, "voxel_edits_per_s": float(vox.sum() / span),
               "batches": st.batches, "max_batch": st.max_batch_seen, "frames": st.frames,
               "frames_skipped": st.frames_skipped, "trace_ms_avg_1920x1080": round(st.trace_ms_avg, 3),
               "gcs": st.gcs, "grows": st.grows, "sequential_spine_edits": st.sequential_spine_edits,
               "handed_back_edits": st.handed_back_edits}
        log(f"[bench] serving: {row}")
        serving.append(row)
        g.garbage_collect()
        return row

    ctx = [g]
    for rate in args.serve_rates:
        reqs, arr = creq[rate]
        serve_row(rate, reqs, arr, 64)
    # The same saturation requests on the context the rate rows have aged
    # (a grown DAG, fuller tables: collections and arena growth fall inside
    # the window), then on a fresh terrain, the state the plain form over the
    # same requests starts from below.
    aged = serve_row("saturation after the rate rows", sat_reqs, np.zeros(len(sat_reqs)), 64)
    g.close()
    g, _, _ = fresh()
    g.autotune_all(edits[:16])
    g.garbage_collect()
    ctx[0] = g
    sat = serve_row("saturation", sat_reqs, np.zeros(len(sat_reqs)), 64)
    g.close()
    sweep = []
    if args.serve_sweep:
        # Sensitivity of the saturation row to the server's sizing: tables
        # sized for twice the items, and larger batches.
        serving_rows = serving
        serving = []
        for scale, mb in ((2.0, 64), (2.0, 128), (2.0, 256)):
            g = svdag_py.MyGPUHashDAG(D, 4, expected_items=int(expected * scale),
                                      workspace_items=min(1 << 28, int(2 * 4 ** max(D - 10, 0) * 400000)))
            g.build_terrain(**tp)
            g.autotune_all(edits[:16])
            g.garbage_collect()
            ctx[0] = g
            row = serve_row(f"saturation, tables x{scale:g}, batches of {mb}", sat_reqs, np.zeros(len(sat_reqs)), mb)
            g.close()
            sweep.append(row)
        serving = serving_rows
    # The plain per-edit form over the saturation row's own requests (on a
    # fresh terrain): voxel edits/s depends on the request mix (a request's
    # voxels grow with r^3, and r is log-uniform), so the serving speedup is
    # also given on one request set.
    g, _, _ = fresh()
    sat_vox = float(sum(workloads.voxel_edit_count(k, c, r, D) for k, m, c, r in sat_reqs))
    t1 = time.perf_counter()
    for e in sat_reqs:
        g.edit(*e)
    same_s = time.perf_counter() - t1
    g.close()
    plain_same = {"requests": len(sat_reqs), "voxel_edits_per_s": sat_vox / same_s,
                  "edits_per_s": len(sat_reqs) / same_s}
    log(f"[bench] plain over the saturation requests: {plain_same}")
    cpu = cpu_edit_baseline(args, terrain_dump, edits, D) if not args.no_cpu_baseline else None
    return {"levels": D, "terrain_build_s": round(build_s, 2), "live_items": gc0.items_live,
            "cpu_baseline": cpu,
            "canonical_hash": f"{hashes['plain']:#018x}", "legs_with_that_hash": list(hashes),
            "plain": plain, "batched": batched, "serving": serving, "shading": shading,
            "serving_saturation_speedup_vs_plain": sat["voxel_edits_per_s"] / plain["voxel_edits_per_s"],
            "plain_over_saturation_requests": plain_same,
            "serving_saturation_speedup_vs_plain_same_requests": sat["voxel_edits_per_s"] / plain_same["voxel_edits_per_s"],
            "serving_aged_saturation_speedup_vs_plain_same_requests":
                aged["voxel_edits_per_s"] / plain_same["voxel_edits_per_s"],
            **({"serving_sweep": sweep} if sweep else {})}


def large_scene_bench(args):
    """Depth-17 scene with 8-bit materials: the fBm terrain of terrain_params(17)
    (base 65536, amplitude 24000, 8 octaves) with height-band materials (bands
    of 512 voxels plus a seeded 0..3 offset per 512-column tile), built one
    16384-wide column tile per edit.  Then `large_copies` copy/paste edits, each
    a 512^3 region (centre = source = a seeded point near the surface, radius
    256) pasted at a seeded point near the surface elsewhere with a 90 degree
    turn about +Y, in groups of 8; after each group the last 8 edits are undone
    and redone.  A collection ends the stream.  Tables start sized for
    `large_expected` items and grow as the allocators run short.

    Checks (outside the timed edits): terrain columns against the height field
    and the band rule; for every copy, seeded destination voxels against the
    copy rule (a set source voxel replaces the destination's value, an empty
    one keeps it); after the undos, the first copy's destination voxels hold
    their values from before the group; after the redos, the last copy's hold
    their values from after it.  Any mismatch fails the leg."""
    import svdag_py
    import workloads

    D, M = args.large_levels, 8
    side = 1 << D
    band_log2, band_seed = 9, 0xBA5E
    tp = dict(workloads.terrain_params(D))
    t0 = time.perf_counter()
    g = svdag_py.MyGPUHashDAG(D, M, expected_items=int(args.large_expected), workspace_items=1 << 26)
    g.build_terrain(**tp, band_log2=band_log2, band_seed=band_seed, tile_log2=D - 3)
    build_s = time.perf_counter() - t0
    gc0 = g.garbage_collect()
    log(f"[bench] depth {D} banded terrain: {build_s:.2f} s, {gc0.items_live} live items")
    rng = np.random.default_rng(10)

    # Terrain columns: the surface voxel and one deeper voxel hold the band
    # material, the voxel above the surface is empty.
    cols = rng.integers(0, side, size=(2000, 2))
    hts = np.array([int(g.terrain_heights(int(x), int(z), 1, 1)[0, 0]) for x, z in cols[:2000]], np.int64)
    xs, zs = cols[:2000, 0], cols[:2000, 1]
    ok = (hts >= 64) & (hts + 1 < side)
    xs, zs, hts = xs[ok], zs[ok], hts[ok]
    for dy, solid in ((0, True), (-63, True), (1, False)):
        ys = hts + dy
        got = g.query(np.stack([xs, ys, zs], axis=1))
        want = (workloads.band_materials(xs, ys, zs, tp["material"], band_log2, band_seed, M) if solid
                else np.full(len(xs), svdag_py.NOT_FOUND, np.uint32))
        if not np.array_equal(got, want):
            raise RuntimeError(f"terrain column check failed at dy={dy}: {int((got != want).sum())} mismatches")
    log(f"[bench] terrain columns checked: {3 * len(xs)} voxels")

    def near_surface():
        x, z = (int(v) for v in rng.integers(256, side - 256, 2))
        y = int(g.terrain_heights(x, z, 1, 1)[0, 0]) + int(rng.integers(-256, 257))
        return (float(x), float(min(max(y, 256), side - 256)), float(z))

    rotation = 1
    copies = [(near_surface(), near_surface()) for _ in range(args.large_copies)]
    lat, undo_ms, redo_ms = [], [], []
    roots = []
    checked = 0
    group_before = None
    after = []
    for i, (src, dst) in enumerate(copies):
        v = workloads.copy_box_voxels(dst, 256.0, 4096, rng, D)
        sv = workloads.copy_source_voxels(v, src, 256.0, src, dst, rotation)
        old = g.query(v)
        srcv = g.query(sv)
        if i % 8 == 0:
            group_before = (v, old)
        s0 = time.perf_counter()
        g.edit(svdag_py.COPY, 0, src, 256.0, src, dst, rotation=rotation)
        lat.append(time.perf_counter() - s0)
        new = g.query(v)
        want = np.where(srcv != svdag_py.NOT_FOUND, srcv, old)
        if not np.array_equal(new, want):
            raise RuntimeError(f"copy {i}: {int((new != want).sum())} of {len(v)} sampled voxels break the copy rule")
        checked += len(v)
        after = (v, new)
        roots.append(g.root)
        if i % 8 == 7:
            for _ in range(8):
                s0 = time.perf_counter()
                g.undo()
                undo_ms.append((time.perf_counter() - s0) * 1e3)
            if not np.array_equal(g.query(group_before[0]), group_before[1]):
                raise RuntimeError("undo of 8 copies did not restore the sampled voxels")
            for _ in range(8):
                s0 = time.perf_counter()
                g.redo()
                redo_ms.append((time.perf_counter() - s0) * 1e3)
            if g.root != roots[-1] or not np.array_equal(g.query(after[0]), after[1]):
                raise RuntimeError("redo did not restore the root and the sampled voxels")
    gc = g.garbage_collect()
    edit_s = sum(lat) + (sum(undo_ms) + sum(redo_ms)) * 1e-3 + gc.gpu_ms * 1e-3
    mem = g.memory_usage()
    g.close()
    vox = float(len(copies)) * 512.0 ** 3
    out = {"levels": D, "material_bits": M, "terrain_build_s": round(build_s, 2), "live_items": gc0.items_live,
           "copies": len(copies), "region": "512^3", "rotation_quarter_turns": rotation,
           "copies_per_s": len(copies) / edit_s, "voxel_edits_per_s": vox / edit_s,
           "copy_latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
           "copy_latency_ms_p99": float(np.percentile(lat, 99) * 1e3),
           "undo_ms_p99": float(np.percentile(undo_ms, 99)) if undo_ms else None,
           "redo_ms_p99": float(np.percentile(redo_ms, 99)) if redo_ms else None,
           "gc_gpu_ms": round(gc.gpu_ms, 3), "live_items_after": gc.items_live,
           "voxels_checked": int(checked + 3 * len(xs)),
           "table_bytes_reserved": int(mem.bytes_reserved), "slabs_used": int(mem.slabs_used),
           "slab_capacity": int(mem.slab_capacity)}
    log(f"[bench] large scene: {out}")
    return out


def reference_format_bench(args):
    """The reference's version-3 file format at the editing workload's depth
    (terrain_params(D), 4 material bits): svdag_write_reference_dag of the
    whole DAG, then svdag_read_reference_dag of those bytes into a fresh
    context.  Host-side work on the host's cores (the writer rebuilds the
    reference's 32-slot tables, the reader walks them and imports through the
    bulk insert).  Checks: both contexts give one canonical hash."""
    import svdag_py
    import workloads

    D = args.edit_levels
    expected = int(64e6 * 4 ** (D - 15)) if D >= 12 else 0
    g = svdag_py.MyGPUHashDAG(D, 4, expected_items=expected)
    g.build_terrain(**workloads.terrain_params(D))
    live = g.garbage_collect().items_live
    t0 = time.perf_counter()
    data = g.write_reference_dag()
    write_s = time.perf_counter() - t0
    log(f"[bench] reference stream written: {len(data) / 1e6:.1f} MB in {write_s:.2f} s")
    g2 = svdag_py.MyGPUHashDAG(D, 4, expected_items=expected)
    t0 = time.perf_counter()
    used = g2.read_reference_dag(data)
    read_s = time.perf_counter() - t0
    log(f"[bench] reference stream read: {read_s:.2f} s")
    h1, h2 = g.canonical()[0], g2.canonical()[0]
    if used != len(data) or h1 != h2:
        raise RuntimeError(f"reference format round trip differs: {h1:#x} vs {h2:#x}, {used} of {len(data)} bytes")
    g2.close()
    g.close()
    return {"levels": D, "material_bits": 4, "live_items": live, "bytes": len(data),
            "write_s": round(write_s, 3), "read_s": round(read_s, 3),
            "write_items_per_s": live / write_s, "read_items_per_s": live / read_s,
            "canonical_hash": f"{h1:#018x}", "host": host_cpu_info()}


def editing_profile(args):
    """The batched-editing stream alone (depth 15 terrain, 256 edits in
    16-edit fused batches, one collection after the stream), for a kernel
    trace of the edit waves and the collection: rocprofv3 --kernel-trace
    --stats -- python3 bench.py --editing-profile."""
    import svdag_py
    import workloads

    D = args.edit_levels
    tp = workloads.terrain_params(D)
    g = svdag_py.MyGPUHashDAG(D, 4, expected_items=int(64e6 * 4 ** (D - 15)) if D >= 12 else 0,
                              workspace_items=min(1 << 28, int(2 * 4 ** max(D - 10, 0) * 400000)))
    g.build_terrain(**tp)
    g.garbage_collect()

    def h(x, z):
        return int(g.terrain_heights(x, z, 1, 1)[0, 0])

    edits = workloads.config3_edits(h, D, n=256, seed=6)
    if args.marked_batch >= 0:
        # A device-to-device copy of known size: counter runs scale FETCH_SIZE
        # and WRITE_SIZE by what the counters read for it.
        nbytes = 256 << 20
        a = svdag_py.DeviceBuffer(nbytes)
        b = svdag_py.DeviceBuffer(nbytes)
        svdag_py.check(svdag_py.lib().svdag_memcpy_d2d(b.ptr, a.ptr, svdag_py.ctypes.c_size_t(nbytes)))
        del a, b
    g.set_batch_mode(args.batch_mode)
    tune_g = g.autotune_all(edits[:16])["lanes_per_item"]
    g.garbage_collect()
    rows = []
    marked = args.marked_batch
    if args.profile_stream:
        # The whole stream in one svdag_edit_stream call, a collection after
        # every batch (each beside the next batch's SVO build).
        _, ms = g.edit_stream(edits, batch=16, collect=True)
        log(f"[profile] edit stream with a collection after each batch: {ms:.3f} ms")
        g.close()
        print(json.dumps({"edit_stream_gpu_ms": round(ms, 3), "lanes_per_item": tune_g}), flush=True)
        return
    for b in range(0, len(edits), 16):
        if b // 16 == marked:
            # Separator launches around the batch (a 1 MiB stream read, kernel
            # probe_stream), so that per-dispatch counter rows of this one batch
            # can be told apart from the build's and the other batches'.
            svdag_py.hbm_probe(svdag_py.PROBE_STREAM, 1 << 20)
        st, ms = g.edit_batch(edits[b:b + 16])
        if b // 16 == marked:
            svdag_py.hbm_probe(svdag_py.PROBE_STREAM, 1 << 20)
        if args.profile_gc_each and b + 16 < len(edits):
            g.garbage_collect(wait=False)
        items = [sum(x.svo_items[l] for x in st) for l in range(D - 1)]
        log(f"[profile] batch {b // 16}: {ms:.3f} ms, SVO items per level {items}")
        rows.append(ms)
    gc = g.garbage_collect()
    log(f"[profile] collection: {gc.gpu_ms:.3f} ms, {gc.items_live} live, {gc.items_freed} freed")
    g.close()
    print(json.dumps({"edit_batches_ms": [round(x, 3) for x in rows], "edit_gpu_ms_total": round(sum(rows), 3),
                      "gc_ms": round(gc.gpu_ms, 3), "lanes_per_item": tune_g}), flush=True)


def host_cpu_info():
    """CPU model, online CPUs (nproc) and the CPUs this process may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_allowed": allowed}


def cpu_edit_baseline(args, dump15, edits15, D):
    """The reference's CPU HashDAG editing path (geometry only) restated in
    C++ (oracle/cpu_hashdag.cpp), on the host cores, with NUM_THREADS = 32
    worker threads (the reference default) and with 1.
    config 1: the depth-10 terrain (imported from the GPU build's canonical
      dump), sphere-add r=64 then sphere-remove r=48; median over 5 runs
      after one warm-up, each on a fresh import.
    stream:   the depth-15 terrain and the 256-edit batched-editing stream
      applied one edit at a time, until `cpu_seconds` of edit time have
      passed (a bounded sample).
    voxel edits/s counts each edit's voxels by the CPU tool's own predicate
    (its sphere has no +0.5 centre); paint / copy do not occur in the stream."""
    import oracle_py
    import svdag_py
    import workloads

    kinds = {svdag_py.SPHERE_ADD: 0, svdag_py.SPHERE_REMOVE: 1, svdag_py.BOX_ADD: 2, svdag_py.BOX_REMOVE: 3}

    def cpu_count(kind, c, r, levels):
        if kind in (svdag_py.SPHERE_ADD, svdag_py.SPHERE_REMOVE):
            return workloads.voxel_edit_count(kind, tuple(x - 0.5 for x in c), r, levels)
        return workloads.voxel_edit_count(kind, c, r, levels)

    def make(levels, dump):
        top = 0
        for _ in range(4):
            d = oracle_py.CpuHashDag(levels, top_span=top)
            try:
                t = time.perf_counter()
                d.import_canonical(dump, threads=16)
                return d, top or 1024, time.perf_counter() - t
            except RuntimeError:
                top = 2 * (top or 1024)
                del d
        raise RuntimeError("CPU HashDAG: the terrain overflows the bucket spans")

    info = host_cpu_info()
    out = {"host": info, "kind": "port"}
    # config 1 at depth 10
    g = svdag_py.MyGPUHashDAG(10, 4)
    g.build_terrain(**workloads.terrain_params(10))
    _, _, dump10 = g.canonical(dump=True)
    g.close()
    c1 = workloads.config1_edits()
    vox1 = sum(cpu_count(k, c, r, 10) for k, m, c, r in c1)
    for threads in (32, 1):
        times = []
        for rep in range(6):
            d, _, _ = make(10, dump10)
            ms = sum(d.edit(kinds[k], c, r, threads)[0] for k, m, c, r in c1)
            del d
            if rep:
                times.append(ms)
        med = float(np.median(times))
        out[f"config1_threads{threads}"] = {"ms_for_both_edits_median": round(med, 3),
                                             "voxel_edits_per_s": vox1 / (med * 1e-3)}
    # the depth-D stream
    if dump15 is not None:
        for threads, budget in ((32, args.cpu_seconds), (1, args.cpu_seconds / 2)):
            d, top, imp_s = make(D, dump15)
            ms_total, vox, n = 0.0, 0, 0
            for k, m, c, r in edits15:
                ms, _ = d.edit(kinds[k], c, r, threads)
                ms_total += ms
                vox += cpu_count(k, c, r, D)
                n += 1
                if ms_total * 1e-3 > budget:
                    break
            del d
            out[f"stream_threads{threads}"] = {"edits": n, "edit_s": round(ms_total * 1e-3, 3),
                                                "voxel_edits_per_s": vox / (ms_total * 1e-3),
                                                "edits_per_s": n / (ms_total * 1e-3),
                                                "import_s": round(imp_s, 2), "top_level_bucket_words": top}
            log(f"[bench] cpu HashDAG stream, {threads} threads: {out[f'stream_threads{threads}']}")
    s32 = out.get("stream_threads32")
    out.update({"value": round((s32 or out["config1_threads32"])["voxel_edits_per_s"] / 1e9, 4),
                "unit": "G voxel edits/s", "cores": 32,
                "sample": (f"{s32['edits']} edits of the depth-{D} batched-editing stream, one at a time, "
                           f"32 threads" if s32 else "config 1 at depth 10, 32 threads")
                + " (reference CPU HashDAG edit path, C++ restatement)"})
    log(f"[bench] cpu HashDAG baseline: {out}")
    return out


def cpu_baseline(items, finds, buckets, args):
    import oracle_py

    threads = args.cpu_threads
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
        "host": host_cpu_info(),
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
    ap.add_argument("--edit-levels", type=int, default=15, help="depth of the editing / serving workloads")
    ap.add_argument("--plain-edits", type=int, default=64)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[50.0, 200.0, 1000.0, 2000.0])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--no-editing", action="store_true", help="slab-hash line only")
    ap.add_argument("--no-probe", action="store_true", help="skip the stream / random-line memory probes")
    ap.add_argument("--no-as-written", action="store_true", help="skip the 96-items-per-bucket find point")
    ap.add_argument("--fill", type=float, default=None,
                    help="head-line fill the find tables are sized for (default: picked by timing 0.5 .. 0.9)")
    ap.add_argument("--no-shading", action="store_true", help="skip the shadow / occlusion timing")
    ap.add_argument("--large-scene", action="store_true", help="add the depth-17 8-bit-material copy/paste leg")
    ap.add_argument("--large-scene-only", action="store_true", help="only the depth-17 copy/paste leg")
    ap.add_argument("--large-copies", type=int, default=32)
    ap.add_argument("--large-levels", type=int, default=17)
    ap.add_argument("--large-expected", type=float, default=4e8, help="items the depth-17 tables start sized for")
    ap.add_argument("--shading-only", action="store_true",
                    help="only the 1920x1080 frame passes on a fresh terrain at --edit-levels")
    ap.add_argument("--serve-sweep", action="store_true",
                    help="also serve the saturation requests with tables sized x2 and batches of 64 / 128 / 256")
    ap.add_argument("--tune-load-factor", action="store_true",
                    help="editing legs also tune the DAG tables' load factor (50/75/95%%; rebuilds the tables)")
    ap.add_argument("--ref-format-only", action="store_true",
                    help="only the reference file-format write / read round trip at --edit-levels")
    ap.add_argument("--editing-profile", action="store_true",
                    help="only the batched-editing stream and its collection (for rocprofv3 --kernel-trace)")
    ap.add_argument("--batch-mode", type=int, default=1, help="editing profile: 1 fused, 2 fused + HIP graphs")
    ap.add_argument("--profile-gc-each", action="store_true",
                    help="editing profile: a collection (not waited for) after every batch")
    ap.add_argument("--profile-stream", action="store_true",
                    help="editing profile: the stream as one svdag_edit_stream call, a collection after each batch")
    ap.add_argument("--marked-batch", type=int, default=-1,
                    help="editing profile: bracket this batch with probe_stream launches (counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this back-end has no multi-GPU path "
              "(the DAG would shard by space across GPUs; not built)", file=sys.stderr)
        sys.exit(2)
    if args.editing_profile:
        editing_profile(args)
        return
    if args.large_scene_only:
        print(json.dumps({"large_scene": large_scene_bench(args)}), flush=True)
        return
    if args.shading_only:
        print(json.dumps({"shading": shading_only(args)}), flush=True)
        return
    if args.ref_format_only:
        print(json.dumps({"reference_format": reference_format_bench(args)}), flush=True)
        return
    out = slab_hash_bench(args)
    if out is None:
        return
    if not args.no_editing:
        out["editing"] = editing_bench(args)
    if args.large_scene:
        out["large_scene"] = large_scene_bench(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
