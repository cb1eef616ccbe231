#!/usr/bin/env python3
"""Benchmark of scan integration (ProbabilisticMap::insertPointCloud) on one
MI355X: the batched form (one captured graph per frame of 64 scans into 64
maps) against the plain per-scan form (one insert call per scan), with the
CPU restatement timed beside it.

Workload (one "step" = one frame): 64 independent maps at 0.1 m voxels, each
receiving one spinning-LiDAR scan of a seeded box scene, point counts
log-uniform in 10k..500k (about 6.9M points a frame), max range 30 m.  Every
timed frame goes into 64 EMPTY maps (the workload as stated: every cell it
touches is created, as in the reference's first insert into a new map); the
map sets are emptied outside the timed events.  Inputs are resident in HBM
before the timed region.  The same frame re-inserted into maps that already
hold it is reported beside it as "steady".

Prints one JSON line.  Extra fields: the plain per-scan rate, the
batched/plain ratio, voxel updates/s, per-frame p50/p99 latency, the
roofline of the dominant scan kernel (the fresh pass, k_scan_fresh_rays, on
empty maps; k_scan_rays on steady frames), of the leaf assembly, and of the
bulk lookup and insert
kernels (VoxelGrid<float>, 2M inserts / 8M lookups), the serving form (16
resident maps, a Poisson stream of depth-frame inserts, lookups and exports;
p50/p99 per request type), and the CPU baselines.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (profiles/traffic.json, scaled by the copy-probe factors), or
    None when no profile has been recorded."""
    try:
        with open(os.path.join(HERE, "profiles", "traffic.json")) as f:
            return json.load(f)[kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None
METRIC = "rays/s + voxel updates/s, batched vs plain per-scan on 1 MI355X; p50/p99 latency"


LOOKUP_BYTES = 28
INSERT_BYTES = 28
GATHER_TABLE = 64 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frame_stats(maps):
    st = [m.stats() for m in maps]
    return {"rays": sum(x["last_rays"] for x in st), "ray_updates": sum(x["last_ray_updates"] for x in st),
            "ray_visits": sum(x["last_ray_visits"] for x in st), "leaves": sum(x["leaf_count"] for x in st)}


def reset_maps(ctx, bx, maps):
    """Empty the maps as new ones are (no roots, no leaves, update_count 1);
    their pools keep their capacity, so a frame into them allocates no device
    memory."""
    for m in maps:
        m.clear(bx.CLEAR_MEMORY)
        m.update_count = 1
    ctx.sync()


def make_map_sets(ctx, bx, scans, dev, max_range, nsets, budget_bytes):
    """Up to nsets sets of one map per scan, as many as fit in budget_bytes.
    Each set is sized by the library's own head-room policy: the frame is
    integrated into it once eagerly and once through a captured graph (as a
    caller's first frames would be), and it is emptied after each, so that
    its pools hold what a fresh frame needs and no timed frame allocates
    device memory or overflows a pool.  Returns the sets, the leaves of one
    fresh frame and the device bytes of one set."""
    descs = lambda maps: [(m, d, o, max_range) for m, d, (_, o) in zip(maps, dev, scans)]
    sets = []
    per_set = 0
    leaves = 0
    g = bx.ScanGraph(ctx, len(scans), sum(len(p) for p, _ in scans))
    while len(sets) < nsets and (not sets or (len(sets) + 1) * per_set <= budget_bytes):
        maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
        bx.insert_batch(ctx, descs(maps))
        leaves = sum(m.stats()["leaf_count"] for m in maps)
        reset_maps(ctx, bx, maps)
        g.set_scans(descs(maps))
        g.launch()
        check_maps(maps, "sizing frames")
        reset_maps(ctx, bx, maps)
        per_set = max(per_set, sum(m.stats()["device_bytes"] for m in maps))
        sets.append(maps)
    g.close()
    ctx.sync()
    return sets, leaves, per_set


def replays(sets):
    return sum(m.stats()["replays"] for maps in sets for m in maps)


def check_maps(maps, what):
    for m in maps:
        if m.sync() != 0:
            raise RuntimeError(f"a map failed during the {what}")


def timed_maps_match(sets):
    """Whether every map of every set, as the timed frames left it, equals the
    oracle's map of its scan: the order-free digests of the cells against
    those committed in tests/golden/digests.json for the default frame (64
    maps, seed 11).  None for another frame."""
    try:
        with open(os.path.join(HERE, "tests", "golden", "digests.json")) as f:
            want = [w["digest"] for w in json.load(f)["maps"]["batch"]]
    except (OSError, KeyError, ValueError):
        return None
    if len(want) != len(sets[0]):
        return None
    return all([str(x) for x in m.digest()] == want[i] for maps in sets for i, m in enumerate(maps))


def run_batched(ctx, bx, scans, dev, steps, warmup, max_range, sets, steady=True):
    """The batched form, one captured graph launch per frame of 64 scans.
    Fresh frames (the workload as stated, and `value`): each timed step
    integrates the frame into a set of 64 empty maps; up to len(sets) steps run
    back to back between two events, the sets being emptied beforehand outside
    the events.  Steady frames: the same frame re-inserted back to back into
    maps that already hold it."""
    descs = [[(m, d, o, max_range) for m, d, (_, o) in zip(maps, dev, scans)] for maps in sets]
    frames = [bx.ScanGraph.prepare(d) for d in descs]
    total_points = sum(len(p) for p, _ in scans)
    nsets = len(sets)
    replays0 = replays(sets)
    # counted frames: exact rays, cell updates and visits, fresh and steady
    reset_maps(ctx, bx, sets[0])
    ctx.counting(True)
    bx.insert_batch(ctx, descs[0])
    fresh_counts = frame_stats(sets[0])
    steady_counts = None
    if steady:
        bx.insert_batch(ctx, descs[0])
        steady_counts = frame_stats(sets[0])
    ctx.counting(False)
    g = bx.ScanGraph(ctx, len(scans), total_points)
    # fresh frames, warm-up then timed
    for ms in sets:
        reset_maps(ctx, bx, ms)
    for k in range(min(warmup, nsets)):
        g.set_scans(frames[k])
        g.launch()
    ctx.sync()
    mem0 = ctx.memory()[0]
    fresh_ms = 0.0
    done = rounds = 0
    while done < steps:
        n = min(nsets, steps - done)
        for ms in sets[:n]:
            reset_maps(ctx, bx, ms)
        ctx.event(0)
        for k in range(n):
            g.set_scans(frames[k])
            g.launch()
        ctx.event(1)
        fresh_ms += ctx.elapsed_ms(0, 1)
        done += n
        rounds += 1
    for ms in sets:
        check_maps(ms, "timed fresh frames")
    fresh_replays = replays(sets) - replays0
    fresh_alloc = ctx.memory()[0] - mem0
    timed_match = timed_maps_match(sets)
    # per-frame latency of fresh frames, one frame between two events
    lat_fresh = []
    for k in range(max(steps, 10)):
        ms = sets[k % nsets]
        reset_maps(ctx, bx, ms)
        ctx.event(2)
        g.set_scans(frames[k % nsets])
        g.launch()
        ctx.event(3)
        lat_fresh.append(ctx.elapsed_ms(2, 3))
    check_maps(sets[0], "latency frames")
    # kernel durations, eager launches with per-kernel events
    names = SCAN_KERNELS
    nprof = 3
    ctx.profile_reset()
    for _ in range(nprof):
        reset_maps(ctx, bx, sets[0])
        ctx.profile(True)
        bx.insert_batch(ctx, descs[0])
        ctx.profile(False)
    kern_fresh = {k: ctx.profile_read(k)[1] / max(ctx.profile_read(k)[0], 1) for k in names}
    out = {
        "points": total_points, "fresh_ms": fresh_ms, "fresh_rounds": rounds, "lat_fresh": lat_fresh,
        "fresh_counts": fresh_counts, "kern_fresh": kern_fresh, "replays_timed": fresh_replays,
        "alloc_timed": fresh_alloc, "steady_ms": None, "timed_match": timed_match,
    }
    if not steady:
        g.close()
        out["replays_all"] = replays(sets) - replays0
        return out
    # steady frames into set 0, which holds the frame after the eager frames above
    g.set_scans(frames[0])
    for _ in range(warmup):
        g.launch()
    ctx.sync()
    ctx.event(0)
    for _ in range(steps):
        g.launch()
    ctx.event(1)
    steady_ms = ctx.elapsed_ms(0, 1)
    check_maps(sets[0], "steady frames")
    lat_steady = []
    for _ in range(max(steps, 10)):
        ctx.event(2)
        g.launch()
        ctx.event(3)
        lat_steady.append(ctx.elapsed_ms(2, 3))
    check_maps(sets[0], "steady latency frames")
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(nprof):
        bx.insert_batch(ctx, descs[0])
    ctx.profile(False)
    kern_steady = {k: ctx.profile_read(k)[1] / max(ctx.profile_read(k)[0], 1) for k in names}
    g.close()
    out.update({"steady_ms": steady_ms, "lat_steady": lat_steady, "steady_counts": steady_counts,
                "kern_steady": kern_steady, "replays_all": replays(sets) - replays0})
    return out


def run_plain(ctx, bx, scans, dev, steps, warmup, max_range, sets):
    """The plain per-scan form, one insert call per scan (64 a frame), into
    fresh maps (as run_batched) and re-inserted into maps holding the frame."""
    nsets = len(sets)
    for ms in sets:
        reset_maps(ctx, bx, ms)
    for k in range(min(warmup, nsets)):
        for m, d, (_, o) in zip(sets[k], dev, scans):
            m.insert_point_cloud(d, o, max_range, sync=False)
    ctx.sync()
    fresh_ms = 0.0
    done = 0
    while done < steps:
        n = min(nsets, steps - done)
        for ms in sets[:n]:
            reset_maps(ctx, bx, ms)
        ctx.event(0)
        for k in range(n):
            for m, d, (_, o) in zip(sets[k], dev, scans):
                m.insert_point_cloud(d, o, max_range, sync=False)
        ctx.event(1)
        fresh_ms += ctx.elapsed_ms(0, 1)
        done += n
    for ms in sets:
        check_maps(ms, "plain fresh frames")
    # per-scan latency, fresh maps (each insert between two events)
    lat = []
    reset_maps(ctx, bx, sets[0])
    for m, d, (_, o) in zip(sets[0][:16], dev[:16], scans[:16]):
        ctx.event(2)
        m.insert_point_cloud(d, o, max_range, sync=False)
        ctx.event(3)
        lat.append(ctx.elapsed_ms(2, 3))
    # steady: set 0 holds the frame (or part of it): fill it, then re-insert
    reset_maps(ctx, bx, sets[0])
    for m, d, (_, o) in zip(sets[0], dev, scans):
        m.insert_point_cloud(d, o, max_range, sync=False)
    for _ in range(warmup):
        for m, d, (_, o) in zip(sets[0], dev, scans):
            m.insert_point_cloud(d, o, max_range, sync=False)
    ctx.sync()
    ctx.event(0)
    for _ in range(steps):
        for m, d, (_, o) in zip(sets[0], dev, scans):
            m.insert_point_cloud(d, o, max_range, sync=False)
    ctx.event(1)
    steady_ms = ctx.elapsed_ms(0, 1)
    check_maps(sets[0], "plain steady frames")
    return {"fresh_ms": fresh_ms, "steady_ms": steady_ms, "lat": lat}


def run_lookups(ctx, bx, scenes, reps=20):
    coords = scenes.clusters()
    vals = scenes.cluster_values(coords)
    q = scenes.lookups(coords)
    g = bx.VoxelGrid(ctx, 0.05, dtype=np.float32)
    dc = ctx.to_device(coords)
    dv = ctx.to_device(vals)
    g.set_values(dc, dv)
    assert g.sync() == 0
    g.clear(bx.CLEAR_MEMORY)
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(3):
        g.clear(bx.CLEAR_MEMORY)
        g.set_values(dc, dv)
    ctx.profile(False)
    set_kern = {k: ctx.profile_read(k)[1] / max(ctx.profile_read(k)[0], 1)
                for k in ("k_op_clear", "k_op_begin", "k_op_resolve", "k_zero_new_leaves", "k_op_apply")}
    set_ms = sum(set_kern.values())
    dq = ctx.to_device(q)
    out_v = ctx.array(len(q), np.float32)
    out_f = ctx.array(len(q), np.uint8)
    g.lookup(dq, out_v, out_f)
    ctx.sync()
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(reps):
        g.lookup(dq, out_v, out_f)
    ctx.profile(False)
    n, ms = ctx.profile_read("k_lookup")
    avg_ms = ms / max(n, 1)
    found = out_f.numpy()
    # the memory system's rate for the hits' cell reads alone: as many scattered
    # 4-byte reads from a 64 MiB table (the map's cells take 67 MB)
    hits = int(found.sum())
    # (the probe library launches on the context's stream but outside its
    # per-kernel profiler: one warm-up launch, then reps launches back to back
    # between two events)
    table = ctx.array(GATHER_TABLE // 4, np.uint32)
    gout = ctx.array(max(hits, 1), np.uint32)

    def gather():
        rc = bx.probes().bxg_gather_probe(ctx._h, table.ptr, GATHER_TABLE, hits, gout.ptr)
        if rc != bx.OK:
            raise RuntimeError(f"gather probe failed: {rc}")

    gather()
    ctx.sync()
    ctx.event(4)
    for _ in range(reps):
        gather()
    ctx.event(5)
    ctx.sync()
    gather_ms = ctx.elapsed_ms(4, 5) / reps
    table.free()
    gout.free()
    for d in (dc, dv, dq, out_v, out_f):
        d.free()
    g.close()
    nq = len(q)
    cpu = cpu_grid_baseline(coords, vals, q, cpu_threads())
    return {
        "cpu_baseline": cpu,
        "lookups": nq, "lookup_ms": avg_ms, "lookups_per_s": nq / (avg_ms / 1e3),
        "hit_frac": float(found.mean()), "inserts": len(coords), "insert_ms": set_ms,
        "gather_ceiling": {"reads": hits, "table_bytes": GATHER_TABLE, "ms": gather_ms,
                           "reads_per_s": hits / (gather_ms / 1e3),
                           "over_lookup_ms": gather_ms / avg_ms},
        "inserts_per_s": len(coords) / (set_ms / 1e3), "insert_kernels_ms": set_kern,
    }


def run_box_scan(ctx, bx, scenes, reps=10):
    """Config 1, the reference's own CPU-sized case: one 100k-point box-room
    scan at 0.05 m, max range 10 m, into a fresh map (pools reserved
    beforehand, so the timed call allocates no device memory).  Returns the
    median per-scan time over reps fresh maps and the oracle's time for the
    same scan on one core."""
    import oracle
    pts, origin = scenes.box_scan()
    d = ctx.to_device(pts)
    ms = []
    for _ in range(reps + 1):
        m = bx.ProbabilisticMap(ctx, 0.05)
        m.reserve(160_000, 4096)
        ctx.sync()
        ctx.event(2)
        m.insert_point_cloud(d, origin, 10.0, sync=False)
        ctx.event(3)
        ms.append(ctx.elapsed_ms(2, 3))
        if m.sync() != 0:
            raise RuntimeError("box scan failed")
        m.close()
    d.free()
    t0 = time.perf_counter()
    om = oracle.ProbMap(0.05)
    om.insert_point_cloud(pts, origin, 10.0)
    cpu_s = time.perf_counter() - t0
    om.close()
    gpu_ms = float(np.median(ms[1:]))
    return {"points": len(pts), "ms": gpu_ms, "rays_per_s": len(pts) / (gpu_ms / 1e3),
            "cpu_ms": cpu_s * 1e3, "cpu_rays_per_s": len(pts) / cpu_s, "cpu_cores": 1}


SCAN_KERNELS = ("k_scan_begin", "k_scan_classify", "k_scan_winners", "k_scan_bucket_reduce",
                "k_scan_bucket_scan", "k_scan_scatter", "k_scan_ray_leaves", "k_scan_endpoints",
                "k_scan_fresh_rays", "k_scan_rays", "k_scan_new_endpoints", "k_scan_new_leaves",
                "k_scan_finalize")


def run_sequence(ctx, bx, scenes, nscans=200, cpu_budget_s=3.0):
    """Config 3: 200 LiDAR scans (64 x 1024 rays, 0.1 m, 50 m) into one map in
    order, one insert each (device points, pools reserved from an untimed
    pass).  The oracle maps the first scans on one core for about
    cpu_budget_s."""
    import oracle
    scans = list(scenes.lidar_sequence(nscans))
    dev = [ctx.to_device(p) for p, _ in scans]
    m = bx.ProbabilisticMap(ctx, 0.1)
    for d, (_, o) in zip(dev, scans):
        m.insert_point_cloud(d, o, 50.0, sync=False)
    if m.sync() != 0:
        raise RuntimeError("sequence sizing pass failed")
    st = m.stats()
    m.close()
    m = bx.ProbabilisticMap(ctx, 0.1)
    m.reserve(int(st["leaf_count"] * 1.25) + 8192, int(st["root_count"] * 2) + 1024)
    ctx.sync()
    ctx.event(2)
    for d, (_, o) in zip(dev, scans):
        m.insert_point_cloud(d, o, 50.0, sync=False)
    ctx.event(3)
    total_ms = ctx.elapsed_ms(2, 3)
    if m.sync() != 0:
        raise RuntimeError("sequence failed")
    leaves = m.stats()["leaf_count"]
    m.close()
    # the same sequence again with per-kernel events (after the timed run):
    # microseconds a scan of each kernel
    m = bx.ProbabilisticMap(ctx, 0.1)
    m.reserve(int(st["leaf_count"] * 1.25) + 8192, int(st["root_count"] * 2) + 1024)
    ctx.sync()
    ctx.profile_reset()
    ctx.profile(True)
    for d, (_, o) in zip(dev, scans):
        m.insert_point_cloud(d, o, 50.0, sync=False)
    if m.sync() != 0:
        raise RuntimeError("profiled sequence failed")
    ctx.profile(False)
    kern_us = {k: ctx.profile_read(k)[1] * 1e3 / nscans for k in SCAN_KERNELS}
    m.close()
    for d in dev:
        d.free()
    npts = sum(len(p) for p, _ in scans)
    om = oracle.ProbMap(0.1)
    t0 = time.perf_counter()
    cpu_pts = cpu_scans = 0
    for pts, o in scans:
        om.insert_point_cloud(pts, o, 50.0)
        cpu_pts += len(pts)
        cpu_scans += 1
        if time.perf_counter() - t0 > cpu_budget_s:
            break
    cpu_s = time.perf_counter() - t0
    om.close()
    return {"scans": nscans, "points": npts, "ms": total_ms, "ms_per_scan": total_ms / nscans,
            "rays_per_s": npts / (total_ms / 1e3), "leaves": leaves,
            "kernels_us_per_scan": {k: v for k, v in kern_us.items() if v > 0},
            "cpu_rays_per_s": cpu_pts / cpu_s, "cpu_cores": 1,
            "cpu_sample": f"the first {cpu_scans} scans ({cpu_pts} points, {cpu_s:.1f} s)"}


SERVE_MAPS = 16
SERVE_FRAMES = 24


def serving_trace(nreq, rate, seed=21, nmaps=SERVE_MAPS):
    """A Poisson request stream: kind (0 insert of a 640x480 depth frame, 1
    lookup batch of 10k..1M coordinates, log-uniform, 2 occupied-voxel
    export), target map, lookup size and the gap before each request (all
    zero for rate None: every request submitted at once).  The lookup
    coordinates are four fixed sets of 1M, of which a batch takes a prefix."""
    rng = np.random.default_rng(seed)
    coords = [np.column_stack([rng.integers(-160, 160, 1_000_000), rng.integers(-160, 160, 1_000_000),
                               rng.integers(-20, 80, 1_000_000)]).astype(np.int32) for _ in range(4)]
    kinds = rng.choice(3, nreq, p=[0.5, 0.4, 0.1])
    which = rng.integers(0, nmaps, nreq)
    sizes = np.exp(rng.uniform(np.log(10_000), np.log(1_000_000), nreq)).astype(np.int64)
    gaps = rng.exponential(1.0 / rate, nreq) if rate else np.zeros(nreq)
    return {"coords": coords, "kinds": kinds, "which": which, "sizes": sizes, "gaps": gaps,
            "rate": rate, "nmaps": nmaps}


def latency_summary(infos, span_s):
    out = {}
    for kind in ("insert", "lookup", "export"):
        lat = np.array([i["latency_ms"] for kd, i, _ in infos if kd == kind])
        # an export's work is the number of cells it wrote, known only once it is done
        work = sum(i["count"] if kd == "export" else n for kd, i, n in infos if kd == kind)
        if len(lat):
            out[kind] = {"count": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)),
                         "p99_ms": float(np.percentile(lat, 99)), "items_per_s": work / span_s}
    return out


def run_dropin(scenes, nscans=200):
    """The drop-in C++ API driven as the reference's ROS node drives it
    (bonxai_amd/bin/dropin_bench): per scan of the 200-scan LiDAR sequence
    (64 x 1024 rays, 0.1 m, 50 m), host points in a std::vector through
    ProbabilisticMap::insertPointCloud (PCIe copy, scan, wait), then
    getOccupiedVoxels.  None where the program was not built."""
    import subprocess
    import tempfile
    exe = os.path.join(HERE, "bonxai_amd", "bin", "dropin_bench")
    if not os.path.exists(exe):
        return None
    scans = list(scenes.lidar_sequence(nscans))
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        path = f.name
        f.write(np.array([0.1, 50.0], np.float64).tobytes())
        f.write(np.array([len(scans)], np.uint32).tobytes())
        for pts, origin in scans:
            f.write(np.array([len(pts)], np.uint32).tobytes())
            f.write(np.asarray(origin, np.float64).tobytes())
            f.write(np.ascontiguousarray(pts, np.float32).tobytes())
    try:
        out = subprocess.run([exe, path], capture_output=True, text=True, timeout=300)
    finally:
        os.unlink(path)
    if out.returncode != 0:
        raise RuntimeError("dropin_bench failed: " + out.stderr.strip())
    r = json.loads(out.stdout)
    r["what"] = ("the C++ drop-in, as the reference's ROS node drives it: per LiDAR scan (config 3), "
                 "ProbabilisticMap::insertPointCloud(std::vector<point>) from host memory, then "
                 "getOccupiedVoxels; PCIe copies included")
    return r


def run_serving(ctx, bx, scenes, trace):
    """Serving form: 16 resident maps at 0.05 m and the request stream of
    `trace` -- inserts of 640x480 depth frames (max range 5 m), lookup
    batches, occupied-voxel exports -- admitted by the engine while earlier
    ones run.  Latency = completion - submission."""
    nmaps, nreq = trace["nmaps"], len(trace["kinds"])
    maps = [bx.ProbabilisticMap(ctx, 0.05) for _ in range(nmaps)]
    frames = []
    for k in range(SERVE_FRAMES):
        pts, origin = scenes.depth_frame(1000 + k)
        frames.append((ctx.to_device(pts), origin))
    big = [ctx.to_device(c) for c in trace["coords"]]
    nout = 2 * 32
    outs = [(ctx.array(1_000_000, np.uint32), ctx.array(1_000_000, np.uint8)) for _ in range(nout)]
    ex = [(ctx.array((8_000_000, 3), np.int32), ctx.array(8_000_000, np.uint32)) for _ in range(4)]
    ctx.sync()
    srv = bx.Server(ctx, max_inflight=32, max_batch_scans=16, streams=4)
    # resident maps: two frames each before the stream starts
    warm = [srv.insert(m, *frames[(2 * i + j) % len(frames)], 5.0) for i, m in enumerate(maps) for j in range(2)]
    for w in warm:
        if srv.wait(w)["status"] != 0:
            raise RuntimeError("serving warm-up insert failed")
    kinds, which, sizes, gaps = trace["kinds"], trace["which"], trace["sizes"], trace["gaps"]
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for k in range(nreq):
        t_next += gaps[k]
        while True:
            dt = t_next - time.perf_counter()
            if dt <= 0:
                break
            if dt > 3e-4:
                time.sleep(dt - 2e-4)
        m = maps[which[k]]
        if kinds[k] == 0:
            d, o = frames[k % len(frames)]
            ids.append(("insert", srv.insert(m, d, o, 5.0), d.shape[0]))
        elif kinds[k] == 1:
            v, f = outs[k % nout]
            ids.append(("lookup", srv.lookup(m, big[k % 4].view(int(sizes[k])), v, f), int(sizes[k])))
        else:
            c, v = ex[k % 4]
            ids.append(("export", srv.export(m, c, v, filter=bx.EXPORT_OCCUPIED), 0))
    t_sub = time.perf_counter() - t0
    infos = [(kind, srv.wait(i), n) for kind, i, n in ids]
    t_all = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    if any(i["status"] != 0 for _, i, _ in infos):
        raise RuntimeError("a serving request failed")
    first = min(i["submit_ns"] for _, i, _ in infos)
    last = max(i["done_ns"] for _, i, _ in infos)
    span = (last - first) / 1e9
    out = {"requests": nreq, "offered_rate_per_s": trace["rate"], "maps": nmaps, "submit_s": t_sub,
           "wall_s": t_all, "requests_per_s": nreq / span,
           "waves": stats["waves"], "scans_per_wave": stats["wave_scans"] / max(stats["waves"], 1),
           "deferred": stats["deferred"], "retried": stats["retried"]}
    out.update(latency_summary(infos, span))
    for d, _ in frames:
        d.free()
    for a in big + [x for pair in outs + ex for x in pair]:
        a.free()
    for m in maps:
        m.close()
    return out


def run_serving_sweep(ctx, bx, scenes, rates, nreq):
    """The serving form at rising offered rates (None: every request at
    once), to show where continuous batching engages and where it saturates."""
    rows = []
    for rate in rates:
        r = run_serving(ctx, bx, scenes, serving_trace(nreq, rate))
        rows.append({k: r[k] for k in ("offered_rate_per_s", "requests", "requests_per_s", "scans_per_wave",
                                      "waves", "deferred") if k in r}
                    | {kind: {k: r[kind][k] for k in ("p50_ms", "p99_ms")} for kind in ("insert", "lookup", "export")
                       if kind in r})
        log(f"serving at {rate or 'unlimited'} req/s offered: {r['requests_per_s']:.0f} req/s, "
            f"{r['scans_per_wave']:.2f} scans a wave")
    return rows


def cpu_serving_baseline(scenes, trace, nthreads, budget_s):
    """The oracle serving the same request stream on the host: the 16 maps
    split over nthreads worker threads (a map's requests in order on its
    worker), each map warmed with two frames as on the GPU, requests taken
    in stream order as fast as the workers go (the CPU cannot keep up with
    the offered rate, so this is its saturated throughput; latency is each
    request's own service time).  Stops once a worker has spent budget_s."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    nmaps = trace["nmaps"]
    frames = [scenes.depth_frame(1000 + k) for k in range(SERVE_FRAMES)]
    maps = [oracle.ProbMap(0.05) for _ in range(nmaps)]
    for i, m in enumerate(maps):
        for j in range(2):
            pts, o = frames[(2 * i + j) % len(frames)]
            m.insert_point_cloud(pts, o, 5.0)
    kinds, which, sizes = trace["kinds"], trace["which"], trace["sizes"]

    def worker(t):
        done = []
        t0 = time.perf_counter()
        for k in range(len(kinds)):
            if which[k] % nthreads != t:
                continue
            if time.perf_counter() - t0 > budget_s:
                break
            m = maps[which[k]]
            s0 = time.perf_counter()
            if kinds[k] == 0:
                pts, o = frames[k % len(frames)]
                m.insert_point_cloud(pts, o, 5.0)
                kind, n = "insert", len(pts)
            elif kinds[k] == 1:
                m.grid.lookup(trace["coords"][k % 4][:int(sizes[k])])
                kind, n = "lookup", int(sizes[k])
            else:
                n = len(m.grid.export(1, 0, sort=False)[0])
                kind = "export"
            done.append((kind, {"latency_ms": (time.perf_counter() - s0) * 1e3, "count": n}, n))
        return done, time.perf_counter() - t0

    with ThreadPoolExecutor(nthreads) as ex:
        t0 = time.perf_counter()
        res = list(ex.map(worker, range(nthreads)))
        wall = time.perf_counter() - t0
    for m in maps:
        m.close()
    infos = [x for d, _ in res for x in d]
    out = {"requests": len(infos), "requests_per_s": len(infos) / wall, "cores": nthreads, "kind": "port",
           "sample": f"the first requests of the stream, each worker for at most {budget_s:.0f} s "
                     f"({len(infos)} requests in {wall:.1f} s), oracle/bx_oracle.c -O2"}
    out.update(latency_summary(infos, wall))
    return out


def cpu_threads():
    """Worker threads for the threaded CPU baseline: the process's CPU share
    (OMP_NUM_THREADS when set), at most 16."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return max(1, min(n, 16))


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline_threads(scans, max_range, budget_s, nthreads):
    """One map per thread (the maps are independent; the oracle's C calls
    release the GIL): scans in frame order, nthreads at a time."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor

    def one(sc):
        pts, origin = sc
        m = oracle.ProbMap(0.1)
        m.insert_point_cloud(pts, origin, max_range)
        m.close()
        return len(pts)

    t0 = time.perf_counter()
    pts_done = maps_done = 0
    with ThreadPoolExecutor(nthreads) as ex:
        for i in range(0, len(scans), nthreads):
            pts_done += sum(ex.map(one, scans[i:i + nthreads]))
            maps_done += len(scans[i:i + nthreads])
            if time.perf_counter() - t0 > budget_s:
                break
    dt = time.perf_counter() - t0
    return pts_done / dt, maps_done, pts_done, dt


def cpu_baseline(scans, max_range, budget_s):
    import oracle
    t0 = time.perf_counter()
    pts_done = 0
    maps_done = 0
    for pts, origin in scans:
        m = oracle.ProbMap(0.1)
        m.insert_point_cloud(pts, origin, max_range)
        m.close()
        pts_done += len(pts)
        maps_done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return pts_done / dt, maps_done, pts_done, dt


def cpu_grid_baseline(coords, vals, q, nthreads):
    """Config 2 on the host's cores, over the reference's own VoxelGrid<float>
    and Accessor where oracle/_ref is built (the loops of the reference's
    benchmark, benchmark_bonxai.cpp:8-44: setValue per coordinate through one
    accessor, then value per query through a const accessor), else over the
    oracle's restatement.  One thread: the 2M inserts, then the 8M lookups.
    nthreads threads: the lookups split into nthreads slices over the one
    grid (a const accessor each; reads only), and the inserts split by root
    key (coordinate >> 5) into one grid per thread, each in input order, so
    every cell gets its writes in order from one thread."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    ref = oracle.ref_available()
    make = (lambda: oracle.RefGrid(0.05, np.float32)) if ref else (lambda: oracle.Grid(0.05, dtype=np.float32))
    g = make()
    t0 = time.perf_counter()
    g.set_values(coords, vals)
    t_set = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.lookup(q)
    t_get = time.perf_counter() - t0
    slices = np.array_split(np.arange(len(q)), nthreads)
    qs = [np.ascontiguousarray(q[s_]) for s_ in slices]
    with ThreadPoolExecutor(nthreads) as ex:
        t0 = time.perf_counter()
        list(ex.map(g.lookup, qs))
        t_get_mt = time.perf_counter() - t0
    g.close()
    r = coords >> 5
    owner = ((r[:, 0].astype(np.int64) * 73856093) ^ (r[:, 1].astype(np.int64) * 19349663)
             ^ (r[:, 2].astype(np.int64) * 83492791)) % nthreads
    parts = [(np.ascontiguousarray(coords[owner == k]), np.ascontiguousarray(vals[owner == k]))
             for k in range(nthreads)]
    grids = [make() for _ in range(nthreads)]
    with ThreadPoolExecutor(nthreads) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda k: grids[k].set_values(*parts[k]), range(nthreads)))
        t_set_mt = time.perf_counter() - t0
    for gg in grids:
        gg.close()
    kind = "reference" if ref else "port"
    return {
        "inserts_per_s": len(coords) / t_set, "lookups_per_s": len(q) / t_get, "cores": 1, "kind": kind,
        "threads": {"inserts_per_s": len(coords) / t_set_mt, "lookups_per_s": len(q) / t_get_mt,
                    "cores": nthreads, "kind": kind},
        "sample": f"the whole workload: {len(coords)} setValue then {len(q)} value lookups, "
                  + ("the reference's VoxelGrid<float> compiled from its headers (oracle/_ref)"
                     if ref else "oracle/bx_oracle.c -O2"),
    }


def cpu_reference_storage(scans, max_range):
    """The same scans with the same integration over the reference's own
    VoxelGrid<Cell> and Accessor (compiled from its headers into oracle/_ref),
    one thread: how the oracle's speed relates to the reference's storage.
    None where oracle/_ref was not built."""
    import oracle
    if not oracle.ref_available():
        return None
    t0 = time.perf_counter()
    pts_done = 0
    for pts, origin in scans:
        m = oracle.RefProb(0.1)
        m.insert_point_cloud(pts, origin, max_range)
        m.close()
        pts_done += len(pts)
    return pts_done / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--maps", type=int, default=64)
    ap.add_argument("--map-sets", type=int, default=20,
                    help="sets of empty maps timed back to back (at most --steps)")
    ap.add_argument("--map-set-gib", type=int, default=160,
                    help="device memory the map sets may take")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-lookups", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-steady", action="store_true",
                    help="skip the steady frames (so that a kernel trace holds fresh frames only)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the single box scan and the 200-scan LiDAR sequence")
    ap.add_argument("--serve-requests", type=int, default=800)
    ap.add_argument("--serve-rate", type=float, default=1500.0, help="offered requests per second")
    ap.add_argument("--ray-blocks-per-cu", type=int, default=0,
                    help="override the tuned workgroups per CU of the ray kernels")
    ap.add_argument("--tuning", action="append", default=[], metavar="KEY=VALUE",
                    help="set a launch setting of the context (repeatable), after the tuned ones")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB five times (the known-size "
                         "reference for scaling HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every figure is for one MI355X; "
                                   "multi-GPU sharding is designed but not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import bonxai_amd as bx
    from bonxai_amd import scenes

    if args.copy_probe:
        with bx.Context(0) as ctx:
            ctx.copy_probe(args.copy_probe << 20, reps=5)
        print(json.dumps({"copy_probe_bytes": args.copy_probe << 20, "reps": 5}))
        return

    max_range = 30.0
    t0 = time.time()
    scans = scenes.batch_scans(nmaps=args.maps, max_range=max_range)
    log(f"generated {len(scans)} scans, {sum(len(p) for p, _ in scans)} points in {time.time() - t0:.1f}s")
    ctx = bx.Context(0, tuned=True)
    if args.ray_blocks_per_cu:
        ctx.set_tuning("ray_blocks_per_cu", args.ray_blocks_per_cu)
    for kv in args.tuning:
        k, v = kv.split("=", 1)
        ctx.set_tuning(k, int(v))
    dev = [ctx.to_device(p) for p, _ in scans]

    t0 = time.time()
    sets, frame_leaves, set_bytes = make_map_sets(ctx, bx, scans, dev, max_range,
                                                  min(args.steps, args.map_sets), args.map_set_gib << 30)
    log(f"{len(sets)} sets of {len(scans)} maps ({frame_leaves} leaves a fresh frame, "
        f"{set_bytes / 2**30:.1f} GiB a set) in {time.time() - t0:.1f}s")
    b = run_batched(ctx, bx, scans, dev, args.steps, args.warmup, max_range, sets, steady=not args.no_steady)
    log(f"batched: fresh {b['fresh_ms'] / args.steps:.3f} ms/frame, steady "
        + (f"{b['steady_ms'] / args.steps:.3f} ms/frame" if b["steady_ms"] is not None else "not run")
        + f", replays {b['replays_all']}")
    p = None
    if not args.no_plain:
        p = run_plain(ctx, bx, scans, dev, args.steps, args.warmup, max_range, sets)
        log(f"plain: fresh {p['fresh_ms'] / args.steps:.3f} ms/frame, steady {p['steady_ms'] / args.steps:.3f}")
    for ms in sets:
        for m in ms:
            m.close()
    del sets
    lk = None
    if not args.no_lookups:
        lk = run_lookups(ctx, bx, scenes)
        log(f"lookups: {lk['lookup_ms']:.3f} ms per 8M; the hits' scattered reads alone "
            f"{lk['gather_ceiling']['ms']:.3f} ms")
    c1 = c3 = None
    if not args.no_configs:
        c1 = run_box_scan(ctx, bx, scenes)
        log(f"box scan: {c1['ms']:.3f} ms ({c1['rays_per_s'] / 1e6:.0f} M rays/s), oracle {c1['cpu_ms']:.0f} ms")
        c3 = run_sequence(ctx, bx, scenes)
        ku = c3["kernels_us_per_scan"]
        log(f"lidar sequence: {c3['ms_per_scan']:.3f} ms/scan ({c3['rays_per_s'] / 1e6:.0f} M rays/s); "
            f"update pass {ku.get('k_scan_rays', 0):.0f} us, leaf pass {ku.get('k_scan_ray_leaves', 0):.0f} us a scan")
        try:
            c3["dropin"] = run_dropin(scenes)
        except (RuntimeError, OSError, ValueError) as e:  # reported, not fatal to the benchmark
            c3["dropin"] = {"error": str(e)}
        if c3["dropin"] and "error" not in c3["dropin"]:
            log(f"drop-in C++ loop: insert p50 {c3['dropin']['insert_ms']['p50']:.3f} ms, with export "
                f"{c3['dropin']['insert_and_export_ms']['p50']:.3f} ms")
    sv = None
    if not args.no_serving:
        sv = run_serving(ctx, bx, scenes, serving_trace(args.serve_requests, args.serve_rate))
        log(f"serving: {sv['requests_per_s']:.0f} req/s, insert p50/p99 "
            f"{sv['insert']['p50_ms']:.2f}/{sv['insert']['p99_ms']:.2f} ms")
        sv["sweep"] = run_serving_sweep(ctx, bx, scenes, [1500.0, 3000.0, 6000.0, 12000.0, None],
                                        args.serve_requests // 2)
        sv["cpu_baseline"] = cpu_serving_baseline(scenes, serving_trace(args.serve_requests, args.serve_rate),
                                                  cpu_threads(), args.cpu_seconds / 2)
        log(f"serving on the host: {sv['cpu_baseline']['requests_per_s']:.1f} req/s "
            f"on {sv['cpu_baseline']['cores']} threads")

    npts = b["points"]
    fc, sc_ = b["fresh_counts"], b.get("steady_counts")
    ms_step = b["fresh_ms"] / args.steps
    rays_s = npts / (ms_step / 1e3)
    steady = b["steady_ms"] is not None
    ms_steady = b["steady_ms"] / args.steps if steady else None
    # voxel updates per frame: endpoint updates (one per traced ray) + ray
    # updates (every distinct cell a ray crossed whose update_id was not cur)
    upd = fc["rays"] + fc["ray_updates"]
    # the dominant kernel of a fresh frame: the fresh pass (maps empty at the
    # scan's start), or the update pass where the fresh pass is switched off.
    # Algorithmic bytes of one fresh-pass launch: per ray its 16-byte record
    # and 4-byte list entry, per updated cell its 1-byte mark, per new leaf
    # its 16-byte origin and 4-byte inner-grid entry.  Of one update-pass
    # launch: per ray 20 bytes, per updated cell one 4-byte read and one
    # 4-byte write (every distinct cell once).
    kf = b["kern_fresh"]
    fresh_pass = kf.get("k_scan_fresh_rays", 0.0) > kf["k_scan_rays"]
    if fresh_pass:
        dom = "k_scan_fresh_rays"
        alg_bytes = 20 * fc["rays"] + 1 * fc["ray_updates"] + 20 * fc["leaves"]
    else:
        dom = "k_scan_rays"
        alg_bytes = 20 * fc["rays"] + 8 * fc["ray_updates"]
    k_ms = kf[dom]
    achieved = alg_bytes / (k_ms / 1e3) / 1e9
    # the leaf assembly: per new leaf 512 mark bytes read, 2048 cell bytes
    # and 64 mask bytes written
    nl_ms = kf.get("k_scan_new_leaves", 0.0)
    nl_bytes = 2624 * fc["leaves"]
    # the whole frame on the integration model: 12 B a point read, 8 B per
    # updated cell, 64 B of mask words per new leaf; counting the leaves'
    # own initialisation (in a fresh frame every updated cell is in a new
    # leaf, which is written whole): 12 B a point, 2112 B a new leaf
    frame_bytes = 12 * npts + 8 * upd + 64 * fc["leaves"]
    frame_bytes_init = 12 * npts + (2048 + 64) * fc["leaves"]
    from bonxai_amd import autotune
    ctx_tuning = {k: ctx.get_tuning(k) for k in autotune.CANDIDATES}
    cpu_rate, cpu_maps, cpu_pts, cpu_dt = cpu_baseline(scans, max_range, args.cpu_seconds)
    ref_rate = cpu_reference_storage(scans[:cpu_maps], max_range)
    nthr = cpu_threads()
    mt_rate, mt_maps, mt_pts, mt_dt = cpu_baseline_threads(scans, max_range, args.cpu_seconds / 2, nthr)
    # visits as counted for CPU loops: every ray cell crossed, plus one
    # endpoint visit per point
    visits = fc["ray_visits"] + npts
    lat = np.array(b["lat_fresh"])
    out = {
        "metric": METRIC,
        "value": rays_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"batched insertPointCloud: {len(scans)} empty maps x 1 LiDAR scan each "
                        "(log-uniform 10k-500k points, box scene), voxel 0.1 m, max range 30 m, "
                        "one captured graph per frame",
            "maps": len(scans), "points_per_frame": npts, "voxel": 0.1, "max_range": max_range,
            "rays_traced_per_frame": fc["rays"], "ray_updates_per_frame": fc["ray_updates"],
            "ray_visits_per_frame": fc["ray_visits"], "leaves_per_frame": fc["leaves"],
            "input_sha256": scenes.digest_scans(scans),
            "timing": f"{args.steps} frames in {b['fresh_rounds']} timed region(s) of back-to-back graph "
                      "launches, each frame into its own set of empty maps; the sets are emptied between "
                      "regions, outside the timed events",
            "tuning": {k: ctx_tuning[k] for k in sorted(ctx_tuning)},
        },
        "rays_per_s_note": "value counts input points (one ray each as insertPointCloud takes them); "
                           "traced_rays_per_s counts the distinct endpoint voxels whose ray was walked",
        "traced_rays_per_s": fc["rays"] / (ms_step / 1e3),
        "voxel_updates_per_s": upd / (ms_step / 1e3),
        "visits_per_s": visits / (ms_step / 1e3),
        "frame_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "replays_after_timed_frames": b["replays_timed"],
        "device_bytes_allocated_by_timed_frames": b["alloc_timed"],
        "timed_maps_equal_oracle": b["timed_match"],
        "kernels_ms": b["kern_fresh"],
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
            "kernel": dom,
            "alg_bytes_per_launch": alg_bytes,
        },
        "new_leaves_roofline": None if nl_ms <= 0 else {
            "bound": "hbm", "kernel": "k_scan_new_leaves", "alg_bytes_per_launch": nl_bytes,
            "achieved": nl_bytes / (nl_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": nl_bytes / (nl_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "traffic": measured_traffic("k_scan_new_leaves"),
        },
        "frame_roofline": {
            "bound": "hbm", "alg_bytes_per_frame": frame_bytes,
            "achieved": frame_bytes / (ms_step / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": frame_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS,
            "alg_bytes_per_frame_with_leaf_init": frame_bytes_init,
            "frac_with_leaf_init": frame_bytes_init / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": {
            "value": cpu_rate, "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": f"the first {cpu_maps} scans of the frame ({cpu_pts} points, {cpu_dt:.1f} s), each "
                      "into an empty map, oracle/bx_oracle.c -O2, one thread",
            "host": host_info(),
            "reference_storage": None if ref_rate is None else {
                "value": ref_rate, "unit": "rays/s", "cores": 1,
                "sample": "the same scans, same integration, over the reference's own VoxelGrid<Cell> "
                          "and Accessor (bonxai.hpp compiled into oracle/_ref)",
                "port_over_reference": cpu_rate / ref_rate},
        },
        "cpu_baseline_threads": {
            "value": mt_rate, "unit": "rays/s", "cores": nthr, "kind": "port",
            "sample": f"the first {mt_maps} scans of the frame ({mt_pts} points, {mt_dt:.1f} s), "
                      f"one empty map per thread, {nthr} threads",
        },
    }
    if steady:
        upd_steady = sc_["rays"] + sc_["ray_updates"]
        k_ms_st = b["kern_steady"]["k_scan_rays"]
        alg_bytes_st = 20 * sc_["rays"] + 8 * sc_["ray_updates"]
        lat_st = np.array(b["lat_steady"])
        out["steady"] = {
            "what": "the same frame re-inserted back to back into the maps that hold it",
            "rays_per_s": npts / (ms_steady / 1e3), "ms_per_frame": ms_steady,
            "voxel_updates_per_s": upd_steady / (ms_steady / 1e3),
            "ray_updates_per_frame": sc_["ray_updates"],
            "frame_latency_ms": {"p50": float(np.percentile(lat_st, 50)),
                                 "p99": float(np.percentile(lat_st, 99))},
            "kernels_ms": b["kern_steady"],
            "roofline": {"bound": "hbm", "kernel": "k_scan_rays", "alg_bytes_per_launch": alg_bytes_st,
                         "achieved": alg_bytes_st / (k_ms_st / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                         "frac": alg_bytes_st / (k_ms_st / 1e3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": measured_traffic("k_scan_rays_steady")},
            "replays": b["replays_all"] - b["replays_timed"],
        }
    if p is not None:
        plain_rays_s = npts / (p["fresh_ms"] / args.steps / 1e3)
        plain_steady = npts / (p["steady_ms"] / args.steps / 1e3)
        out["plain"] = {"rays_per_s": plain_rays_s, "ms_per_frame": p["fresh_ms"] / args.steps,
                        "steady_rays_per_s": plain_steady, "steady_ms_per_frame": p["steady_ms"] / args.steps,
                        "scan_latency_ms": {"p50": float(np.percentile(p["lat"], 50)),
                                            "p99": float(np.percentile(p["lat"], 99))}}
        out["batched_over_plain"] = rays_s / plain_rays_s
        if steady:
            out["steady"]["batched_over_plain"] = (npts / (ms_steady / 1e3)) / plain_steady
    if lk is not None:
        # algorithmic bytes of a lookup, the declared model: 12-byte coordinate,
        # 8-byte mask word and 4-byte cell read, 4-byte value written (root and
        # inner-node reads are cache-resident and not counted)
        lk_bytes = LOOKUP_BYTES * lk["lookups"]
        lk_gbs = lk_bytes / (lk["lookup_ms"] / 1e3) / 1e9
        out["lookup"] = {**lk, "roofline": {"bound": "hbm", "achieved": lk_gbs, "peak": HBM_PEAK_GBS,
                                            "unit": "GB/s", "frac": lk_gbs / HBM_PEAK_GBS,
                                            "traffic": measured_traffic("k_lookup"),
                                            "bytes_per_lookup": LOOKUP_BYTES}}
        # algorithmic bytes of a setValue, the declared model: 12-byte
        # coordinate and 4-byte value read, 4-byte cell write, 8-byte mask-word
        # read-modify-write
        ins_gbs = INSERT_BYTES * lk["inserts"] / (lk["insert_ms"] / 1e3) / 1e9
        out["lookup"]["insert_roofline"] = {"bound": "hbm", "achieved": ins_gbs, "peak": HBM_PEAK_GBS,
                                            "unit": "GB/s", "frac": ins_gbs / HBM_PEAK_GBS,
                                            "bytes_per_insert": INSERT_BYTES,
                                            "kernels": "k_op_clear + k_op_begin + k_op_resolve + k_zero_new_leaves + k_op_apply"}
    if c1 is not None:
        out["box_scan"] = c1
        out["lidar_sequence"] = c3
    if sv is not None:
        out["serving"] = sv
    for d in dev:
        d.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
