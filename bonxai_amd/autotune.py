"""Autotuner of the batched scan-integration launch.

Sweeps, on a frame of uneven scans into many maps (captured-graph launches,
as the batched form runs them; by default each frame into emptied maps, the
benchmark's workload):
  ray_block           workgroup size of the two work-stealing ray walks
  ray_blocks_per_cu   workgroups per CU of the cell-update walk
  leaf_blocks_per_cu  workgroups per CU of the leaf-allocation walk
  ray_chunk           most rays a wavefront takes per cursor atomic
                      (rays per lane = ray_chunk / 64)
  ep_table_factor     endpoint-table slots per point (inverse load factor)
  root_slack          root-directory slots per root (inverse load factor)
  ray_window          cells a ray may run ahead of its wavefront's slowest
                      ray in the update pass (1 << 20: no window)
  leaf_window         the same in leaves for the leaf pass
  fresh_pass          maps empty at a scan's start: one create-and-mark walk
                      (1) or the leaf and update walks (0)
  fresh_blocks_per_cu workgroups per CU of the create-and-mark walk
  fresh_window        its lockstep window, in cells
  ray_order           winner order: direction only (0), or classes of ray
                      length and direction (1..5)
by coordinate descent (the fresh-pass and shared keys on frames into
emptied maps, the workload; the update- and leaf-pass keys on frames
re-inserted into filled maps) from the defaults, and records the fastest setting per
device (architecture name + CU count) in bonxai_amd/tuning.json.
Context(tuned=True) applies the recorded setting of its device.

Usage: python -m bonxai_amd.autotune [--maps 64] [--passes 2] [--state both|fresh|steady] [--out PATH]
"""
import argparse
import json
import os
import sys
import time


HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_FILE = os.path.join(HERE, "tuning.json")

CANDIDATES = {
    "ray_block": [64, 128, 256],
    "ray_blocks_per_cu": [2, 4, 6, 8, 12, 16],
    "leaf_blocks_per_cu": [4, 6, 8, 12, 16, 32],
    "ray_chunk": [64, 128, 256, 512],
    "ep_table_factor": [2, 4],
    "root_slack": [2, 4],
    "ray_window": [4, 8, 16, 32, 64, 1 << 20],
    "leaf_window": [8, 16, 32, 64, 1 << 20],
    "fresh_pass": [1, 0],
    "fresh_blocks_per_cu": [2, 4, 6, 8, 12, 16],
    "fresh_window": [4, 8, 16, 32, 64, 1 << 20],
    "ray_order": [3, 4, 2, 5, 1, 0],
}
# keys searched on each kind of frame: frames into empty maps (the workload)
# run the fresh pass; frames re-inserted into filled maps the leaf and
# update passes.  Keys of both are searched on fresh frames.
FRESH_KEYS = ["ray_block", "ray_chunk", "ep_table_factor", "root_slack", "fresh_pass",
              "fresh_blocks_per_cu", "fresh_window", "ray_order"]
STEADY_KEYS = ["ray_blocks_per_cu", "leaf_blocks_per_cu", "ray_window", "leaf_window"]


def device_key(ctx):
    info = ctx.device_info()
    return f"{info['arch']}|{info['cu_count']}"


def load(ctx, path=TUNING_FILE):
    """The recorded setting of ctx's device, or None."""
    try:
        with open(path) as f:
            return json.load(f).get(device_key(ctx), {}).get("settings")
    except (OSError, ValueError):
        return None


def apply(ctx, settings):
    for k, v in settings.items():
        ctx.set_tuning(k, int(v))


def empty(ctx, bx, maps):
    """Empty the maps (pools keep their capacity), as new maps are."""
    for m in maps:
        m.clear(bx.CLEAR_MEMORY)
        m.update_count = 1
    ctx.sync()


def frame_ms(ctx, bx, descs, npoints, maps, fresh=True, reps=5, warmup=2):
    """Milliseconds per frame of a graph captured with the current tuning:
    each frame into emptied maps (fresh), or re-inserted into maps that hold
    it (steady)."""
    g = bx.ScanGraph(ctx, len(descs), npoints)
    g.set_scans(descs)
    for _ in range(warmup):
        if fresh:
            empty(ctx, bx, maps)
        g.launch()
    ctx.sync()
    total = 0.0
    if fresh:
        for _ in range(reps):
            empty(ctx, bx, maps)
            ctx.event(0)
            g.launch()
            ctx.event(1)
            total += ctx.elapsed_ms(0, 1)
    else:
        ctx.event(0)
        for _ in range(reps):
            g.launch()
        ctx.event(1)
        total = ctx.elapsed_ms(0, 1)
    g.close()
    for m in maps:
        if m.sync() != 0:
            raise RuntimeError("a map failed during tuning")
    return total / reps


def search(ctx, bx, scans, passes=2, fresh=True, log=print, keys=None, start=None):
    max_range = 30.0
    maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    dev = [ctx.to_device(p) for p, _ in scans]
    descs = [(m, d, o, max_range) for m, d, (_, o) in zip(maps, dev, scans)]
    npoints = sum(len(p) for p, _ in scans)
    bx.insert_batch(ctx, descs)  # sizes the pools (the library's head-room policy)
    empty(ctx, bx, maps)
    bx.insert_batch(ctx, descs)
    best = {k: ctx.get_tuning(k) for k in CANDIDATES}
    if start:
        best.update(start)
    apply(ctx, best)

    def measure():
        return frame_ms(ctx, bx, descs, npoints, maps, fresh)

    default_ms = best_ms = measure()
    table = [(dict(best), best_ms)]
    log(f"default {best} {best_ms:.3f} ms")
    for p in range(passes):
        for key, values in CANDIDATES.items():
            if keys is not None and key not in keys:
                continue
            for v in values:
                if v == best[key]:
                    continue
                trial = dict(best, **{key: v})
                apply(ctx, trial)
                ms = measure()
                table.append((trial, ms))
                log(f"pass {p} {key}={v}: {ms:.3f} ms")
                if ms < best_ms * 0.99:
                    best, best_ms = trial, ms
            apply(ctx, best)
    for m in maps:
        if m.sync() != 0:
            raise RuntimeError("a map failed during tuning")
        m.close()
    for d in dev:
        d.free()
    return best, best_ms, default_ms, table


def main(argv=None):
    ap = argparse.ArgumentParser(description="autotune the batched scan launch")
    ap.add_argument("--maps", type=int, default=64)
    ap.add_argument("--state", choices=["both", "fresh", "steady"], default="both",
                    help="time frames into emptied maps (fresh, the workload), re-inserted ones, "
                         "or both: the fresh-pass keys on fresh frames, then the update-pass keys on "
                         "re-inserted ones")
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    import bonxai_amd as bx
    from bonxai_amd import scenes
    scans = scenes.batch_scans(nmaps=args.maps, max_range=30.0)
    ctx = bx.Context(0)
    t0 = time.time()
    log = lambda *a: print(*a, flush=True)  # noqa: E731
    extra = {}
    if args.state == "both":
        best, best_ms, default_ms, table = search(ctx, bx, scans, args.passes, fresh=True, log=log,
                                                  keys=FRESH_KEYS)
        best_st, st_ms, st_default, table_st = search(ctx, bx, scans, args.passes, fresh=False, log=log,
                                                      keys=STEADY_KEYS, start=best)
        best = best_st
        table = table + table_st
        extra = {"steady_frame_ms": st_ms, "default_steady_frame_ms": st_default}
    else:
        best, best_ms, default_ms, table = search(ctx, bx, scans, args.passes, fresh=args.state == "fresh",
                                                  log=log)
    rec = {}
    try:
        with open(args.out) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        pass
    rec[device_key(ctx)] = {
        "settings": best, "frame_ms": best_ms, "default_frame_ms": default_ms,
        "workload": f"{len(scans)} LiDAR scans into {len(scans)} "
                    f"{ {'fresh': 'empty', 'steady': 'filled', 'both': 'empty (then filled)'}[args.state] } "
                    f"maps, 0.1 m, 30 m, {sum(len(p) for p, _ in scans)} points, captured graph",
        "evaluations": len(table), "seconds": time.time() - t0, **extra,
    }
    with open(args.out, "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
    print(json.dumps(rec[device_key(ctx)]), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
