"""bxg_insert_point_cloud_noalloc, the allocation-free form of
bxg_insert_point_cloud, through the C ABI: two calls in a row (the second
with fewer points) give maps bit-identical to those of the allocating entry
point on the same inputs; no pool grows; a scan the reserved pools cannot
hold changes no cell and succeeds once more is reserved; a call beyond the
maximum size is refused on the host."""
import numpy as np
import pytest

from conftest import sorted_cells

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    ac, aw = sorted_cells(*a.cells())
    bc, bw = sorted_cells(*b.cells())
    assert ac.shape == bc.shape, f"{what}: {len(ac)} vs {len(bc)} cells"
    assert np.array_equal(ac, bc) and np.array_equal(aw, bw), what
    assert a.update_count == b.update_count, what


def test_noalloc_equals_allocating_entry_point(ctx):
    import bonxai_amd as bx
    from bonxai_amd import scenes
    print("noalloc: two box scans, 200k then 50k points, 0.05 m", flush=True)
    scans = [scenes.box_scan(seed=3, n=200_000), scenes.box_scan(seed=4, n=50_000)]
    a = bx.ProbabilisticMap(ctx, 0.05)
    b = bx.ProbabilisticMap(ctx, 0.05)
    a.reserve(400_000, 8192)
    ctx.sync()
    dev = [ctx.to_device(p) for p, _ in scans]
    held = a.stats()["device_bytes"]
    for k, (d, (_, origin)) in enumerate(zip(dev, scans)):
        assert a.insert_point_cloud_noalloc(d, origin, 10.0) == bx.OK, ctx.last_error()
        b.insert_point_cloud(d, origin, 10.0)  # the allocating entry point
        _same(a, b, f"call {k}")
    st = a.stats()
    assert st["device_bytes"] == held  # the pools did not grow
    assert st["replays"] == 0 and st["status"] == 0
    for d in dev:
        d.free()
    a.close()
    b.close()


def test_noalloc_out_of_pool_then_reserved(ctx):
    import bonxai_amd as bx
    from bonxai_amd import scenes
    print("noalloc: LiDAR scans into a map with a small pool", flush=True)
    scans = list(scenes.lidar_sequence(2))
    dev = [ctx.to_device(p) for p, _ in scans]
    a = bx.ProbabilisticMap(ctx, 0.1)
    b = bx.ProbabilisticMap(ctx, 0.1)
    a.reserve(1024, 1024)
    ctx.sync()
    before = a.stats()
    rc = a.insert_point_cloud_noalloc(dev[0], scans[0][1], 50.0)
    assert rc == bx.ERR_OUT_OF_POOL
    after = a.stats()
    assert after["status"] == 0 and after["update_count"] == before["update_count"]
    assert after["leaf_capacity"] == before["leaf_capacity"]
    assert len(a.cells()[0]) == 0  # no cell changed
    a.reserve(400_000, 8192)
    ctx.sync()
    for k, (d, (_, origin)) in enumerate(zip(dev, scans)):
        assert a.insert_point_cloud_noalloc(d, origin, 50.0) == bx.OK, ctx.last_error()
        b.insert_point_cloud(d, origin, 50.0)
        _same(a, b, f"LiDAR scan {k}")
    for d in dev:
        d.free()
    a.close()
    b.close()


def test_noalloc_refuses_more_than_the_maximum(ctx):
    import ctypes
    import bonxai_amd as bx
    m = bx.ProbabilisticMap(ctx, 0.1)
    d = ctx.array((16, 3), np.float32)
    o = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    # refused on the host before any device work (the points are not read)
    rc = bx.lib().bxg_insert_point_cloud_noalloc(m._h, d.ptr, 524_289, 0, o, 10.0)
    assert rc == bx.ERR_INVALID_ARG
    assert "BXG_NOALLOC_MAX_POINTS" in ctx.last_error()
    d.free()
    m.close()
