"""GPU parity when leaves are created in pool memory that held other leaves.

clear(CLEAR_MEMORY) and pool growth leave the pool's old bytes in place; a
leaf is initialised when it is created (the scan's leaf assembly, the bulk
calls' zeroing pass), as the reference's Grid constructor initialises a new
leaf.  Each test first fills the pool with different cells, so that a leaf
created later over stale bytes would show them.
"""
import numpy as np
import pytest

from conftest import assert_same_cells

pytestmark = pytest.mark.gpu


def test_scans_into_cleared_pool(ctx):
    """A LiDAR scan fills the map; after clear(CLEAR_MEMORY) other scans
    (other origins and points, one of them counted) go into the same pool:
    every cell, the counts and the update ids equal the oracle's map that
    received only those scans."""
    import bonxai_amd as bx
    import oracle
    from bonxai_amd import scenes
    scans = list(scenes.lidar_sequence(12))
    m = bx.ProbabilisticMap(ctx, 0.1)
    m.insert_point_cloud(*scans[0], 50.0)
    m.insert_point_cloud(*scans[1], 50.0)
    filled = m.stats()["leaf_count"]
    m.clear(bx.CLEAR_MEMORY)
    m.update_count = 1
    om = oracle.ProbMap(0.1)
    for k in (11, 6):
        m.insert_point_cloud(*scans[k], 50.0)
        om.insert_point_cloud(*scans[k], 50.0)
        assert_same_cells(m.cells(), om.cells(), f"scan {k} into the cleared pool")
    assert m.update_count == om.update_count
    assert m.digest() == om.digest()
    print(f"{filled} leaves before the clear, {m.stats()['leaf_count']} after", flush=True)
    # the counted form marks cells with atomics: same map, exact counts
    m.clear(bx.CLEAR_MEMORY)
    m.update_count = 1
    om2 = oracle.ProbMap(0.1)
    ctx.counting(True)
    m.insert_point_cloud(*scans[3], 50.0)
    ctx.counting(False)
    om2.insert_point_cloud(*scans[3], 50.0)
    assert_same_cells(m.cells(), om2.cells(), "counted scan into the cleared pool")
    st = m.stats()
    ost = om2.stats()
    assert st["last_rays"] == ost["rays"] and st["last_ray_updates"] == ost["ray_updates"]
    assert st["last_ray_visits"] == ost["visits"]
    m.close()
    om.close()
    om2.close()


def test_bulk_calls_into_cleared_pool(ctx):
    """setValue, setCellOn and setCellOff after clear(CLEAR_MEMORY), over
    leaves that held other values: was_on of every input, the values and the
    on cells equal the oracle's fresh grid."""
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(17)
    g = bx.VoxelGrid(ctx, 0.1, dtype=np.int32)
    fill = rng.integers(-60, 60, (200_000, 3)).astype(np.int32)
    g.set_values(fill, rng.integers(1, 10**9, len(fill)).astype(np.int32))
    g.clear(bx.CLEAR_MEMORY)
    og = oracle.Grid(0.1, dtype=np.int32)
    a = rng.integers(-60, 60, (50_000, 3)).astype(np.int32)
    va = rng.integers(-10**9, 10**9, len(a)).astype(np.int32)
    assert np.array_equal(g.set_values(a, va, True), og.set_values(a, va))
    b = rng.integers(-70, 70, (50_000, 3)).astype(np.int32)
    vb = rng.integers(-10**9, 10**9, len(b)).astype(np.int32)
    assert np.array_equal(g.set_cells_on(b, vb, True), og.set_cells_on(b, vb))
    c = rng.integers(-70, 70, (20_000, 3)).astype(np.int32)
    assert np.array_equal(g.set_cells_off(c, True), og.set_cells_off(c))
    assert_same_cells(g.export(), og.export(), "bulk calls into the cleared pool")
    q = rng.integers(-75, 75, (100_000, 3)).astype(np.int32)
    v, f = g.lookup(q)
    ov, of = og.lookup(q)
    assert np.array_equal(f, of) and np.array_equal(v[f], ov[of])
    g.close()
    og.close()


def test_failed_scans_into_reused_pool(ctx):
    """A map filled, cleared and shrunk by release_unused_memory, then given
    asynchronous scans that overflow its small pool: the leaves the failed
    runs created are written off, the scans run again after growth, and the
    map equals the oracle's."""
    import bonxai_amd as bx
    import oracle
    from bonxai_amd import scenes
    scans = list(scenes.lidar_sequence(6))
    m = bx.ProbabilisticMap(ctx, 0.1)
    for p, o in scans[:2]:
        m.insert_point_cloud(p, o, 50.0)
    m.clear(bx.CLEAR_MEMORY)
    m.update_count = 1
    m.release_unused_memory()
    om = oracle.ProbMap(0.1)
    tiny, o0 = scans[2][0][:300], scans[2][1]
    for _ in range(2):
        m.insert_point_cloud(tiny, o0, 50.0)
        om.insert_point_cloud(tiny, o0, 50.0)
    cap0 = m.stats()["leaf_capacity"]
    dev = [ctx.to_device(p) for p, _ in scans[3:]]
    for d, (p, o) in zip(dev, scans[3:]):
        m.insert_point_cloud(d, o, 50.0, sync=False)
        om.insert_point_cloud(p, o, 50.0)
    assert m.sync() == 0
    st = m.stats()
    print(f"leaf capacity {cap0} -> {st['leaf_capacity']}, replays {st['replays']}", flush=True)
    assert st["leaf_count"] > cap0
    assert_same_cells(m.cells(), om.cells(), "failed scans into a reused pool")
    assert m.update_count == om.update_count
    for d in dev:
        d.free()
    m.close()
    om.close()
