"""GPU parity of ProbabilisticMap scan integration against the oracle.

Bit-exact bar: the same allocated (on) coordinate set, the same raw CellT
word ((log-odds << 4) | update_id) in every cell, and the same update_count,
after every scan.
"""
import numpy as np
import pytest

from conftest import assert_same_cells

pytestmark = pytest.mark.gpu


def _oracle():
    import oracle
    return oracle


def _gpu_cells(m):
    c, w = m.cells()
    return c, w


def test_fp64_front_end_is_ieee(ctx):
    """sqrt, division, the unfused (a/b)*c + c, floor(a*b) and the ordered
    squared norm of the scan front end round exactly as IEEE-754 double."""
    rng = np.random.default_rng(5)
    n = 1 << 20
    a = np.concatenate([rng.random(n // 2) * 1e4, np.exp(rng.normal(0, 20, n // 2))])
    b = np.concatenate([rng.random(n // 2) * 50 + 1e-3, np.exp(rng.normal(0, 5, n // 2))])
    c = rng.normal(0, 30, n)
    # values landing next to voxel boundaries: k*0.05 +- a few ulp
    k = rng.integers(-10**6, 10**6, 4096).astype(np.float64) * 0.05
    a[:4096] = np.nextafter(k, k + np.where(rng.random(4096) < 0.5, -1, 1))
    b[:4096] = 20.0
    out = ctx.fp64_probe(a, b, c)
    with np.errstate(all="ignore"):
        q = a / b
        exp = np.stack([np.sqrt(a), q, c + (q * c), np.floor(a * b), (a * a + b * b) + c * c], 1)
    for j, name in enumerate(["sqrt", "div", "c+(a/b)*c", "floor(a*b)", "sqnorm"]):
        assert np.array_equal(out[:, j].view(np.uint64), exp[:, j].view(np.uint64)), name


def test_box_scan_single(ctx):
    """One 100k-point box-room scan at 0.05 m, max range 10 m."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    pts, origin = scenes.box_scan()
    om = ora.ProbMap(0.05)
    om.insert_point_cloud(pts, origin, 10.0)
    m = bx.ProbabilisticMap(ctx, 0.05)
    m.insert_point_cloud(pts, origin, 10.0)
    assert_same_cells(_gpu_cells(m), om.cells(), "box scan")
    assert m.update_count == om.update_count == 2
    assert m.digest() == om.digest()
    m.close()
    om.close()


def test_box_scan_clipped_and_repeated(ctx):
    """Max range 2.5 m (most points clip to misses) and the same scan three
    more times, through update_count 2 -> 3 -> 1 -> 2."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    pts, origin = scenes.box_scan(seed=9, n=30_000)
    om = ora.ProbMap(0.05)
    m = bx.ProbabilisticMap(ctx, 0.05)
    for k in range(4):
        om.insert_point_cloud(pts, origin, 2.5)
        m.insert_point_cloud(pts, origin, 2.5)
        assert_same_cells(_gpu_cells(m), om.cells(), f"scan {k}")
        assert m.update_count == om.update_count
    m.close()
    om.close()


def test_lidar_sequence_prefix(ctx):
    """The first scans of the sequential LiDAR workload (0.1 m, 50 m), with a
    full comparison after each scan."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    om = ora.ProbMap(0.1)
    m = bx.ProbabilisticMap(ctx, 0.1)
    for k, (pts, origin) in enumerate(scenes.lidar_sequence(3)):
        om.insert_point_cloud(pts, origin, 50.0)
        m.insert_point_cloud(pts, origin, 50.0)
        assert m.digest() == om.digest(), f"scan {k}"
    assert_same_cells(_gpu_cells(m), om.cells(), "after 3 scans")
    m.close()
    om.close()


def test_lidar_sequence_against_digests(ctx):
    """Sequential LiDAR mapping into one map (64 x 1024 rays a scan, 0.1 m,
    50 m max range, 1 m/scan): the map's order-free digest and the scan's
    endpoint/ray statistics (rays, cells updated, cells crossed) after every
    scan equal the oracle's, committed in
    tests/golden/digests.json (all 200 scans when the full set was generated,
    else its first scans)."""
    import json
    import os
    import bonxai_amd as bx
    from bonxai_amd import scenes
    with open(os.path.join(os.path.dirname(__file__), "golden", "digests.json")) as f:
        want = json.load(f)["maps"]["lidar_sequence"]
    m = bx.ProbabilisticMap(ctx, 0.1)
    ctx.counting(True)
    try:
        for k, (pts, origin) in enumerate(scenes.lidar_sequence(len(want))):
            m.insert_point_cloud(pts, origin, 50.0)
            got = [str(x) for x in m.digest()]
            assert got == want[k]["digest"], f"scan {k}"
            st = m.stats()
            assert st["last_rays"] == want[k]["stats"]["rays"], f"scan {k} rays"
            assert st["last_ray_updates"] == want[k]["stats"]["ray_updates"], f"scan {k} ray updates"
            assert st["last_ray_visits"] == want[k]["stats"]["visits"], f"scan {k} ray visits"
    finally:
        ctx.counting(False)
    print(f"{len(want)} scans, {m.stats()['leaf_count']} leaves", flush=True)
    m.close()


def test_batched_frame_against_digests(ctx):
    """The batched workload (64 maps, one scan each, log-uniform 10k..500k
    points, 0.1 m, 30 m) in one batched pass: every map's digest and scan
    statistics equal the oracle's in tests/golden/digests.json (all 64 maps
    when the full set was generated, else the first ones)."""
    import json
    import os
    import bonxai_amd as bx
    from bonxai_amd import scenes
    with open(os.path.join(os.path.dirname(__file__), "golden", "digests.json")) as f:
        want = json.load(f)["maps"]["batch"]
    scans = scenes.batch_scans()
    maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    dev = [ctx.to_device(p) for p, _ in scans]
    ctx.counting(True)
    try:
        bx.insert_batch(ctx, [(m, d, o, 30.0) for m, d, (_, o) in zip(maps, dev, scans)])
        for w in want:
            m = maps[w["map"]]
            assert m.sync() == 0
            assert [str(x) for x in m.digest()] == w["digest"], f"map {w['map']}"
            st = m.stats()
            assert st["last_rays"] == w["stats"]["rays"] and st["last_ray_updates"] == w["stats"]["ray_updates"]
            assert st["last_ray_visits"] == w["stats"]["visits"], f"map {w['map']} ray visits"
    finally:
        ctx.counting(False)
    print(f"{len(want)} maps checked", flush=True)
    for m in maps:
        m.close()
    for d in dev:
        d.free()


def test_id_cycle_and_frozen_cells(ctx):
    """Eight small scans from alternating origins exercise the 3-value
    update_id cycle, including cells frozen because their id equals the
    scan's id (probabilistic_map.cpp:103-105)."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    om = ora.ProbMap(0.1)
    m = bx.ProbabilisticMap(ctx, 0.1)
    for k in range(8):
        pts, origin = scenes.lidar_scan(3, 77 + k, float(k % 3) * 2.0, 16, 64, far_range=30.0)
        rng_keep = np.random.default_rng(k).random(len(pts)) < (0.5 if k % 2 else 1.0)
        pts = pts[rng_keep]
        om.insert_point_cloud(pts, origin, 20.0)
        m.insert_point_cloud(pts, origin, 20.0)
        assert_same_cells(_gpu_cells(m), om.cells(), f"scan {k}")
        assert m.update_count == om.update_count
    m.close()
    om.close()


def test_kat_first_index_wins(ctx):
    """A hit and a clipped miss landing in the same voxel: the first point in
    input order decides the update (probabilistic_map.cpp:30-54)."""
    import bonxai_amd as bx
    origin = np.zeros(3)
    R = 1.0
    hit = np.array([0.99, 0.0, 0.0], np.float32)      # |v| < R: hit at x=0.99
    miss = np.array([1.5, 0.0, 0.0], np.float32)      # clipped to x=1.0 -> voxel 10
    near = np.array([0.999, 0.001, 0.0], np.float32)  # hit in voxel 9
    for order, kind in (((hit, miss), "hit first"), ((miss, hit), "miss first")):
        # voxel of the hit: floor(0.99*10)=9; of the clipped miss: floor(1.0*10)=10.
        # Use a second hit in the miss voxel to share it.
        pts = np.stack([order[0], order[1], near])
        m = bx.ProbabilisticMap(ctx, 0.1)
        m.insert_point_cloud(pts, origin, R)
        om = _oracle().ProbMap(0.1)
        om.insert_point_cloud(pts, origin, R)
        assert_same_cells(_gpu_cells(m), om.cells(), kind)
        m.close()
        om.close()
    # same voxel, both kinds: point at 1.04 (hit, R=1.05) and a far point
    # clipped onto 1.05*(direction) -> both in voxel 10 at 0.1 m
    R = 1.05
    a = np.array([1.04, 0.0, 0.0], np.float32)
    b = np.array([5.0, 0.0, 0.0], np.float32)
    for pts, expect in ((np.stack([a, b]), 847297), (np.stack([b, a]), -405465)):
        m = bx.ProbabilisticMap(ctx, 0.1)
        m.insert_point_cloud(pts, origin, R)
        c, w = m.cells()
        sel = np.all(c == [10, 0, 0], axis=1)
        assert sel.sum() == 1
        assert (w[sel].astype(np.int32) >> 4)[0] == expect
        m.close()


def test_kat_empty_scan_and_origin_voxel(ctx):
    """An empty scan still advances update_count; a point in the origin
    voxel is an endpoint without a ray."""
    import bonxai_amd as bx
    m = bx.ProbabilisticMap(ctx, 0.1)
    m.insert_point_cloud(np.zeros((0, 3), np.float32), np.zeros(3), 5.0)
    assert m.update_count == 2
    m.insert_point_cloud(np.array([[0.05, 0.05, 0.05]], np.float32), np.zeros(3), 5.0)
    c, w = m.cells()
    assert c.tolist() == [[0, 0, 0]]
    assert int(w[0]) == ((847297 << 4) | 2)
    assert m.update_count == 3
    m.close()


def test_negative_coords_and_root_crossings(ctx):
    """Rays crossing root (32-voxel) and leaf boundaries in every octant."""
    import bonxai_amd as bx
    ora = _oracle()
    rng = np.random.default_rng(3)
    pts = (rng.random((20_000, 3)) - 0.5) * 14.0
    origin = np.array([0.013, -0.021, 0.004])
    om = ora.ProbMap(0.1)
    m = bx.ProbabilisticMap(ctx, 0.1)
    for k in range(2):
        om.insert_point_cloud(pts.astype(np.float32), origin + k, 6.0)
        m.insert_point_cloud(pts.astype(np.float32), origin + k, 6.0)
    assert_same_cells(_gpu_cells(m), om.cells(), "octants")
    m.close()
    om.close()


def test_double_points(ctx):
    """insertPointCloud with double points (Eigen::Vector3d callers)."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    pts, origin = scenes.box_scan(seed=4, n=20_000)
    p64 = pts.astype(np.float64) + 1e-9
    om = ora.ProbMap(0.05)
    om.insert_point_cloud(p64, origin, 4.0)
    m = bx.ProbabilisticMap(ctx, 0.05)
    m.insert_point_cloud(p64, origin, 4.0)
    assert_same_cells(_gpu_cells(m), om.cells(), "f64 points")
    m.close()
    om.close()


def test_pool_growth_replay(ctx):
    """A scan that overflows a tiny node pool is re-run after the pool grows,
    and the result is still exact (cells are untouched by the failed run)."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    pts, origin = scenes.box_scan(seed=5)
    m = bx.ProbabilisticMap(ctx, 0.05)
    st0 = m.stats()
    assert st0["leaf_capacity"] < 5000
    m.insert_point_cloud(pts, origin, 10.0)
    om = ora.ProbMap(0.05)
    om.insert_point_cloud(pts, origin, 10.0)
    assert_same_cells(_gpu_cells(m), om.cells(), "after growth")
    assert m.stats()["leaf_count"] == om.grid.leaf_count()
    m.close()
    om.close()


@pytest.mark.parametrize("bits", [(1, 2), (2, 4), (3, 1)])
def test_other_leaf_sizes_same_cells(ctx, bits):
    """A scan's cells and values do not depend on how the grid is cut into
    leaves: maps with 4^3, 16^3 and 2^3 leaves (the ray and leaf walks at
    those sizes) equal the oracle's 8^3-leaf map, for a box scan and a LiDAR
    scan."""
    import bonxai_amd as bx
    from bonxai_amd import scenes

    class Map(bx.ProbabilisticMap):
        def __init__(self, c, res, ib, lb):
            bx.VoxelGrid.__init__(self, c, res, ib, lb, np.uint32, _flags=bx.MAP_PROBABILISTIC)

    ora = _oracle()
    pts, origin = scenes.box_scan(seed=9, n=30_000)
    lidar, lo = next(iter(scenes.lidar_sequence(1)))
    for res, (p, o, r) in ((0.05, (pts, origin, 10.0)), (0.1, (lidar, lo, 50.0))):
        m = Map(ctx, res, *bits)
        om = ora.ProbMap(res)
        for _ in range(2):
            m.insert_point_cloud(p, o, r)
            om.insert_point_cloud(p, o, r)
        assert_same_cells(_gpu_cells(m), om.cells(), f"bits {bits} res {res}")
        assert m.digest() == om.digest()
        m.close()
        om.close()


def test_async_inserts_replay_after_overflow(ctx):
    """Asynchronous device inserts whose head-room estimate is far too small
    (two tiny scans first teach it a small growth): the scans that overflowed
    the node pool, and every later one, run again in order when the map is
    next confirmed, and the map equals the oracle's."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    scans = list(scenes.lidar_sequence(4))
    m = bx.ProbabilisticMap(ctx, 0.1)
    om = ora.ProbMap(0.1)
    tiny, o0 = scans[0][0][:300], scans[0][1]
    for _ in range(2):
        m.insert_point_cloud(tiny, o0, 50.0)
        om.insert_point_cloud(tiny, o0, 50.0)
    cap0 = m.stats()["leaf_capacity"]
    dev = [ctx.to_device(p) for p, _ in scans]
    for d, (p, o) in zip(dev, scans):
        m.insert_point_cloud(d, o, 50.0, sync=False)
        om.insert_point_cloud(p, o, 50.0)
    assert m.sync() == 0
    st = m.stats()
    print(f"leaf capacity {cap0} -> {st['leaf_capacity']}, {st['leaf_count']} leaves", flush=True)
    assert st["leaf_count"] > cap0  # the async scans could not fit the pool they were given
    assert_same_cells(_gpu_cells(m), om.cells(), "async inserts after overflow")
    assert m.update_count == om.update_count
    assert m.digest() == om.digest()
    for d in dev:
        d.free()
    m.close()
    om.close()


def test_occupied_free_and_classify(ctx):
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    pts, origin = scenes.box_scan(seed=6, n=50_000)
    om = ora.ProbMap(0.05)
    om.insert_point_cloud(pts, origin, 10.0)
    m = bx.ProbabilisticMap(ctx, 0.05)
    m.insert_point_cloud(pts, origin, 10.0)
    from conftest import sorted_cells
    occ = sorted_cells(m.get_occupied_voxels())[0]
    free = sorted_cells(m.get_free_voxels())[0]
    assert np.array_equal(occ, om.occupied())
    assert np.array_equal(free, om.free())
    rng = np.random.default_rng(1)
    q = np.concatenate([occ[:1000], free[:1000], rng.integers(-120, 120, (2000, 3))]).astype(np.int32)
    assert np.array_equal(m.classify(q), om.classify(q))
    m.close()
    om.close()


def test_add_hit_miss_points(ctx):
    """addHitPoint / addMissPoint update the endpoint at once and trace the
    ray on the next insertPointCloud (probabilistic_map.hpp:95-103)."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    pts, origin = scenes.box_scan(seed=8, n=5_000)
    om = ora.ProbMap(0.05)
    m = bx.ProbabilisticMap(ctx, 0.05)
    for p, hit in (((1.0, 0.5, 0.2), True), ((-2.0, 1.0, 1.0), False), ((1.0, 0.5, 0.2), False)):
        (om.add_hit_point if hit else om.add_miss_point)(np.array(p))
        (m.add_hit_point if hit else m.add_miss_point)(p)
    assert_same_cells(_gpu_cells(m), om.cells(), "after add points")
    om.insert_point_cloud(pts, origin, 10.0)
    m.insert_point_cloud(pts, origin, 10.0)
    assert_same_cells(_gpu_cells(m), om.cells(), "after the next insert")
    m.close()
    om.close()


def test_batch_matches_sequential(ctx):
    """Eight uneven scans into eight maps in one batched pass equal eight
    independent oracle runs."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    scans = scenes.batch_scans(nmaps=8, lo=2_000, hi=40_000)
    maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    dev = [ctx.to_device(p) for p, _ in scans]
    bx.insert_batch(ctx, [(m, d, o, 30.0) for m, d, (_, o) in zip(maps, dev, scans)])
    for i, (m, (p, o)) in enumerate(zip(maps, scans)):
        assert m.sync() == 0
        om = ora.ProbMap(0.1)
        om.insert_point_cloud(p, o, 30.0)
        assert_same_cells(_gpu_cells(m), om.cells(), f"map {i}")
        om.close()
    for m in maps:
        m.close()
    for d in dev:
        d.free()


def test_batch_mixed_empty_and_filled_maps(ctx):
    """One batch whose maps alternate between empty ones (their rays take the
    fresh pass, whose winners end the list) and ones that already hold a
    scan (leaf and update passes), eager and through a captured graph, also
    counted: every map and its ray counts equal the oracle's."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    scans = scenes.batch_scans(nmaps=7, lo=3_000, hi=40_000, seed=21)
    prior = scenes.batch_scans(nmaps=7, lo=3_000, hi=20_000, seed=22)
    filled = [i % 2 == 1 for i in range(len(scans))]
    for use_graph in (False, True):
        maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
        models = [ora.ProbMap(0.1) for _ in scans]
        for m, om, f, (p, o) in zip(maps, models, filled, prior):
            if f:
                m.insert_point_cloud(p, o, 30.0)
                om.insert_point_cloud(p, o, 30.0)
        dev = [ctx.to_device(p) for p, _ in scans]
        descs = [(m, d, o, 30.0) for m, d, (_, o) in zip(maps, dev, scans)]
        ctx.counting(True)
        if use_graph:
            g = bx.ScanGraph(ctx, len(scans), sum(len(p) for p, _ in scans))
            g.set_scans(descs)
            g.launch()
            ctx.sync()
            g.close()
        else:
            bx.insert_batch(ctx, descs)
        ctx.counting(False)
        for i, (m, om, (p, o)) in enumerate(zip(maps, models, scans)):
            assert m.sync() == 0
            om.insert_point_cloud(p, o, 30.0)
            what = f"map {i} ({'filled' if filled[i] else 'empty'}, graph {use_graph})"
            assert_same_cells(_gpu_cells(m), om.cells(), what)
            st, ost = m.stats(), om.stats()
            assert st["last_rays"] == ost["rays"] and st["last_ray_updates"] == ost["ray_updates"], what
            assert st["last_ray_visits"] == ost["visits"], what
            assert m.update_count == om.update_count
            om.close()
            m.close()
        for d in dev:
            d.free()


def test_batch_of_more_scans_than_bucket_tags(ctx):
    """A wave of 70 scans (more than the 64 bucket tags, so the winner list is
    not split by pipeline and every walk skips the other pipeline's rays),
    empty and filled maps mixed: every map equals the oracle's."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    scans = scenes.batch_scans(nmaps=70, lo=1_000, hi=6_000, seed=23)
    maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    models = [ora.ProbMap(0.1) for _ in scans]
    for i in range(0, len(scans), 3):  # every third map holds a scan already
        p, o = scans[(i + 1) % len(scans)]
        maps[i].insert_point_cloud(p, o, 30.0)
        models[i].insert_point_cloud(p, o, 30.0)
    dev = [ctx.to_device(p) for p, _ in scans]
    bx.insert_batch(ctx, [(m, d, o, 30.0) for m, d, (_, o) in zip(maps, dev, scans)])
    for i, (m, om, (p, o)) in enumerate(zip(maps, models, scans)):
        assert m.sync() == 0
        om.insert_point_cloud(p, o, 30.0)
        assert_same_cells(_gpu_cells(m), om.cells(), f"map {i}")
        om.close()
        m.close()
    for d in dev:
        d.free()


@pytest.mark.parametrize("setting", [
    {"ray_block": 64, "ray_chunk": 64, "ray_blocks_per_cu": 16, "leaf_blocks_per_cu": 4,
     "small_ray_blocks_per_cu": 8},
    {"ray_block": 128, "ray_chunk": 512, "ep_table_factor": 4, "root_slack": 4, "ray_window": 1 << 20,
     "leaf_window": 1},
    {"ray_block": 256, "ray_chunk": 1024, "ray_blocks_per_cu": 1, "leaf_blocks_per_cu": 1,
     "small_ray_blocks_per_cu": 1, "ray_window": 1, "leaf_window": 1 << 20},
    {"fresh_pass": 0, "ray_order": 0},
    {"fresh_pass": 1, "ray_window": 1, "ray_block": 64, "fresh_blocks_per_cu": 1, "fresh_window": 1,
     "ray_order": 2},
    {"fresh_blocks_per_cu": 16, "fresh_window": 1 << 20, "ray_chunk": 64, "ray_order": 5},
    {"ray_order": 4, "fresh_pass": 0},
])
def test_tuning_settings_keep_results(setting):
    """Every launch setting the autotuner may pick (workgroup size, rays per
    lane, workgroups per CU, table load factors, lockstep windows, the fresh
    pass on or off for maps empty at a scan's start, the order of the rays)
    gives the same maps, in
    the batched call and in a captured graph."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    c = bx.Context(0)
    for k, v in setting.items():
        c.set_tuning(k, v)
        assert c.get_tuning(k) == v
    scans = scenes.batch_scans(nmaps=6, lo=2_000, hi=30_000, seed=13)
    maps = [bx.ProbabilisticMap(c, 0.1) for _ in scans]
    dev = [c.to_device(p) for p, _ in scans]
    descs = [(m, d, o, 30.0) for m, d, (_, o) in zip(maps, dev, scans)]
    bx.insert_batch(c, descs)
    g = bx.ScanGraph(c, len(scans), sum(len(p) for p, _ in scans))
    g.set_scans(descs)
    g.launch()
    c.sync()
    g.close()
    for i, (m, (p, o)) in enumerate(zip(maps, scans)):
        assert m.sync() == 0
        om = ora.ProbMap(0.1)
        om.insert_point_cloud(p, o, 30.0)
        om.insert_point_cloud(p, o, 30.0)
        assert_same_cells(_gpu_cells(m), om.cells(), f"map {i}")
        assert m.update_count == om.update_count
        om.close()
    for m in maps:
        m.close()
    for d in dev:
        d.free()
    c.close()


def test_batch_same_map_keeps_order(ctx):
    """Scans naming the same map in one batch run in input order."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    seq = list(scenes.lidar_sequence(3, nazim=128))
    m = bx.ProbabilisticMap(ctx, 0.1)
    other = bx.ProbabilisticMap(ctx, 0.1)
    dev = [ctx.to_device(p) for p, _ in seq]
    bx.insert_batch(ctx, [(m, dev[0], seq[0][1], 50.0), (other, dev[1], seq[1][1], 50.0),
                          (m, dev[1], seq[1][1], 50.0), (m, dev[2], seq[2][1], 50.0)])
    assert m.sync() == 0
    om = ora.ProbMap(0.1)
    for p, o in seq:
        om.insert_point_cloud(p, o, 50.0)
    assert_same_cells(_gpu_cells(m), om.cells(), "same-map batch")
    assert m.update_count == om.update_count == 1
    for x in (m, other):
        x.close()
    for d in dev:
        d.free()


def test_graph_frames_replay_after_overflow(ctx):
    """Two launches of a captured frame whose scans overflow their maps'
    pools (two tiny scans first teach each map a small growth): confirming
    the maps re-runs the skipped scans of both frames in order, and every map
    equals the oracle's."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    ora = _oracle()
    scans = scenes.batch_scans(nmaps=4, lo=40_000, hi=120_000, seed=17)
    maps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    models = [ora.ProbMap(0.1) for _ in scans]
    for m, om, (p, o) in zip(maps, models, scans):
        for _ in range(2):
            m.insert_point_cloud(p[:300], o, 30.0)
            om.insert_point_cloud(p[:300], o, 30.0)
    caps = [m.stats()["leaf_capacity"] for m in maps]
    dev = [ctx.to_device(p) for p, _ in scans]
    g = bx.ScanGraph(ctx, 4, sum(len(p) for p, _ in scans))
    g.set_scans([(m, d, o, 30.0) for m, d, (_, o) in zip(maps, dev, scans)])
    for _ in range(2):
        g.launch()
        for om, (p, o) in zip(models, scans):
            om.insert_point_cloud(p, o, 30.0)
    for i, (m, om) in enumerate(zip(maps, models)):
        assert m.sync() == 0
        assert m.stats()["leaf_count"] > caps[i]  # the frame did not fit the pools it was given
        assert_same_cells(_gpu_cells(m), om.cells(), f"map {i}")
        assert m.update_count == om.update_count
    g.close()
    for m, om in zip(maps, models):
        m.close()
        om.close()
    for d in dev:
        d.free()


def test_graph_replays_equal_eager(ctx):
    """A captured frame graph replayed four times (with eager work and
    synchronisation between replays) equals four eager batched inserts."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    scans = scenes.batch_scans(nmaps=6, lo=2_000, hi=20_000)
    dev = [ctx.to_device(p) for p, _ in scans]
    gmaps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    emaps = [bx.ProbabilisticMap(ctx, 0.1) for _ in scans]
    g = bx.ScanGraph(ctx, 8, sum(len(p) for p, _ in scans))
    g.set_scans([(m, d, o, 30.0) for m, d, (_, o) in zip(gmaps, dev, scans)])
    for rep in range(4):
        g.launch()
        bx.insert_batch(ctx, [(m, d, o, 30.0) for m, d, (_, o) in zip(emaps, dev, scans)])
        ctx.sync()
        for i, (a, b) in enumerate(zip(gmaps, emaps)):
            assert a.sync() == 0, f"replay {rep} map {i} reported a device error"
            assert a.digest() == b.digest(), f"replay {rep} map {i}"
            assert a.update_count == b.update_count
    g.close()
    for m in gmaps + emaps:
        m.close()
    for d in dev:
        d.free()
