"""The reference server's ingest pre-step fused into the scan
(bonxai_server.cpp:147-179): non-finite points dropped, the rest moved by a
float sensor-to-world matrix, origin = the matrix's translation, then
insertPointCloud.  Compared with the oracle's restatement bit-exactly; the
float association of the transform is the SSE one of PCL (parity unpinned:
PCL is absent), and the points whose endpoint would differ under the scalar
association are counted and printed."""
import numpy as np
import pytest

from conftest import assert_same_cells, sorted_cells

pytestmark = pytest.mark.gpu


def rigid(rng):
    """A random float rotation + translation (row-major 4x4)."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    t = np.eye(4, dtype=np.float32)
    t[:3, :3] = r
    t[:3, 3] = rng.uniform(-20, 20, 3)
    return t


def sensor_cloud(rng, n, seed):
    """A depth frame moved into its own sensor frame, as PointXYZ (x y z pad),
    with NaN and Inf entries."""
    from bonxai_amd import scenes
    pts, origin = scenes.depth_frame(seed)
    pts = pts[:n]
    out = np.zeros((len(pts), 4), np.float32)
    out[:, :3] = pts - origin.astype(np.float32)
    bad = rng.choice(len(out), len(out) // 50, replace=False)
    out[bad[::3], 0] = np.nan
    out[bad[1::3], 1] = np.inf
    out[bad[2::3], 2] = -np.inf
    return out


def test_ingest_host_matches_oracle(ctx):
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(4)
    m = bx.ProbabilisticMap(ctx, 0.05)
    om = oracle.ProbMap(0.05)
    amb = 0
    for k in range(3):
        cloud = sensor_cloud(rng, 120_000, 40 + k)
        tf = rigid(rng)
        amb += oracle.tf_ambiguity(cloud, tf, 5.0, 0.05)
        m.insert_point_cloud_tf(cloud, tf, 5.0)
        om.insert_point_cloud_tf(cloud, tf, 5.0)
        assert_same_cells(m.cells(), om.cells(), f"frame {k}")
    print(f"points whose endpoint depends on the transform's association: {amb}", flush=True)
    # these frames hold points whose endpoint voxel differs between the two
    # float associations (1, 3 and 5 of them), so the equal maps show the GPU
    # computes the SSE association, not the scalar one
    assert amb > 0
    assert m.update_count == om.update_count
    m.close()
    om.close()


def test_ingest_device_batch_matches_oracle(ctx):
    """Device PointXYZ points (16-byte stride), several maps in one batched
    pass, each with its own matrix, next to a plain scan."""
    import bonxai_amd as bx
    import oracle
    from bonxai_amd import scenes
    rng = np.random.default_rng(6)
    maps = [bx.ProbabilisticMap(ctx, 0.05) for _ in range(4)]
    models = [oracle.ProbMap(0.05) for _ in range(4)]
    descs, keep = [], []
    for i in range(3):
        cloud = sensor_cloud(rng, 60_000, 60 + i)
        tf = rigid(rng)
        d = ctx.to_device(cloud)
        keep.append(d)
        descs.append((maps[i], d, None, 5.0, tf))
        models[i].insert_point_cloud_tf(cloud, tf, 5.0)
        m = maps[i]
        m.insert_point_cloud_tf(d, tf, 5.0)  # device form, alone
        models[i].insert_point_cloud_tf(cloud, tf, 5.0)
    pts, origin = scenes.box_scan(n=30_000, seed=9)
    d = ctx.to_device(pts)
    keep.append(d)
    descs.append((maps[3], d, origin, 10.0))
    models[3].insert_point_cloud(pts, origin, 10.0)
    bx.insert_batch(ctx, descs)
    for i, (m, om) in enumerate(zip(maps, models)):
        assert m.sync() == 0
        assert_same_cells(m.cells(), om.cells(), f"map {i}")
        m.close()
        om.close()
    for d in keep:
        d.free()


def test_ingest_identity_equals_plain_insert(ctx):
    """With the identity rotation the pre-step is the finite filter alone: the
    same map as a plain insert of the finite points from the translation."""
    import bonxai_amd as bx
    rng = np.random.default_rng(8)
    cloud = sensor_cloud(rng, 80_000, 77)
    tf = np.eye(4, dtype=np.float32)
    tf[:3, 3] = [1.5, -2.25, 0.75]
    a = bx.ProbabilisticMap(ctx, 0.05)
    a.insert_point_cloud_tf(cloud, tf, 5.0)
    fin = np.isfinite(cloud[:, :3]).all(1)
    moved = cloud[fin, :3] + tf[:3, 3]  # x*1 + (0 + (0 + t)) == x + t in float
    b = bx.ProbabilisticMap(ctx, 0.05)
    b.insert_point_cloud(moved.astype(np.float32), tf[:3, 3].astype(np.float64), 5.0)
    assert_same_cells(a.cells(), sorted_cells(*b.cells()), "identity transform")
    a.close()
    b.close()


def test_ingest_graph_frames_match_oracle(ctx):
    """Transformed scans (PointXYZ stride, a matrix per map) in a captured
    frame graph launched twice: the library keeps its own copy of each matrix
    (the arrays it was given are zeroed after set_scans), and every map
    equals the oracle's after both frames."""
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(12)
    maps = [bx.ProbabilisticMap(ctx, 0.05) for _ in range(3)]
    models = [oracle.ProbMap(0.05) for _ in range(3)]
    clouds, tfs, dev = [], [], []
    for i in range(3):
        cloud = sensor_cloud(rng, 40_000, 80 + i)
        clouds.append(cloud)
        tfs.append(rigid(rng))
        dev.append(ctx.to_device(cloud))
    g = bx.ScanGraph(ctx, 3, sum(len(c) for c in clouds))
    g.set_scans([(m, d, None, 5.0, tf.copy()) for m, d, tf in zip(maps, dev, tfs)])
    for k in g._keep:  # the matrices handed over: their contents no longer matter
        if isinstance(k, np.ndarray) and k.size == 16:
            k[:] = 0
    for _ in range(2):
        g.launch()
        for om, cloud, tf in zip(models, clouds, tfs):
            om.insert_point_cloud_tf(cloud, tf, 5.0)
    for i, (m, om) in enumerate(zip(maps, models)):
        assert m.sync() == 0
        assert_same_cells(m.cells(), om.cells(), f"map {i}")
        assert m.update_count == om.update_count
    g.close()
    for m, om, d in zip(maps, models, dev):
        m.close()
        om.close()
        d.free()


def lidar_sensor_cloud(rng, pts, origin):
    """A LiDAR scan moved into its own sensor frame, as PointXYZ, with a few
    NaN entries."""
    out = np.zeros((len(pts), 4), np.float32)
    out[:, :3] = pts - origin.astype(np.float32)
    bad = rng.choice(len(out), len(out) // 100, replace=False)
    out[bad, 1] = np.nan
    return out


def test_async_tf_inserts_replay_after_overflow(ctx):
    """Asynchronous transformed inserts of device points whose head-room
    estimate is far too small (two tiny scans first teach it a small growth;
    each LiDAR scan, an eighth of its beams at 20 m, then needs about 18k
    leaves, twice the pool they leave): the scans that
    overflowed the node pool run again when the map is confirmed, with the
    matrices they were given.  Each matrix array is overwritten with NaN
    right after its call returns, so a replay that read the caller's memory
    instead of the library's copy would move every point off the map."""
    import bonxai_amd as bx
    import oracle
    from bonxai_amd import scenes
    rng = np.random.default_rng(14)
    scans = list(scenes.lidar_sequence(4))
    m = bx.ProbabilisticMap(ctx, 0.1)
    om = oracle.ProbMap(0.1)
    tiny = lidar_sensor_cloud(rng, *scans[0])[:300]
    tf0 = rigid(rng)
    for _ in range(2):
        m.insert_point_cloud_tf(tiny, tf0, 20.0)
        om.insert_point_cloud_tf(tiny, tf0, 20.0)
    cap0 = m.stats()["leaf_capacity"]
    dev = []
    for pts, origin in scans:
        cloud = lidar_sensor_cloud(rng, pts[::8], origin)
        tf = rigid(rng)
        d = ctx.to_device(cloud)
        dev.append(d)
        handed = tf.copy()
        m.insert_point_cloud_tf(d, handed, 20.0, sync=False)
        handed[:] = np.nan
        om.insert_point_cloud_tf(cloud, tf, 20.0)
    assert m.sync() == 0
    st = m.stats()
    print(f"leaf capacity {cap0} -> {st['leaf_capacity']}, replays {st['replays']}", flush=True)
    assert st["leaf_count"] > cap0 and st["replays"] > 0
    assert_same_cells(m.cells(), om.cells(), "async transformed inserts after overflow")
    assert m.update_count == om.update_count
    for d in dev:
        d.free()
    m.close()
    om.close()
