"""The serving engine (bxg_server_*): a mixed stream of inserts (depth
frames), lookups and occupied/free exports over resident maps, submitted
while earlier requests run.  Every result must equal the same request run
alone: each map's requests replayed one at a time, in order, through the
plain API.  The final maps are also checked against the oracle."""
import random
import time

import numpy as np
import pytest

from conftest import assert_same_cells, sorted_cells

pytestmark = pytest.mark.gpu

RES = 0.05
RANGE = 5.0


def make_requests(nmaps, n_per_map, seed):
    """A submission order over nmaps maps: ('insert', map, points, origin) /
    ('lookup', map, coords) / ('export', map, filter)."""
    from bonxai_amd import scenes
    rng = np.random.default_rng(seed)
    reqs = []
    frame = 0
    for k in range(n_per_map * nmaps):
        m = int(rng.integers(nmaps))
        kind = rng.choice(["insert", "insert", "lookup", "export"])
        if kind == "insert":
            pts, origin = scenes.depth_frame(100 + frame)
            frame += 1
            reqs.append(("insert", m, pts[::int(rng.integers(1, 5))].copy(), origin))
        elif kind == "lookup":
            c = rng.integers(-120, 120, (int(rng.integers(1_000, 60_000)), 3)).astype(np.int32)
            reqs.append(("lookup", m, c))
        else:
            reqs.append(("export", m, int(rng.choice([1, 2]))))
    return reqs


def test_serving_results_equal_alone(ctx):
    import bonxai_amd as bx
    import oracle
    nmaps = 4
    reqs = make_requests(nmaps, 6, seed=3)
    maps = [bx.ProbabilisticMap(ctx, RES) for _ in range(nmaps)]
    dev = []
    for r in reqs:  # inputs and outputs in HBM before the stream starts
        if r[0] == "insert":
            dev.append((ctx.to_device(r[2]),))
        elif r[0] == "lookup":
            dev.append((ctx.to_device(r[2]), ctx.array(len(r[2]), np.uint32), ctx.array(len(r[2]), np.uint8)))
        else:
            dev.append((ctx.array((4_000_000, 3), np.int32), ctx.array(4_000_000, np.uint32)))
    ctx.sync()
    srv = bx.Server(ctx, max_inflight=8, max_batch_scans=4, streams=2)
    pyrng = random.Random(7)
    ids = []
    for r, d in zip(reqs, dev):
        m = maps[r[1]]
        if r[0] == "insert":
            ids.append(srv.insert(m, d[0], r[3], RANGE))
        elif r[0] == "lookup":
            ids.append(srv.lookup(m, d[0], d[1], d[2]))
        else:
            ids.append(srv.export(m, d[0], d[1], filter=r[2]))
        time.sleep(pyrng.random() * 0.002)  # arrivals while earlier requests run
    infos = [srv.wait(i) for i in ids]
    stats = srv.stats()
    srv.close()
    assert all(i["status"] == 0 for i in infos), infos
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    print("server stats", stats, flush=True)

    # the same requests alone, one at a time, map by map, on fresh maps
    alone = [bx.ProbabilisticMap(ctx, RES) for _ in range(nmaps)]
    models = [oracle.ProbMap(RES) for _ in range(nmaps)]
    for k, (r, d) in enumerate(zip(reqs, dev)):
        m, om = alone[r[1]], models[r[1]]
        if r[0] == "insert":
            m.insert_point_cloud(r[2], r[3], RANGE)
            om.insert_point_cloud(r[2], r[3], RANGE)
        elif r[0] == "lookup":
            vals, found = m.lookup(r[2])
            got_v, got_f = d[1].numpy(), d[2].numpy().astype(bool)
            assert np.array_equal(got_f, found), f"request {k}: lookup found flags"
            assert np.array_equal(got_v, vals), f"request {k}: lookup values"
            ov, of = om.grid.lookup(r[2])
            assert np.array_equal(of, found) and np.array_equal(ov[of], vals[found])
        else:
            n = infos[k]["count"]
            got = sorted_cells(d[0].numpy(n), d[1].numpy(n))
            want = sorted_cells(*m.export(r[2]))
            assert np.array_equal(got[0], want[0]), f"request {k}: export coords"
            assert np.array_equal(got[1], want[1]), f"request {k}: export values"
            assert np.array_equal(want[0], om.grid.export(r[2], 0)[0])
    for i in range(nmaps):
        assert_same_cells(maps[i].export(), sorted_cells(*alone[i].export()), f"map {i} final")
        assert_same_cells(maps[i].export(), models[i].cells(), f"map {i} vs oracle")
        assert maps[i].update_count == alone[i].update_count == models[i].update_count
        models[i].close()
    for d in dev:
        for a in d:
            a.free()
    for m in maps + alone:
        m.close()


def test_serving_budget_refusal():
    """An insert whose pools cannot grow within the budget, with nothing in
    flight to wait for, fails with BXG_ERR_BUDGET; the engine keeps serving."""
    import bonxai_amd as bx
    from bonxai_amd import scenes
    c = bx.Context(0, budget_bytes=48 << 20)
    m = bx.ProbabilisticMap(c, RES)
    pts, origin = scenes.depth_frame(5)
    dp = c.to_device(pts)
    q = c.to_device(np.zeros((10, 3), np.int32))
    qv, qf = c.array(10, np.uint32), c.array(10, np.uint8)
    srv = bx.Server(c)
    a = srv.wait(srv.insert(m, dp, origin, 50.0))  # a 50 m reach needs far more than 48 MiB
    b = srv.wait(srv.lookup(m, q, qv, qf))
    srv.close()
    assert a["status"] == bx.ERR_BUDGET
    assert b["status"] == 0 and not qf.numpy().any()
    for x in (dp, q, qv, qf):
        x.free()
    m.close()
    c.close()
