"""CPU checks that pin the oracle (the CPU restatement in oracle/bx_oracle.c)
before it is trusted as the parity checker of the GPU path.

Pins, by part of the reference path:
  posToCoord (bonxai.hpp:404-410)        closed form floor(x * (1/res)) in
                                          numpy float64; the reference's own
                                          posToCoord where oracle/_ref exists
  RayIterator (probabilistic_map.hpp:162-203)
                                          the closed form
                                          o + sgn(d) * floor((2k|d| + m) / 2m),
                                          exhaustively, plus invariants
  logods defaults (probabilistic_map.hpp:34-36,56-64)
                                          closed form
  insertPointCloud (probabilistic_map.hpp:141-160,
  probabilistic_map.cpp:30-106)          an independent order-free model
                                          (first-index winner per voxel, set
                                          of ray cells) on tiny scenes; the
                                          hand-derived known-answer scenes;
                                          fixtures written by scan integration
                                          over the reference's own VoxelGrid
  VoxelGrid ops (bonxai.hpp:449-743)     the reference's own VoxelGrid
                                          (fixtures, and live where built)
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from bonxai_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEFAULTS = dict(miss=-405465, hit=847297, cmin=-1992430, cmax=3476099, thr=0)


# ------------------------------------------------------------ closed forms

def closed_form_ray(o, e):
    o = np.asarray(o, np.int64)
    d = np.asarray(e, np.int64) - o
    m = int(np.abs(d).max())
    if m == 0:
        return np.zeros((0, 3), np.int64)
    sgn = np.where(d < 0, -1, 1)
    k = np.arange(m)[:, None]
    return o + sgn * ((2 * k * np.abs(d) + m) // (2 * m))


def test_pos_to_coord_closed_form():
    rng = np.random.default_rng(0)
    for res in (0.05, 0.1, 0.01, 0.2, 0.005):
        inv = 1.0 / res
        k = rng.integers(-10**6, 10**6, 20000).astype(np.float64)
        x = np.concatenate([k * res, np.nextafter(k * res, np.inf), np.nextafter(k * res, -np.inf),
                            rng.normal(0, 100, 20000), [0.0, -0.0, 1e-300, -1e-300]])
        x3 = np.ascontiguousarray(x[: (x.size // 3) * 3])
        out = np.empty(x3.size, np.int32)
        oracle.lib().bxo_pos_to_coord(oracle._p(x3), x3.size // 3, inv, oracle._p(out))
        assert np.array_equal(out, np.floor(x3 * inv).astype(np.int32))
        if oracle.ref_available():
            ref = np.empty(x3.size, np.int32)
            oracle.ref().ref_pos_to_coord(res, oracle._p(x3), x3.size // 3, oracle._p(ref))
            assert np.array_equal(out, ref)


def test_ray_matches_closed_form_exhaustive():
    rng = np.random.default_rng(1)
    r = np.arange(-12, 13)
    deltas = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    for d in deltas:
        o = rng.integers(-100, 100, 3)
        ray = oracle.compute_ray(o, o + d)
        exp = closed_form_ray(o, o + d)
        assert np.array_equal(ray, exp), (o, d)
        m = int(np.abs(d).max())
        assert len(ray) == m
        if m:
            assert np.array_equal(ray[0], o)
            assert not np.any(np.all(ray == o + d, axis=1))
            steps = np.abs(np.diff(ray, axis=0))
            assert steps.max(initial=0) <= 1
    for _ in range(2000):  # long rays
        o = rng.integers(-10**5, 10**5, 3)
        e = o + rng.integers(-700, 700, 3)
        assert np.array_equal(oracle.compute_ray(o, e), closed_form_ray(o, e))


def test_logods_defaults_closed_form():
    def logods(p):
        p = float(np.float32(p))
        return int(1e6 * math.log(p / (1.0 - p)))
    assert logods(0.4) == DEFAULTS["miss"]
    assert logods(0.7) == DEFAULTS["hit"]
    assert logods(0.12) == DEFAULTS["cmin"]
    assert logods(0.97) == DEFAULTS["cmax"]
    assert logods(0.5) == DEFAULTS["thr"] == 0
    import bonxai_amd as bx
    assert bx.ProbabilisticMap.logods(0.7) == DEFAULTS["hit"]


# ------------------------------------------------ independent scan model

class SetModel:
    """Order-free statement of one scan (first-index winner per voxel,
    endpoint updates before the set of ray cells), in plain Python."""

    def __init__(self, res):
        self.res = res
        self.inv = 1.0 / res
        self.cells = {}
        self.cur = 1

    def insert(self, pts, origin, R):
        F = [float(v) for v in origin]
        first = {}
        for P in pts:
            T = [float(v) for v in P]
            v = [T[a] - F[a] for a in range(3)]
            sq = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
            if sq >= R * R:
                s = math.sqrt(sq)
                q = [F[a] + ((v[a] / s) * R) for a in range(3)]
                hit = False
            else:
                q, hit = T, True
            e = tuple(math.floor(q[a] * self.inv) for a in range(3))
            first.setdefault(e, hit)
        marked = [e for e in first if self.cells.get(e, (0, 0))[1] != self.cur]
        for e in marked:
            p, _ = self.cells.get(e, (0, 0))
            p = min(p + DEFAULTS["hit"], DEFAULTS["cmax"]) if first[e] else max(p + DEFAULTS["miss"], DEFAULTS["cmin"])
            self.cells[e] = (p, self.cur)
        o = tuple(math.floor(F[a] * self.inv) for a in range(3))
        ray_cells = set()
        for e in marked:
            for c in closed_form_ray(o, e):
                ray_cells.add(tuple(int(x) for x in c))
        for c in ray_cells:
            p, i = self.cells.get(c, (0, 0))
            if i != self.cur:
                self.cells[c] = (max(p + DEFAULTS["miss"], DEFAULTS["cmin"]), self.cur)
        self.cur = self.cur % 3 + 1

    def dump(self):
        keys = sorted(self.cells)
        coords = np.array(keys, np.int32).reshape(-1, 3)
        words = np.array([((self.cells[k][0] & 0xFFFFFFF) << 4) | (self.cells[k][1] & 15) for k in keys],
                         np.uint32)
        return coords, words


def test_oracle_matches_order_free_model():
    rng = np.random.default_rng(7)
    for trial in range(25):
        res = [0.1, 0.25, 0.05][trial % 3]
        om = oracle.ProbMap(res)
        sm = SetModel(res)
        for scan in range(1 + trial % 5):
            n = int(rng.integers(0, 60))
            origin = rng.normal(0, 0.5, 3)
            pts = (origin + rng.normal(0, 1.5, (n, 3))).astype(np.float32)
            if n > 4:
                pts[: n // 4] = pts[n // 4: 2 * (n // 4)]  # duplicates, both kinds
            R = float(rng.choice([0.8, 1.5, 3.0, 100.0]))
            om.insert_point_cloud(pts, origin, R)
            sm.insert(pts, origin, R)
            oc, ow = om.cells()
            mc, mw = sm.dump()
            assert np.array_equal(oc, mc), (trial, scan)
            assert np.array_equal(ow, mw), (trial, scan)
            assert om.update_count == sm.cur
        om.close()


class LeafPhasedModel(SetModel):
    """The phase order of the GPU pipeline, in plain Python: endpoints in
    leaves that existed before the scan are updated first; then the ray cells
    in random order, a cell of a leaf the scan created being stored as the
    miss update of zero without being read (an old leaf's cell is read and
    skipped when its update_id is cur); then the endpoints in the new leaves
    are written as the endpoint update of zero over whatever the rays stored;
    a new leaf's on cells are its nonzero words.  Leaves of 8^3 cells."""

    def __init__(self, res, seed):
        super().__init__(res)
        self.leaves = set()
        self.rng = np.random.default_rng(seed)

    @staticmethod
    def leaf(c):
        return (c[0] >> 3, c[1] >> 3, c[2] >> 3)

    def insert(self, pts, origin, R):
        F = [float(v) for v in origin]
        first = {}
        for P in pts:
            T = [float(v) for v in P]
            v = [T[a] - F[a] for a in range(3)]
            sq = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
            if sq >= R * R:
                s = math.sqrt(sq)
                q = [F[a] + ((v[a] / s) * R) for a in range(3)]
                hit = False
            else:
                q, hit = T, True
            e = tuple(math.floor(q[a] * self.inv) for a in range(3))
            first.setdefault(e, hit)
        old = set(self.leaves)
        marked = [e for e in first if self.cells.get(e, (0, 0))[1] != self.cur]
        o = tuple(math.floor(F[a] * self.inv) for a in range(3))
        rays = [[tuple(int(x) for x in c) for c in closed_form_ray(o, e)] for e in marked]
        # every endpoint's leaf and every ray cell's leaf exist before any cell is written
        self.leaves.update(self.leaf(e) for e in first)
        self.leaves.update(self.leaf(c) for r in rays for c in r)

        def upd(p, hit):
            return min(p + DEFAULTS["hit"], DEFAULTS["cmax"]) if hit else max(p + DEFAULTS["miss"], DEFAULTS["cmin"])

        for e in marked:
            if self.leaf(e) in old:
                p, _ = self.cells.get(e, (0, 0))
                self.cells[e] = (upd(p, first[e]), self.cur)
        fresh = (max(DEFAULTS["miss"], DEFAULTS["cmin"]), self.cur)
        for k in self.rng.permutation(len(rays)):
            for c in rays[k]:
                if self.leaf(c) not in old:
                    self.cells[c] = fresh
                else:
                    p, i = self.cells.get(c, (0, 0))
                    if i != self.cur:
                        self.cells[c] = (upd(p, False), self.cur)
        for e in marked:
            if self.leaf(e) not in old:
                self.cells[e] = (upd(0, first[e]), self.cur)
        self.cur = self.cur % 3 + 1


def test_new_leaf_phase_order_matches_oracle():
    """The reordered pipeline (old endpoints, rays with blind stores in new
    leaves, then new endpoints) gives the oracle's cells, on sequences whose
    scans mix new and old leaves, duplicate voxels and clipped points."""
    rng = np.random.default_rng(23)
    for trial in range(30):
        res = [0.1, 0.25, 0.05][trial % 3]
        om = oracle.ProbMap(res)
        lm = LeafPhasedModel(res, trial)
        origin = rng.normal(0, 0.5, 3)
        for scan in range(1 + trial % 6):
            n = int(rng.integers(0, 80))
            origin = origin + rng.normal(0, 0.4, 3)
            pts = (origin + rng.normal(0, 1.5, (n, 3))).astype(np.float32)
            if n > 4:
                pts[: n // 4] = pts[n // 4: 2 * (n // 4)]
            R = float(rng.choice([0.8, 1.5, 3.0, 100.0]))
            om.insert_point_cloud(pts, origin, R)
            lm.insert(pts, origin, R)
            oc, ow = om.cells()
            mc, mw = lm.dump()
            assert np.array_equal(oc, mc), (trial, scan)
            assert np.array_equal(ow, mw), (trial, scan)
        om.close()


def test_kat_by_hand():
    """Known answers derived by hand from probabilistic_map.cpp:30-106."""
    hit, miss = DEFAULTS["hit"], DEFAULTS["miss"]
    z = np.zeros(3)
    # first point of a voxel decides hit vs miss
    a = np.array([[1.04, 0.0, 0.0]], np.float32)
    b = np.array([[5.0, 0.0, 0.0]], np.float32)
    for pts, p in ((np.concatenate([a, b]), hit), (np.concatenate([b, a]), miss)):
        m = oracle.ProbMap(0.1)
        m.insert_point_cloud(pts, z, 1.05)
        c, w = m.cells()
        assert c.tolist() == [[k, 0, 0] for k in range(11)]
        assert (w[-1].astype(np.int32) >> 4) == p and (w[-1] & 15) == 1
        assert np.all((w[:-1].astype(np.int32) >> 4) == miss)
        m.close()
    # frozen endpoint: touched in scan 1 (id 1), untouched in scans 2 and 3,
    # endpoint again in scan 4 (id 1 again): no update and no ray
    m = oracle.ProbMap(0.1)
    p0 = np.array([[0.55, 0.0, 0.0]], np.float32)
    far = np.array([[0.0, 3.05, 0.0]], np.float32)
    m.insert_point_cloud(p0, z, 5.0)
    m.insert_point_cloud(far, z, 5.0)
    m.insert_point_cloud(far, z, 5.0)
    m.insert_point_cloud(np.concatenate([p0, far]), z, 5.0)
    c, w = m.cells()
    cells = {tuple(k): (int(v.astype(np.int32)) >> 4, int(v) & 15) for k, v in zip(c, w)}
    assert cells[(5, 0, 0)] == (hit, 1)
    for k in range(1, 5):
        assert cells[(k, 0, 0)] == (miss, 1)
    assert cells[(0, 30, 0)][0] == min(3 * hit, DEFAULTS["cmax"])
    assert m.update_count == 2
    m.close()
    # empty scan advances the counter; an origin-voxel point has no ray
    m = oracle.ProbMap(0.1)
    m.insert_point_cloud(np.zeros((0, 3), np.float32), z, 1.0)
    assert m.update_count == 2
    m.insert_point_cloud(np.array([[0.05, 0.05, 0.05]], np.float32), z, 1.0)
    c, w = m.cells()
    assert c.tolist() == [[0, 0, 0]] and int(w[0]) == ((hit << 4) | 2)
    m.close()


# ------------------------------------------------------------- fixtures

def _kat_scenarios():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.kat_scenarios()


def test_kat_fixtures_reproduced():
    kat = np.load(os.path.join(GOLDEN, "kat.npz"))
    for name, (scans, res) in _kat_scenarios().items():
        m = oracle.ProbMap(res)
        for pts, origin, R in scans:
            m.insert_point_cloud(pts, origin, R)
        c, w = m.cells()
        assert np.array_equal(c, kat[f"{name}_coords"]), name
        assert np.array_equal(w, kat[f"{name}_words"]), name
        assert m.update_count == int(kat[f"{name}_cur"][0]), name
        m.close()


def test_grid_fixture_reproduced():
    fx = np.load(os.path.join(GOLDEN, "grid.npz"))
    coords = scenes.clusters(n=100_000)
    vals = scenes.cluster_values(coords)
    g = oracle.Grid(0.05, dtype=np.float32)
    assert np.array_equal(g.set_values(coords, vals), fx["was_on"])
    v, f = g.lookup(scenes.lookups(coords, n=200_000))
    assert np.array_equal(v, fx["lookup_vals"]) and np.array_equal(f, fx["lookup_found"])
    c, cv = g.export()
    assert np.array_equal(c, fx["cells"]) and np.array_equal(cv, fx["cell_vals"])
    g.close()


def test_input_digests():
    d = json.load(open(os.path.join(GOLDEN, "digests.json")))["inputs"]
    assert scenes.fnv1a(scenes.box_scan()[0]) == d["box_scan"]
    c = scenes.clusters()
    assert scenes.fnv1a(c) == d["clusters"]
    assert scenes.fnv1a(scenes.lookups(c)) == d["lookups"]
    for k, (pts, _) in enumerate(scenes.lidar_sequence(3)):
        assert scenes.fnv1a(pts) == d[f"lidar_{k}"]
    assert [int(x) for x in scenes.batch_sizes()] == d["batch_sizes"]
    for i, (pts, _) in enumerate(scenes.batch_scans()[:4]):
        assert scenes.fnv1a(pts) == d[f"batch_{i}"]
    for s in range(2):
        assert scenes.fnv1a(scenes.depth_frame(s)[0]) == d[f"depth_{s}"]


# SHA-256 of the batched benchmark frame's inputs (64 scans, points and
# origins), the value bench.py reports as config.input_sha256
BATCH_FRAME_SHA256 = "5de68a0120b80bb244587663ae876a9adfdaee3f362d4998564efc90f122cfc6"


def test_batch_frame_sha256():
    assert scenes.digest_scans(scenes.batch_scans()) == BATCH_FRAME_SHA256


def test_box_scan_digest():
    d = json.load(open(os.path.join(GOLDEN, "digests.json")))["maps"]["box_scan"]
    pts, origin = scenes.box_scan()
    m = oracle.ProbMap(0.05)
    m.insert_point_cloud(pts, origin, 10.0)
    assert list(map(str, m.digest())) == d["digest"]
    assert m.stats() == d["stats"]
    m.close()


def test_squared_norm_association_is_immaterial():
    """The endpoint of no point of the workloads changes under the other
    association of Eigen's squaredNorm (x*x + (y*y + z*z)): every point of the
    box scan, of all 200 LiDAR scans of the sequence, of all 64 scans of the
    batched frame, and of the depth frames the serving benchmark and tests
    insert (seeds 0..3, 100..139 and 1000..1039)."""
    pts, origin = scenes.box_scan()
    assert oracle.endpoints(pts, origin, 10.0, 0.05)[2] == 0
    for k, (pts, origin) in enumerate(scenes.lidar_sequence(200)):
        assert oracle.endpoints(pts, origin, 50.0, 0.1)[2] == 0, f"lidar scan {k}"
    for k, (pts, origin) in enumerate(scenes.batch_scans()):
        assert oracle.endpoints(pts, origin, 30.0, 0.1)[2] == 0, f"batch scan {k}"
    for seed in [*range(4), *range(100, 140), *range(1000, 1040)]:
        pts, origin = scenes.depth_frame(seed)
        assert oracle.endpoints(pts, origin, 5.0, 0.05)[2] == 0, f"depth frame {seed}"


# ----------------------------------------- the reference's own VoxelGrid

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.int64, None])
def test_grid_ops_match_reference(dtype):
    rng = np.random.default_rng(11)
    r = oracle.RefGrid(0.1, dtype)
    g = oracle.Grid(0.1, dtype=dtype)
    for step in range(6):
        a = rng.integers(-70, 70, (4000, 3)).astype(np.int32)
        v = (rng.integers(-1000, 1000, 4000)).astype(dtype) if dtype is not None else None
        op = step % 3
        if op == 0:
            assert np.array_equal(r.set_values(a, v), g.set_values(a, v))
        elif op == 1:
            assert np.array_equal(r.set_cells_off(a), g.set_cells_off(a))
        else:
            assert np.array_equal(r.set_cells_on(a, v), g.set_cells_on(a, v))
        q = rng.integers(-80, 80, (4000, 3)).astype(np.int32)
        rv, rf = r.lookup(q)
        gv, gf = g.lookup(q)
        assert np.array_equal(rf, gf)
        if dtype is not None:
            assert np.array_equal(rv, gv)
        rc, rcv = r.export()
        gc, gcv = g.export()
        assert np.array_equal(rc, gc)
        if dtype is not None:
            assert np.array_equal(rcv, gcv)
        assert r.active_cells_count() == g.active_cells_count()
    r.clear(1)
    g.clear(1)
    assert r.active_cells_count() == g.active_cells_count() == 0
    r.close()
    g.close()


@needs_ref
def test_scan_integration_over_reference_storage():
    """Scan integration over the reference's VoxelGrid<CellT> equals the
    restatement on the box scan and three LiDAR scans."""
    pts, origin = scenes.box_scan(seed=12, n=30_000)
    rp, om = oracle.RefProb(0.05), oracle.ProbMap(0.05)
    rp.insert_point_cloud(pts, origin, 10.0)
    om.insert_point_cloud(pts, origin, 10.0)
    assert all(np.array_equal(a, b) for a, b in zip(rp.cells(), om.cells()))
    rp, om = oracle.RefProb(0.1), oracle.ProbMap(0.1)
    for pts, origin in scenes.lidar_sequence(3, nazim=256):
        rp.insert_point_cloud(pts, origin, 50.0)
        om.insert_point_cloud(pts, origin, 50.0)
    assert all(np.array_equal(a, b) for a, b in zip(rp.cells(), om.cells()))
    assert rp.update_count == om.update_count


@needs_ref
def test_cpu_restatement_speed_against_reference_storage(capsys):
    """The CPU baseline is the restatement; this relates its speed to the
    reference's own VoxelGrid storage (unordered_map roots, shared_ptr leaves,
    block allocator) running the same integration on the same inputs: the
    full 100k-point box scan, three 65k-ray LiDAR scans, and 200k clustered
    setValue calls with 800k lookups. Results must agree. Each side runs three
    times, alternating, and its fastest run counts; the ratios are printed
    (run with -s)."""
    import json
    import time

    def best(runs):
        times = {"reference_storage_s": [], "restatement_s": []}
        for _ in range(3):
            for key, f in runs:
                t = time.perf_counter()
                after = f()
                times[key].append(time.perf_counter() - t)
                after()  # results are read outside the timed region
        out = {k: min(v) for k, v in times.items()}
        out["ratio"] = out["reference_storage_s"] / out["restatement_s"]
        return out

    box = scenes.box_scan()
    lidar = list(scenes.lidar_sequence(3))

    def scan_run(make, scans, rng):
        def run():
            m = make()
            for pts, origin in scans:
                m.insert_point_cloud(pts, origin, rng)

            def after():
                if len(cells) < 2:
                    cells.append(m.cells())
                m.close()
            return after
        return run

    out = {}
    for name, scans, res, rng in (("box_scan_100k", [box], 0.05, 10.0),
                                  ("lidar_3_scans", lidar, 0.1, 50.0)):
        cells = []
        out[name] = best([("reference_storage_s", scan_run(lambda: oracle.RefProb(res), scans, rng)),
                          ("restatement_s", scan_run(lambda: oracle.ProbMap(res), scans, rng))])
        assert all(np.array_equal(a, b) for a, b in zip(cells[0], cells[1]))
    coords = scenes.clusters(n=200_000)
    vals = scenes.cluster_values(coords)
    q = scenes.lookups(coords, n=800_000)
    res = {}

    def grid_run(key, make):
        def run():
            g = make()
            g.set_values(coords, vals)
            res[key] = g.lookup(q)
            return g.close
        return run

    out["grid_200k_set_800k_lookup"] = best([
        ("reference_storage_s", grid_run("r", lambda: oracle.RefGrid(0.05, np.float32))),
        ("restatement_s", grid_run("g", lambda: oracle.Grid(0.05, dtype=np.float32)))])
    assert np.array_equal(res["r"][1], res["g"][1]) and np.array_equal(res["r"][0], res["g"][0])
    with capsys.disabled():
        print("\ncpu timing, reference storage vs restatement:", json.dumps(out))
