"""The C++ drop-in (include/bonxai/*.hpp, include/bonxai_map/probabilistic_map.hpp)
driven by the programs of bonxai_amd/cpp/ on the GPU and compared with the
oracle and with streams written by the reference's own Serialize.

Scripts are replayed by the programs; every value a call returned is checked:
  - dropin_probmap: insertPointCloud (float and double point types),
    addHitPoint/addMissPoint, isOccupied/isFree/isUnknown, const accessors,
    writes through Accessor::value(coord, true), forEachCell,
    getOccupiedVoxels/getFreeVoxels (CoordT and Point3D), update count;
  - dropin_grid: VoxelGrid<float> accessors (setValue, value, value(create)
    with writes through the pointer, setCellOn/Off, isCellOn), forEachCell
    writes, clear, releaseUnusedMemory, Serialize/Deserialize, batch calls;
  - dropin_serialization: the reference example's round trip, its stream
    against the reference's (tests/golden/serial_cube_int.bin), and a stream
    of the reference's loaded, dumped and written again.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from bonxai_amd import scenes

from conftest import sorted_cells

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bonxai_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.gpu


def run(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (run __graft_entry__.build())")
    print(f"running {name} {' '.join(map(str, args))}", flush=True)
    p = subprocess.run([exe, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout
    return p.stdout


def parse_stream(data, vb):
    """A grid stream (serialization.hpp:77-118) as (header, {root: bytes})."""
    nl = data.index(b"\n")
    header = data[:nl].decode()
    inner_bits, leaf_bits = (int(x) for x in header.split("<")[1].split(">")[0].split(",")[-2:])
    inner_words = max((1 << (3 * inner_bits)) >> 6, 1)
    leaf_words = max((1 << (3 * leaf_bits)) >> 6, 1)
    pos = nl + 1
    (nroots,) = struct.unpack_from("<I", data, pos)
    pos += 4
    roots = {}
    for _ in range(nroots):
        start = pos
        key = struct.unpack_from("<3i", data, pos)
        pos += 12
        inner = struct.unpack_from(f"<{inner_words}Q", data, pos)
        pos += 8 * inner_words
        nleaves = sum(bin(w).count("1") for w in inner)
        for _ in range(nleaves):
            lw = struct.unpack_from(f"<{leaf_words}Q", data, pos)
            pos += 8 * leaf_words
            pos += vb * sum(bin(w).count("1") for w in lw)
        assert key not in roots
        roots[key] = data[start:pos]
    assert pos == len(data), "trailing bytes"
    return header, roots


def stream_cells(data, vb, dtype):
    """Every on cell of a stream: sorted coords and values."""
    header, roots = parse_stream(data, vb)
    inner_bits, leaf_bits = (int(x) for x in header.split("<")[1].split(">")[0].split(",")[-2:])
    inner_words = max((1 << (3 * inner_bits)) >> 6, 1)
    leaf_words = max((1 << (3 * leaf_bits)) >> 6, 1)
    im, lm = (1 << inner_bits) - 1, (1 << leaf_bits) - 1
    coords, vals = [], []
    for blob in roots.values():
        rx, ry, rz = struct.unpack_from("<3i", blob, 0)
        pos = 12
        inner = struct.unpack_from(f"<{inner_words}Q", blob, pos)
        pos += 8 * inner_words
        for ii in range(1 << (3 * inner_bits)):
            if not (inner[ii >> 6] >> (ii & 63)) & 1:
                continue
            ox = rx | ((ii & im) << leaf_bits)
            oy = ry | (((ii >> inner_bits) & im) << leaf_bits)
            oz = rz | (((ii >> (2 * inner_bits)) & im) << leaf_bits)
            lw = struct.unpack_from(f"<{leaf_words}Q", blob, pos)
            pos += 8 * leaf_words
            for li in range(1 << (3 * leaf_bits)):
                if (lw[li >> 6] >> (li & 63)) & 1:
                    coords.append((ox | (li & lm), oy | ((li >> leaf_bits) & lm),
                                   oz | ((li >> (2 * leaf_bits)) & lm)))
                    vals.append(blob[pos:pos + vb])
                    pos += vb
    c = np.array(coords, np.int32).reshape(-1, 3)
    v = np.frombuffer(b"".join(vals), dtype)
    return sorted_cells(c, v)


# ------------------------------------------------------------ serialization


def test_serialization_example_roundtrip(tmp_path):
    out = tmp_path / "cube.bin"
    log = run("dropin_serialization", out)
    assert "Round trip looks good!" in log
    mine = out.read_bytes()
    with open(os.path.join(GOLDEN, "serial_cube_int.bin"), "rb") as f:
        ref = f.read()
    # byte-identical per root; the root order is the writer's hash order
    assert parse_stream(mine, 4) == parse_stream(ref, 4)
    assert len(mine) == len(ref)


def test_serialization_loads_reference_stream(tmp_path):
    src = os.path.join(GOLDEN, "serial_clusters_float.bin")
    cells, again = tmp_path / "cells.bin", tmp_path / "again.bin"
    run("dropin_serialization", "--load", src, cells, again)
    with open(src, "rb") as f:
        ref = f.read()
    want_c, want_v = stream_cells(ref, 4, np.float32)
    raw = cells.read_bytes()
    (n,) = struct.unpack_from("<Q", raw, 0)
    rec = np.frombuffer(raw[8:], dtype=[("c", "<i4", 3), ("v", "<f4")], count=n)
    got_c, got_v = sorted_cells(np.array(rec["c"]), np.array(rec["v"]))
    assert np.array_equal(got_c, want_c) and np.array_equal(got_v.view(np.uint32), want_v.view(np.uint32))
    assert parse_stream(again.read_bytes(), 4) == parse_stream(ref, 4)


# ------------------------------------------------------------ ProbabilisticMap


class Script:
    def __init__(self):
        self.parts = []
        self.nops = 0

    def op(self, kind, payload=b""):
        self.parts.append(struct.pack("<I", kind) + payload)
        self.nops += 1

    def coords(self, c):
        c = np.ascontiguousarray(c, np.int32)
        return struct.pack("<I", len(c)) + c.tobytes()

    def bytes(self, head):
        return head + struct.pack("<I", self.nops) + b"".join(self.parts)


def query_coords(rng, cells, n=300):
    """Coords on cells of the map, next to them, and far away (no leaf)."""
    pick = cells[rng.integers(0, len(cells), n // 3)]
    near = pick + rng.integers(-2, 3, (n // 3, 3)).astype(np.int32)
    far = rng.integers(-5000, 5000, (n - 2 * (n // 3), 3)).astype(np.int32)
    return np.concatenate([pick, near, far]).astype(np.int32)


def test_probabilistic_map_dropin(tmp_path):
    res = 0.05
    rng = np.random.default_rng(5)
    model = oracle.ProbMap(res)
    sc = Script()
    expect_q = []

    def insert(pts, origin, rng_m, f64):
        if f64:
            pts = np.ascontiguousarray(pts, np.float64)
            sc.op(1, struct.pack("<I4d", len(pts), *origin, rng_m) + pts.tobytes())
        else:
            # a float point type has a float origin (insertPointCloud takes const PointT&)
            origin = tuple(float(np.float32(x)) for x in origin)
            pts = np.ascontiguousarray(pts, np.float32)
            sc.op(0, struct.pack("<I4d", len(pts), *origin, rng_m) + pts.tobytes())
        model.insert_point_cloud(pts, np.array(origin, np.float64), rng_m)

    def query(q):
        sc.op(4, sc.coords(q))
        st = model.classify(q)
        vals, found = model.grid.lookup(q)
        expect_q.append((st, np.where(found, vals.view(np.int32).astype(np.int64), -1)))

    a, oa = scenes.box_scan(n=20_000, seed=3)
    insert(a, oa, 10.0, False)
    c0, _ = model.cells()
    query(query_coords(rng, c0))
    b, ob = scenes.box_scan(n=15_000, seed=4)
    insert(b.astype(np.float64) + 0.013, (float(ob[0]) + 0.2, float(ob[1]), float(ob[2])), 10.0, True)
    for xyz, hit in (((0.31, -0.42, 0.77), True), ((1.5, 1.25, -0.5), False), ((-2.0, 0.1, 0.3), True)):
        sc.op(2 if hit else 3, struct.pack("<3d", *xyz))
        (model.add_hit_point if hit else model.add_miss_point)(np.array(xyz))
    c1, _ = model.cells()
    query(query_coords(rng, c1))
    # writes through Accessor::value(coord, true): existing and new cells
    wc = np.concatenate([c1[rng.integers(0, len(c1), 50)],
                         rng.integers(-300, 300, (50, 3)).astype(np.int32)])
    wc = np.unique(wc, axis=0)
    words = ((rng.integers(-400000, 400000, len(wc)) << 4) | rng.integers(0, 4, len(wc))).astype(np.int32)
    sc.op(5, sc.coords(wc) + words.tobytes())
    model.grid.set_values(wc, words.view(np.uint32))
    c, _ = scenes.box_scan(n=10_000, seed=6)
    insert(c, oa, 4.0, False)
    # the server's ingest pre-step: sensor-frame PointXYZ-like points with
    # non-finite entries, a float rigid transform, origin from the matrix
    d, _ = scenes.box_scan(n=12_000, seed=8)
    xyzw = np.zeros((len(d), 4), np.float32)
    xyzw[:, :3] = d - d.mean(0)
    xyzw[::97, 1] = np.nan
    xyzw[5::131, 2] = np.inf
    ang = 0.4
    tf = np.array([[np.cos(ang), -np.sin(ang), 0, 0.35], [np.sin(ang), np.cos(ang), 0, -0.2],
                   [0, 0, 1, 0.9], [0, 0, 0, 1]], np.float32)
    sc.op(6, struct.pack("<I", len(xyzw)) + tf.tobytes() + struct.pack("<d", 6.0) + xyzw.tobytes())
    model.insert_point_cloud_tf(xyzw, tf, 6.0)
    c2, _ = model.cells()
    query(query_coords(rng, c2))

    script, result = tmp_path / "script.bin", tmp_path / "result.bin"
    script.write_bytes(sc.bytes(struct.pack("<d", res)))
    run("dropin_probmap", script, result)
    check_probmap_result(result.read_bytes(), expect_q, model, res)
    model.close()


def check_probmap_result(raw, expect_q, model, res):
    """Compare a dropin_probmap result with the oracle model's state: every
    query op's outputs (expect_q: oracle states and words, in order), then
    forEachCell, getOccupied/FreeVoxels, the Point3D overload,
    activeCellsCount and the update count."""
    pos = 0
    for st, word in expect_q:
        n = len(st)
        rec = np.frombuffer(raw, dtype=[("o", "i1"), ("f", "i1"), ("u", "i1"), ("s", "i1"), ("w", "<i8")],
                            count=n, offset=pos)
        pos += rec.nbytes
        assert np.array_equal(rec["o"], (st == 1).astype(np.int8))
        assert np.array_equal(rec["f"], (st == -1).astype(np.int8))
        assert np.array_equal(rec["u"], (st == 0).astype(np.int8))
        assert np.array_equal(rec["s"], st)
        assert np.array_equal(rec["w"], word)
    (n,) = struct.unpack_from("<Q", raw, pos)
    pos += 8
    rec = np.frombuffer(raw, dtype=[("c", "<i4", 3), ("w", "<i4")], count=n, offset=pos)
    pos += rec.nbytes
    got_c, got_w = sorted_cells(np.array(rec["c"]), np.array(rec["w"]).view(np.uint32))
    want_c, want_w = model.cells()
    assert np.array_equal(got_c, want_c) and np.array_equal(got_w, want_w)
    for want in (model.occupied(), model.free()):
        (n,) = struct.unpack_from("<Q", raw, pos)
        pos += 8
        got = np.frombuffer(raw, "<i4", 3 * n, pos).reshape(-1, 3)
        pos += 12 * n
        assert np.array_equal(sorted_cells(got, None)[0], want)
    (n,) = struct.unpack_from("<Q", raw, pos)
    pos += 8
    pts = np.frombuffer(raw, "<f8", 3 * n, pos).reshape(-1, 3)
    pos += 24 * n
    occ = model.occupied()
    assert n == len(occ)
    want_pts = occ.astype(np.float64) * res
    got_sorted = pts[np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0]))]
    assert np.array_equal(got_sorted, want_pts)
    active, upd = struct.unpack_from("<QI", raw, pos)
    assert active == len(want_c)
    assert upd == model.update_count


def test_checkpoint_resume_and_local_refresh(tmp_path):
    """ProbabilisticMap::saveCheckpoint / loadCheckpoint: a map saved after four
    scans (update_count 2) and restored into a new map continues exactly as
    the uninterrupted map over three more scans; the checkpoint opens with the
    bytes of Serialize(grid()).  The scans have a 2 m reach from moving
    origins, so most cached pages lie outside a scan's reach and are not
    re-read, and an addHitPoint endpoint 4 m away extends the next scan's
    reach: the accessors' values after every scan equal the oracle's."""
    res = 0.05
    rng = np.random.default_rng(17)
    model = oracle.ProbMap(res)
    sc = Script()
    expect_q = []
    seq = []  # (points, origin) of every scan, for the sensitivity check

    def insert(pts, origin):
        origin = tuple(float(np.float32(x)) for x in origin)
        pts = np.ascontiguousarray(pts, np.float32)
        sc.op(0, struct.pack("<I4d", len(pts), *origin, 2.0) + pts.tobytes())
        model.insert_point_cloud(pts, np.array(origin, np.float64), 2.0)
        seq.append((pts, np.array(origin, np.float64)))

    def query():
        c, _ = model.cells()
        q = np.concatenate([c, query_coords(rng, c, 3000)]).astype(np.int32)
        sc.op(4, sc.coords(q))
        st = model.classify(q)
        vals, found = model.grid.lookup(q)
        expect_q.append((st, np.where(found, vals.view(np.int32).astype(np.int64), -1)))

    room, o1 = scenes.box_scan(n=20_000, seed=11)
    o2 = o1 + np.array([1.5, 0.5, 0.0])
    insert(room[:10_000], o1)
    query()
    insert(room[10_000:], o2)
    query()
    hit = (float(o1[0]) - 4.0, float(o1[1]) + 0.3, float(o1[2]))
    sc.op(2, struct.pack("<3d", *hit))
    model.add_hit_point(np.array(hit))
    insert(room[::3], o1)
    query()
    assert model.update_count == 1
    insert(room[1::3], o2)
    assert model.update_count == 2
    sc.op(7)
    for k, o in enumerate((o1, o2, o1)):
        insert(room[(k % 3)::2], o)
        query()
    script, result = tmp_path / "script.bin", tmp_path / "result.bin"
    script.write_bytes(sc.bytes(struct.pack("<d", res)))
    out = run("dropin_probmap", script, result)
    assert "checkpoint" in out and "update_count 2" in out
    check_probmap_result(result.read_bytes(), expect_q, model, res)
    # the scenario is sensitive to the counter: restarting it at 1 after the
    # fourth scan gives another map
    alt = oracle.ProbMap(res)
    for k, (pts, origin) in enumerate(seq[:4]):
        if k == 2:
            alt.add_hit_point(np.array(hit))
        alt.insert_point_cloud(pts, origin, 2.0)
    alt.update_count = 1
    for pts, origin in seq[4:]:
        alt.insert_point_cloud(pts, origin, 2.0)
    a_c, a_w = alt.cells()
    m_c, m_w = model.cells()
    assert not (np.array_equal(a_c, m_c) and np.array_equal(a_w, m_w))
    alt.close()
    model.close()


# ------------------------------------------------------------ VoxelGrid<float>


def test_voxel_grid_dropin(tmp_path):
    rng = np.random.default_rng(9)
    model = oracle.Grid(0.1, 2, 3, np.float32)
    sc = Script()
    expect = []  # (dtype, values) of every per-op output, in order

    def pool(n):
        base = rng.integers(-40, 40, (n, 3)).astype(np.int32)
        return base

    def vals(n):
        return rng.standard_normal(n).astype(np.float32)

    def fmt(c, v=None):
        return sc.coords(c) + (np.ascontiguousarray(v, np.float32).tobytes() if v is not None else b"")

    for step in range(2):
        c = pool(400)
        v = vals(400)
        sc.op(0, fmt(c, v))  # setValue, one at a time (duplicates included)
        expect.append(("i1", np.array([model.set_values(c[i:i + 1], v[i:i + 1])[0] for i in range(len(c))],
                                      np.int8)))
        q = np.concatenate([c[:100], pool(100)])
        sc.op(1, fmt(q))
        mv, mf = model.lookup(q)
        expect.append(("rec", (mf.astype(np.int8), np.where(mf, mv, np.float32(0)))))
        c2 = np.concatenate([c[100:150], pool(50)])
        v2 = vals(len(c2))
        sc.op(2, fmt(c2, v2))  # value(create) then write through the pointer
        before = []
        for i in range(len(c2)):
            pv, pf = model.lookup(c2[i:i + 1])
            before.append(pv[0] if pf[0] else np.float32(0))
            model.set_values(c2[i:i + 1], v2[i:i + 1])
        expect.append(("f4", np.array(before, np.float32)))
        c3 = np.concatenate([c[150:200], pool(50)])
        d3 = vals(len(c3))
        sc.op(3, fmt(c3, d3))
        expect.append(("i1", np.array([model.set_cells_on(c3[i:i + 1], d3[i:i + 1])[0]
                                       for i in range(len(c3))], np.int8)))
        c4 = np.concatenate([c[200:300], pool(50)])
        sc.op(4, fmt(c4))
        expect.append(("i1", np.array([model.set_cells_off(c4[i:i + 1])[0] for i in range(len(c4))], np.int8)))
        c5 = np.concatenate([c[300:350], pool(100)])
        v5 = vals(len(c5))
        sc.op(5, fmt(c5, v5))  # GPU batch
        expect.append(("i1", model.set_values(c5, v5).astype(np.int8)))
        sc.op(6)  # forEachCell: v = 2 v + 1
        ec, ev = model.export()
        model.set_values(ec, np.float32(2) * ev + np.float32(1))
        q = np.concatenate([c[:200], pool(50)])
        sc.op(11, fmt(q))
        expect.append(("i1", model.lookup(q)[1].astype(np.int8)))
        sc.op(10)  # Serialize / Deserialize into the same grid
        sc.op(8)  # releaseUnusedMemory
        model.release_unused()
    sc.op(7)  # every cell off, leaves kept
    model.clear(1)
    c = pool(200)
    v = vals(200)
    sc.op(0, fmt(c, v))
    expect.append(("i1", np.array([model.set_values(c[i:i + 1], v[i:i + 1])[0] for i in range(len(c))], np.int8)))
    sc.op(8)
    model.release_unused()
    sc.op(9)  # free everything
    model.clear(0)
    c = pool(300)
    v = vals(300)
    sc.op(5, fmt(c, v))
    expect.append(("i1", model.set_values(c, v).astype(np.int8)))
    sc.op(6)
    ec, ev = model.export()
    model.set_values(ec, np.float32(2) * ev + np.float32(1))

    script, result = tmp_path / "script.bin", tmp_path / "result.bin"
    script.write_bytes(sc.bytes(struct.pack("<dII", 0.1, 2, 3)))
    run("dropin_grid", script, result)
    raw = result.read_bytes()
    pos = 0
    for kind, want in expect:
        if kind == "rec":
            f, v = want
            rec = np.frombuffer(raw, dtype=[("f", "i1"), ("v", "<f4")], count=len(f), offset=pos)
            pos += rec.nbytes
            assert np.array_equal(rec["f"], f)
            assert np.array_equal(rec["v"].view(np.uint32), v.astype(np.float32).view(np.uint32))
        else:
            got = np.frombuffer(raw, kind, len(want), pos)
            pos += got.nbytes
            if kind == "f4":
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
            else:
                assert np.array_equal(got, want)
    (n,) = struct.unpack_from("<Q", raw, pos)
    pos += 8
    rec = np.frombuffer(raw, dtype=[("c", "<i4", 3), ("v", "<f4")], count=n, offset=pos)
    pos += rec.nbytes
    got_c, got_v = sorted_cells(np.array(rec["c"]), np.array(rec["v"]))
    want_c, want_v = model.export()
    assert np.array_equal(got_c, want_c)
    assert np.array_equal(got_v.view(np.uint32), want_v.view(np.uint32))
    (active,) = struct.unpack_from("<Q", raw, pos)
    assert active == model.active_cells_count() == len(want_c)
    model.close()
