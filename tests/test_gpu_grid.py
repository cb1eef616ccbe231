"""GPU parity of the VoxelGrid<T> bulk operations against the oracle
(setValue / setCellOn / setCellOff / value / isCellOn / forEachCell /
activeCellsCount / clear), bit-exact."""
import numpy as np
import pytest

from conftest import assert_same_cells, sorted_cells

pytestmark = pytest.mark.gpu


def test_set_and_lookup_clusters(ctx):
    """Reduced VoxelGrid<float> workload: 200k clustered inserts with value =
    float(coord hash), then 400k lookups (half hits)."""
    import bonxai_amd as bx
    import oracle
    from bonxai_amd import scenes
    coords = scenes.clusters(n=200_000)
    vals = scenes.cluster_values(coords)
    g = bx.VoxelGrid(ctx, 0.05, dtype=np.float32)
    og = oracle.Grid(0.05, dtype=np.float32)
    was = g.set_values(coords, vals, return_was_on=True)
    assert np.array_equal(was, og.set_values(coords, vals))
    q = scenes.lookups(coords, n=400_000)
    v, f = g.lookup(q)
    ov, of = og.lookup(q)
    assert np.array_equal(f, of)
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32))
    assert_same_cells(g.export(), og.export(), "export")
    assert g.active_cells_count() == og.active_cells_count()
    assert g.digest() == og.digest()
    g.close()
    og.close()


def test_duplicates_last_write_wins(ctx):
    """Repeated coordinates with different values: the last value stays and
    was_on is the prior state for the first occurrence, true afterwards."""
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(0)
    coords = rng.integers(-40, 40, (50_000, 3)).astype(np.int32)
    vals = rng.integers(-10**9, 10**9, 50_000).astype(np.int32)
    g = bx.VoxelGrid(ctx, 0.1, dtype=np.int32)
    og = oracle.Grid(0.1, dtype=np.int32)
    for k in range(2):
        sl = slice(k * 25_000, (k + 1) * 25_000)
        assert np.array_equal(g.set_values(coords[sl], vals[sl], True), og.set_values(coords[sl], vals[sl]))
    assert_same_cells(g.export(), og.export(), "dups")
    g.close()
    og.close()


def test_set_values_without_was_on(ctx):
    """setValue without was_on (the call keeps only each cell's last input):
    repeated coordinates keep their last value, over three calls into float,
    int32 and binary grids of several shapes, and an out-of-range coordinate
    fails the call and leaves the cells unchanged."""
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(7)
    coords = rng.integers(-60, 60, (60_000, 3)).astype(np.int32)
    ints = rng.integers(-10**9, 10**9, 60_000).astype(np.int32)
    for dtype, ib, lb in ((np.int32, 2, 3), (np.float32, 1, 2), (np.int32, 4, 4), (None, 2, 3)):
        vals = None if dtype is None else (ints if dtype == np.int32 else (ints % 1000).astype(np.float32))
        g = bx.VoxelGrid(ctx, 0.1, ib, lb, dtype=dtype)
        og = oracle.Grid(0.1, ib, lb, dtype=dtype)
        for k in range(3):
            sl = slice(k * 20_000, (k + 1) * 20_000)
            assert g.set_values(coords[sl], None if vals is None else vals[sl]) is None
            og.set_values(coords[sl], None if vals is None else vals[sl])
        what = f"dtype {dtype}, bits {ib}/{lb}"
        if dtype is None:
            assert np.array_equal(sorted_cells(g.export()[0])[0], og.export()[0]), what
        else:
            assert_same_cells(g.export(), og.export(), what)
        assert g.active_cells_count() == og.active_cells_count(), what
        g.close()
        og.close()
    g = bx.VoxelGrid(ctx, 0.1, dtype=np.int32)
    g.set_values(coords[:1000], ints[:1000])
    before = sorted_cells(*g.export())
    bad = np.concatenate([coords[1000:2000], np.array([[1 << 30, 0, 0]], np.int32)])
    with pytest.raises(bx.BonxaiError) as e:
        g.set_values(bad, ints[:1001])
    assert e.value.code == bx.ERR_COORD_RANGE
    g.clear_error()
    assert_same_cells(g.export(), before, "a failed call")
    g.close()


def test_cells_on_off(ctx):
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(1)
    a = rng.integers(-30, 30, (20_000, 3)).astype(np.int32)
    b = rng.integers(-30, 30, (20_000, 3)).astype(np.int32)
    d = rng.integers(0, 1000, 20_000).astype(np.int64)
    g = bx.VoxelGrid(ctx, 0.1, dtype=np.int64)
    og = oracle.Grid(0.1, dtype=np.int64)
    assert np.array_equal(g.set_cells_on(a, d, True), og.set_cells_on(a, d))
    assert np.array_equal(g.set_cells_off(b, True), og.set_cells_off(b))
    assert np.array_equal(g.set_cells_on(b, d, True), og.set_cells_on(b, d))
    assert_same_cells(g.export(), og.export(), "on/off")
    v, f = g.lookup(b)
    ov, of = og.lookup(b)
    assert np.array_equal(f, of) and np.array_equal(v, ov)
    g.close()
    og.close()


def test_binary_grid(ctx):
    """BinaryVoxelGrid (VoxelGrid<EmptyVoxel>, bonxai.hpp:335)."""
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(2)
    a = rng.integers(-100, 100, (30_000, 3)).astype(np.int32)
    g = bx.VoxelGrid(ctx, 0.1, dtype=None)
    og = oracle.Grid(0.1, dtype=None)
    assert np.array_equal(g.set_values(a, None, True), og.set_values(a, None))
    q = rng.integers(-100, 100, (30_000, 3)).astype(np.int32)
    assert np.array_equal(g.lookup(q)[1], og.lookup(q)[1])
    assert np.array_equal(sorted_cells(g.export()[0])[0], og.export()[0])
    g.close()
    og.close()


@pytest.mark.parametrize("bits", [(1, 1), (2, 3), (3, 2), (4, 4)])
def test_other_grid_shapes(ctx, bits):
    import bonxai_amd as bx
    import oracle
    ib, lb = bits
    rng = np.random.default_rng(ib * 10 + lb)
    a = rng.integers(-500, 500, (20_000, 3)).astype(np.int32)
    v = rng.random(20_000).astype(np.float32)
    g = bx.VoxelGrid(ctx, 0.1, ib, lb, dtype=np.float32)
    og = oracle.Grid(0.1, ib, lb, dtype=np.float32)
    assert np.array_equal(g.set_values(a, v, True), og.set_values(a, v))
    assert_same_cells(g.export(), og.export(), f"bits {bits}")
    assert g.active_cells_count() == og.active_cells_count()
    g.close()
    og.close()


def test_clear_options(ctx):
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(3)
    a = rng.integers(-50, 50, (10_000, 3)).astype(np.int32)
    v = rng.random(10_000).astype(np.float32)
    g = bx.VoxelGrid(ctx, 0.1)
    og = oracle.Grid(0.1)
    g.set_values(a, v)
    og.set_values(a, v)
    g.clear(bx.SET_ALL_CELLS_OFF)
    og.clear(1)
    assert g.active_cells_count() == 0 == og.active_cells_count()
    b = a[:5000]
    assert np.array_equal(g.set_cells_on(b, None, True), og.set_cells_on(b))
    assert_same_cells(g.export(), og.export(), "after SET_ALL_CELLS_OFF")
    g.clear(bx.CLEAR_MEMORY)
    assert g.active_cells_count() == 0
    assert g.stats()["leaf_count"] == 0
    g.set_values(a, v)
    og.clear(0)
    og.set_values(a, v)
    assert_same_cells(g.export(), og.export(), "after CLEAR_MEMORY")
    g.close()
    og.close()


@pytest.mark.parametrize("dtype", [np.float32, None])
def test_release_unused_memory(ctx, dtype):
    """releaseUnusedMemory (bonxai.hpp:367-387): leaves whose cells are all off
    and roots left without leaves go; the rest (values included) stays, and
    the map keeps working after the pools are compacted."""
    import bonxai_amd as bx
    import oracle
    rng = np.random.default_rng(8)
    a = rng.integers(-400, 400, (60_000, 3)).astype(np.int32)
    v = rng.random(len(a)).astype(np.float32)
    g = bx.VoxelGrid(ctx, 0.1, dtype=dtype)
    og = oracle.Grid(0.1, dtype=dtype)
    if dtype is None:
        g.set_cells_on(a)
        og.set_cells_on(a)
    else:
        g.set_values(a, v)
        og.set_values(a, v)
    # every cell of a third of the leaves off, plus scattered cells elsewhere
    leaf = a & ~7
    drop = (np.abs(leaf).sum(1) // 8) % 3 == 0
    off = np.concatenate([a[drop], a[~drop][::7]])
    assert np.array_equal(g.set_cells_off(off, True), og.set_cells_off(off))
    before = g.stats()
    g.release_unused_memory()
    og.release_unused()
    st = g.stats()
    assert st["leaf_count"] == og.leaf_count() < before["leaf_count"]
    assert st["root_count"] == og.root_count() <= before["root_count"]
    assert st["leaf_capacity"] < before["leaf_capacity"]
    assert_same_cells(g.export(), og.export(), "after releaseUnusedMemory")
    assert g.digest() == og.digest()
    # new leaves and values into the compacted pools
    b = rng.integers(-600, 600, (20_000, 3)).astype(np.int32)
    w = rng.random(len(b)).astype(np.float32)
    if dtype is None:
        assert np.array_equal(g.set_cells_on(b, None, True), og.set_cells_on(b))
    else:
        assert np.array_equal(g.set_values(b, w, True), og.set_values(b, w))
    assert_same_cells(g.export(), og.export(), "after refill")
    vals, found = g.lookup(a)
    ovals, ofound = og.lookup(a)
    assert np.array_equal(found, ofound)
    if dtype is not None:
        assert np.array_equal(vals.view(np.uint32), ovals.view(np.uint32))
    # releasing everything
    g.clear(bx.SET_ALL_CELLS_OFF)
    g.release_unused_memory()
    assert g.stats()["leaf_count"] == 0 and g.stats()["root_count"] == 0
    g.close()
    og.close()


def test_leaf_paging_round_trip(ctx):
    """Whole-leaf read/write (the host accessor's paging path)."""
    import bonxai_amd as bx
    rng = np.random.default_rng(4)
    a = rng.integers(-60, 60, (5_000, 3)).astype(np.int32)
    v = rng.random(5_000).astype(np.float32)
    g = bx.VoxelGrid(ctx, 0.1)
    g.set_values(a, v)
    origins = np.unique(a & ~7, axis=0)
    masks, cells, exists = g.read_leaves(origins)
    assert exists.all()
    h = bx.VoxelGrid(ctx, 0.1)
    h.write_leaves(origins, masks, cells)
    assert_same_cells(h.export(), sorted_cells(*g.export()), "paged copy")
    g.close()
    h.close()


def test_coord_range_error(ctx):
    import bonxai_amd as bx
    g = bx.VoxelGrid(ctx, 0.1)
    with pytest.raises(bx.BonxaiError) as e:
        g.set_values(np.array([[1 << 30, 0, 0]], np.int32), np.ones(1, np.float32))
    assert e.value.code == bx.ERR_COORD_RANGE
    g.set_values(np.array([[5, 5, 5]], np.int32), np.ones(1, np.float32))
    assert g.active_cells_count() == 1
    g.close()


def _probe_mix(i):
    """The gather probe's index hash on uint64 (wrapping) arithmetic."""
    x = i.astype(np.uint64)
    x ^= x >> np.uint64(31)
    x *= np.uint64(0x7fb5d329728ea185)
    x ^= x >> np.uint64(27)
    x *= np.uint64(0x81dadef4bc2dd44d)
    return x ^ (x >> np.uint64(33))


def test_gather_probe_reads_table(ctx):
    """The scattered-read probe behind the lookup ceiling reads the words its
    index hash names, including a ragged last workgroup; bad sizes are errors."""
    import bonxai_amd as bx
    words = 1 << 16
    table_host = (np.arange(words, dtype=np.uint32) * np.uint32(2654435761))
    table = ctx.to_device(table_host)
    n = 100_003
    out = ctx.gather_probe(words * 4, n, table=table)
    idx = _probe_mix(np.arange(n, dtype=np.uint64)) & np.uint64(words - 1)
    assert np.array_equal(out.numpy(), table_host[idx.astype(np.int64)])
    out.free()
    for bad in (words * 4 - 4, 0, 6):
        with pytest.raises(bx.BonxaiError):
            ctx.gather_probe(bad, 16, table=table)
    with pytest.raises(ValueError):
        ctx.gather_probe(words * 8, 16, table=table)  # larger than the table: refused on the host
    table.free()
