"""The leaf-granular walk of the GPU leaf pass (LeafWalk in
bonxai_amd/csrc/bxg_kernels.hip), restated with the device's 32-bit state
and its 2^30 saturation, against the cell walk of RayIterator
(probabilistic_map.hpp:162-203): the closed form cell_k = o + sgn(d) *
floor((2k|d| + m) / (2m)), itself checked equal to the oracle's ray in
test_dropin_host.py.  Short rays: the walk's leaves and entry cells equal the
cell walk's, exactly.  Long rays (up to 2^24 cells, coordinates to 2^25):
every leaf the walk reports holds its entry cell, the cell before it lies in
the previous leaf, and the walk ends in the leaf of the last cell."""
import random

import numpy as np

import oracle

M32 = (1 << 32) - 1
NEVER = 0x7FFFFFFF


def _cell(o, d, s, m, k):
    return tuple(o[i] + s[i] * ((2 * k * d[i] + m) // (2 * m)) for i in range(3))


def leaf_walk(o, e, lb, k0, k1, cap=None):
    """(leaf origin, entry cell) pairs for cells [k0, k1) of the ray o -> e."""
    d = [e[i] - o[i] for i in range(3)]
    s = [(-1 if v < 0 else 1) for v in d]
    d = [abs(v) for v in d]
    m = max(d)
    L = 1 << lb
    lm = L - 1
    c0 = _cell(o, d, s, m, k0) if m else tuple(o)  # m = 0: the endpoint is the origin's voxel
    st = []
    for a in range(3):
        n = (2 * k0 * d[a] + m) // (2 * m) if k0 else 0
        c = c0[a]
        b = 2 * (d[a] if d[a] > 0 else 1)
        if d[a] == 0:
            st.append([c & ~lm, s[a] * L, NEVER, 0, 0, 0, 0, b])
            continue
        N = n + ((L - (c & lm)) if s[a] > 0 else ((c & lm) + 1))
        A = (2 * N - 1) * m
        q64 = A // b
        Q64 = (L * m) // d[a]
        q = min(q64, 1 << 30)
        r = (A - q64 * b) & M32
        Q = min(Q64, 1 << 30)
        R = (2 * (L * m - Q64 * d[a])) & M32
        st.append([c & ~lm, s[a] * L, (q + (1 if r > 0 else 0)) & M32, q, r, Q, R, b])
    out, k = [], k0
    while cap is None or len(out) < cap:
        out.append((tuple(x[0] for x in st), k))
        kn = min(x[2] for x in st)
        if kn >= k1:
            break
        for x in st:
            if x[2] == kn:
                x[0] += x[1]
                x[3] = (x[3] + x[5]) & M32
                x[4] = (x[4] + x[6]) & M32
                if x[4] >= x[7]:
                    x[3] = (x[3] + 1) & M32
                    x[4] -= x[7]
                x[2] = (x[3] + (1 if x[4] > 0 else 0)) if x[3] < (1 << 30) else NEVER
        k = kn
    return out


def test_short_rays_equal_cell_walk():
    rng = random.Random(3)
    for _ in range(1500):
        lb = rng.choice([1, 2, 3, 4])
        o = [rng.randint(-300, 300) for _ in range(3)]
        e = [o[i] + rng.randint(-200, 200) for i in range(3)]
        cells = oracle.compute_ray(np.array(o), np.array(e))
        m = len(cells)
        if m == 0:
            continue
        k0 = rng.randint(0, m - 1) if rng.random() < 0.5 else 0
        k1 = rng.randint(k0 + 1, m) if rng.random() < 0.5 else m
        want = []
        for k in range(k0, k1):
            leaf = tuple(int(v) & ~((1 << lb) - 1) for v in cells[k])
            if not want or want[-1][0] != leaf:
                want.append((leaf, k))
        assert leaf_walk(o, e, lb, k0, k1) == want, (o, e, lb, k0, k1)


def test_leaf_pass_reaches_the_endpoint_leaf():
    """The leaf pass walks cells [k0, m + 1) of a winner's ray, its m cells and
    then the endpoint (cell m is e), so that the endpoint's leaf is created
    with the ray's; also for an endpoint in the origin's voxel (m = 0) and for
    the endpoint-only walk [m, m + 1) of an addHitPoint / addMissPoint scan."""
    rng = random.Random(7)
    for _ in range(3000):
        lb = rng.choice([1, 2, 3, 4])
        o = [rng.randint(-300, 300) for _ in range(3)]
        e = [o[i] + rng.randint(-40, 40) for i in range(3)] if rng.random() < 0.9 else list(o)
        d = [abs(e[i] - o[i]) for i in range(3)]
        s = [(-1 if e[i] < o[i] else 1) for i in range(3)]
        m = max(d)
        cells = [_cell(o, d, s, m, k) for k in range(m + 1)] if m else [tuple(o)]
        assert cells[-1] == tuple(e)
        r = rng.random()
        k0 = 0 if r < 0.4 else (m if r < 0.6 else rng.randint(0, m))
        want = []
        for k in range(k0, m + 1):
            leaf = tuple(v & ~((1 << lb) - 1) for v in cells[k])
            if not want or want[-1][0] != leaf:
                want.append((leaf, k))
        assert leaf_walk(o, e, lb, k0, m + 1) == want, (o, e, lb, k0)


def test_long_rays_against_closed_form():
    rng = random.Random(5)
    R = 1 << 25
    for _ in range(150):
        lb = rng.choice([1, 2, 3, 4])
        o = [rng.randint(-R + 1, R - 1) for _ in range(3)]
        span = rng.choice([1000, 100_000, 1 << 24])
        e = [max(-R, min(R - 1, o[i] + rng.randint(-span, span))) for i in range(3)]
        if rng.random() < 0.3:
            e[rng.randint(0, 2)] = o[rng.randint(0, 2)]
        d = [abs(e[i] - o[i]) for i in range(3)]
        s = [(-1 if e[i] < o[i] else 1) for i in range(3)]
        m = max(d)
        if m == 0:
            continue
        k0 = rng.randint(0, m - 1) if rng.random() < 0.5 else 0
        k1 = rng.randint(k0 + 1, m) if rng.random() < 0.5 else m
        lv = leaf_walk(o, e, lb, k0, k1, cap=400)
        L = 1 << lb
        for j, (leaf, k) in enumerate(lv):
            assert tuple(v & ~(L - 1) for v in _cell(o, d, s, m, k)) == leaf
            if j:
                assert tuple(v & ~(L - 1) for v in _cell(o, d, s, m, k - 1)) != leaf
        if len(lv) < 400:
            assert tuple(v & ~(L - 1) for v in _cell(o, d, s, m, k1 - 1)) == lv[-1][0]


def test_segment_start_equals_stepping():
    """RayWalk::seek (a segment of a ray starting mid-way, used when a call has
    fewer rays than lanes): position n = floor((2k|d| + m) / (2m)) steps along
    each axis and error k|d| - n m, the state the error walk of
    probabilistic_map.hpp:162-203 reaches after k steps."""
    rng = random.Random(1)
    for _ in range(3000):
        o = [rng.randint(-50, 50) for _ in range(3)]
        e = [o[i] + rng.randint(-60, 60) for i in range(3)]
        d = [e[i] - o[i] for i in range(3)]
        s = [(-1 if v < 0 else 1) for v in d]
        d = [abs(v) for v in d]
        m = max(d)
        c, er = list(o), [0, 0, 0]
        for k in range(m):
            n = [(2 * k * d[i] + m) // (2 * m) for i in range(3)]
            assert [o[i] + s[i] * n[i] for i in range(3)] == c
            assert [k * d[i] - n[i] * m for i in range(3)] == er
            for i in range(3):
                er[i] += d[i]
                if 2 * er[i] >= m:
                    c[i] += s[i]
                    er[i] -= m


def test_step_threshold_identity():
    """RayWalk::advance tests e >= h with h = (m + 1) >> 1 in place of the
    reference's 2 e >= m (probabilistic_map.hpp:162-203): equal for every
    error the walk holds (-m <= e <= m after the add) and every length m."""
    for m in range(1, 600):
        h = (m + 1) >> 1
        e = np.arange(-2 * m, 2 * m + 1)
        assert np.array_equal(2 * e >= m, e >= h)
    for m in (2**20 - 1, 2**20, 2**24 + 1, 2**26):
        h = (m + 1) >> 1
        e = np.array([-m, -1, 0, h - 2, h - 1, h, h + 1, m - 1, m], dtype=np.int64)
        assert np.array_equal(2 * e >= m, e >= h)
