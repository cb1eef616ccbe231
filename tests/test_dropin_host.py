"""Host-side logic of the C++ drop-in headers, on the CPU: the program
tests/dropin_host_checks.cpp is compiled with g++ against include/ (no GPU
call, no device library needed) and its output is compared with closed forms
and with the oracle's ray (probabilistic_map.hpp:162-203 restated)."""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checks(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    exe = d / "dropin_host_checks"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "dropin_host_checks.cpp"),
                    "-o", str(exe)], check=True)
    rng = np.random.default_rng(17)
    o = rng.integers(-1000, 1000, (400, 3))
    e = o + rng.integers(-40, 41, (400, 3))
    e[:10] = o[:10]  # empty rays
    e[10:20, 1:] = o[10:20, 1:]  # axis-aligned
    pairs = np.stack([o, e], 1).astype(np.int32)
    rays = d / "rays.bin"
    rays.write_bytes(struct.pack("<I", len(pairs)) + pairs.tobytes())
    out = subprocess.run([str(exe), str(rays)], check=True, stdout=subprocess.PIPE, text=True).stdout
    lines = {}
    for line in out.splitlines():
        key, _, rest = line.partition(" ")
        lines.setdefault(key, []).append(rest)
    return lines, pairs


def test_logods_and_options(checks):
    lines, _ = checks
    # the reference's Options defaults (probabilistic_map.hpp:56-64)
    assert lines["logods"][0].split() == ["-405465", "847297", "-1992430", "3476099", "0"]
    assert lines["options"][0].split() == ["-405465", "847297", "-1992430", "3476099", "0"]
    # logods(p) = int(1e6 * log(p / (1 - p))) in double, p a float
    for p, want in ((0.4, -405465), (0.7, 847297), (0.12, -1992430), (0.97, 3476099)):
        pf = float(np.float32(p))
        assert int(1e6 * math.log(pf / (1.0 - pf))) == want


def test_cell_layout(checks):
    lines, _ = checks
    # update_id in the low 4 bits, probability_log in the high 28
    assert int(lines["cell"][0]) == (((-123456) << 4) & 0xFFFFFFFF) | ((-3) & 0xF)


def test_header_parse(checks):
    lines, _ = checks
    assert lines["header"][0].split() == ["Bonxai::ProbabilisticMap::CellT", "2", "3", "0.05"]
    assert lines["header"][1].split() == ["int", "1", "4", "0.1"]


def test_mask(checks):
    lines, _ = checks
    # bit n in word n >> 6 (mask.hpp:329-365); max(bits / 64, 1) words
    assert lines["mask"][0].split() == ["8", "4", "1", "1", "0", "on:", "0", "64", "511"]
    assert lines["mask1"][0].split() == ["1", "8"]


def test_convert_coord_hash(checks):
    lines, _ = checks
    assert [float(x) for x in lines["convert"][0].split()] == [1.5, -2.25, 3, 0.1, 0.2, 0.3, 7, 8, 9]
    want = [math.floor(-0.3 * 20.0), math.floor(0.05 * 20.0), math.floor(1.0 * 20.0)]
    assert [int(x) for x in lines["pos2coord"][0].split()] == want

    def h(x, y, z):  # grid_coord.hpp:209-218
        return ((x * 73856093) ^ (y * 19349669) ^ (z * 83492791)) & ((1 << 20) - 1)

    assert [int(x) for x in lines["hash"][0].split()] == [h(-7, 12345, 99), h(1 << 20, -(1 << 20), 3)]


def test_compute_ray_matches_oracle(checks):
    lines, pairs = checks
    assert len(lines["ray"]) == len(pairs)
    for k, (o, e) in enumerate(pairs):
        head, _, cells = lines["ray"][k].partition(":")
        assert int(head) == k
        got = np.array(cells.split(), np.int32).reshape(-1, 3)
        want = oracle.compute_ray(o, e)
        assert np.array_equal(got, want), (o, e)
        # length max|delta|, origin first, end excluded, empty iff o == e
        assert len(got) == int(np.abs(e - o).max())
        if len(got):
            assert np.array_equal(got[0], o) and not (got == e).all(1).any()


@pytest.fixture(scope="module")
def pcd_check(tmp_path_factory):
    d = tmp_path_factory.mktemp("pcd")
    exe = d / "dropin_pcd_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "dropin_pcd_check.cpp"), "-o", str(exe)], check=True)
    return exe, d


def _run_pcd(exe, d, src):
    out = subprocess.run([str(exe), str(src), str(d / "pts.bin"), str(d / "copy.pcd")], check=True,
                         stdout=subprocess.PIPE, text=True).stdout
    return out, np.fromfile(d / "pts.bin", np.float64).reshape(-1, 3)


def test_pcd_reads_reference_data(pcd_check):
    """ReadPointsFromPCD (pcl_utils.cpp:10-26) on the reference's data/apple.pcd
    (ascii, fields x y z rgb imX imY): every point, as float widened to double."""
    exe, d = pcd_check
    out, pts = _run_pcd(exe, d, os.path.join(ROOT, "tests", "golden", "apple.pcd"))
    want = np.load(os.path.join(ROOT, "tests", "golden", "apple_xyz.npy"))
    assert pts.shape == want.shape
    assert np.array_equal(pts, want.astype(np.float64))
    assert "reread equal 1" in out
    assert "missing file: File doesn't exist" in out


def test_pcd_reads_binary(pcd_check):
    """A binary PCD with x y z among other fields (float32 and float64)."""
    exe, d = pcd_check
    rng = np.random.default_rng(1)
    n = 500
    rec = np.zeros(n, dtype=[("intensity", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f8"), ("ring", "<u2")])
    for k in ("x", "y", "z"):
        rec[k] = rng.uniform(-50, 50, n)
    hdr = ("# .PCD v0.7\nVERSION 0.7\nFIELDS intensity x y z ring\nSIZE 4 4 4 8 2\nTYPE F F F F U\n"
           f"COUNT 1 1 1 1 1\nWIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS {n}\nDATA binary\n")
    src = d / "bin.pcd"
    src.write_bytes(hdr.encode() + rec.tobytes())
    out, pts = _run_pcd(exe, d, src)
    want = np.stack([rec["x"].astype(np.float64), rec["y"].astype(np.float64),
                     rec["z"].astype(np.float32).astype(np.float64)], 1)
    assert np.array_equal(pts, want)
    assert "reread equal 1" in out
