"""Seeded synthetic workloads (ctypes over bonxai_amd/libbxscenes.so).

Every generator is deterministic and libm-free (see scenes.c), so the same
seed gives bit-identical float32 inputs on every machine running the same
build.  The workload functions below fix the sizes of the benchmark
configurations:

  box_scan          single 100k-point scan of an 8 x 6 x 3 m box room, origin
                    at the centre (voxel 0.05 m, max range 10 m)
  clusters          2M int32 voxel coords from 64 Gaussian clusters
                    (sigma 0.5 m, centres in a 100 m cube, voxel 0.05 m),
                    values float(std::hash<CoordT>); 8M lookups, half hits
  lidar_sequence    spinning-LiDAR scans (64 beams x 1024 azimuths,
                    -25..+15 deg) of a ground plane + 300 boxes along a
                    smooth path, 1 m per scan (voxel 0.1 m, max range 50 m)
  batch_scans       64 maps, one LiDAR-scene scan each of log-uniform
                    10k..500k points (voxel 0.1 m, max range 30 m)
  depth_frame       640 x 480 pinhole depth frame (fx = fy = 525,
                    cx = 319.5, cy = 239.5) in a closed box room (max 5 m)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "libbxscenes.so")
_lib = None

_D = ctypes.c_double
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        h = ctypes.CDLL(LIB_PATH)
        h.bxs_box_scan.argtypes = [_U64, _U64, _D, _D, _D, _D, _D, _P, _P]
        h.bxs_clusters.argtypes = [_U64, _U64, _U32, _D, _D, _D, _P]
        h.bxs_lookups.argtypes = [_U64, _U64, _P, _U64, _D, _D, _D, _P]
        h.bxs_lidar_scan.argtypes = [_U64, _U32, _U64, _D, _U32, _U32, _D, _D, _D, _D, _D, _D, _D, _P, _P]
        h.bxs_depth_frame.argtypes = [_U64, _U32, _U32, _D, _D, _D, _D, _U32, _P, _P]
        h.bxs_fnv1a.argtypes = [_P, _U64]
        h.bxs_fnv1a.restype = _U64
        h.bxs_sincos_array.argtypes = [_P, _U64, _P, _P]
        _lib = h
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def fnv1a(a):
    a = np.ascontiguousarray(a)
    return int(lib().bxs_fnv1a(_p(a), a.nbytes))


def digest_scans(scans):
    """SHA-256 (hex) over every scan's float32 points and float64 origin, in
    order: identical on every machine that builds the same inputs."""
    import hashlib
    h = hashlib.sha256()
    for pts, origin in scans:
        h.update(np.ascontiguousarray(pts, np.float32).tobytes())
        h.update(np.ascontiguousarray(origin, np.float64).tobytes())
    return h.hexdigest()


def sincos(x):
    x = np.ascontiguousarray(x, np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().bxs_sincos_array(_p(x), x.size, _p(s), _p(c))
    return s, c


def box_scan(seed=1, n=100_000, size=(8.0, 6.0, 3.0), sigma=0.01, inside_frac=0.10):
    """(points (n,3) float32, origin (3,) float64)."""
    out = np.empty((n, 3), np.float32)
    origin = np.empty(3, np.float64)
    lib().bxs_box_scan(seed, n, size[0], size[1], size[2], sigma, inside_frac, _p(out), _p(origin))
    return out, origin


def clusters(seed=2, n=2_000_000, nclusters=64, extent=100.0, sigma=0.5, voxel=0.05):
    out = np.empty((n, 3), np.int32)
    lib().bxs_clusters(seed, n, nclusters, extent, sigma, voxel, _p(out))
    return out


def lookups(inserted, seed=3, n=8_000_000, hit_frac=0.5, extent=100.0, voxel=0.05):
    ins = np.ascontiguousarray(inserted, np.int32)
    out = np.empty((n, 3), np.int32)
    lib().bxs_lookups(seed, n, _p(ins), len(ins), hit_frac, extent, voxel, _p(out))
    return out


def coord_hash(coords):
    """std::hash<Bonxai::CoordT> (grid_coord.hpp:209-218): the low 20 bits of
    x*73856093 ^ y*19349669 ^ z*83492791 in 64-bit arithmetic."""
    c = np.asarray(coords, np.int64)
    h = (c[:, 0] * 73856093) ^ (c[:, 1] * 19349669) ^ (c[:, 2] * 83492791)
    return (h & ((1 << 20) - 1)).astype(np.uint64)


def cluster_values(coords):
    """Values of the VoxelGrid<float> insert workload: float(coord hash)."""
    return coord_hash(coords).astype(np.float32)


def lidar_scan(world_seed, scan_seed, s, nbeams=64, nazim=1024, elev=(-25.0, 15.0),
               cast_range=120.0, far_range=75.0, range_sigma=0.02, dropout=0.05,
               sensor_h=1.8, nboxes=300):
    """One scan at arc length s metres: (points (nbeams*nazim,3) f32, origin)."""
    out = np.empty((nbeams * nazim, 3), np.float32)
    origin = np.empty(3, np.float64)
    lib().bxs_lidar_scan(world_seed, nboxes, scan_seed, s, nbeams, nazim, elev[0], elev[1],
                         cast_range, far_range, range_sigma, dropout, sensor_h, _p(out), _p(origin))
    return out, origin


def lidar_sequence(nscans=200, world_seed=7, nbeams=64, nazim=1024, max_range=50.0, start=0):
    """Scans start..start+nscans-1 of the sequential mapping workload, 1 m
    apart; no-return rays and dropouts at 1.5 x max_range."""
    for k in range(start, start + nscans):
        yield lidar_scan(world_seed, 1000 + k, float(k), nbeams, nazim,
                         far_range=1.5 * max_range)


def batch_sizes(nmaps=64, seed=11, lo=10_000, hi=500_000):
    """Point counts log-uniform in [lo, hi], multiples of 64 (64 beams)."""
    rng = np.random.default_rng(seed)
    u = rng.random(nmaps)
    n = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    return (np.round(n / 64.0) * 64).astype(np.int64)


def batch_scans(nmaps=64, seed=11, max_range=30.0, lo=10_000, hi=500_000):
    """The batched workload: one scan per map with its own scene seed and
    pose; list of (points, origin)."""
    out = []
    for i, n in enumerate(batch_sizes(nmaps, seed, lo, hi)):
        nazim = int(n) // 64
        pts, origin = lidar_scan(100 + i, 5000 + i, float(3 * i % 97), 64, nazim,
                                 far_range=1.5 * max_range)
        out.append((pts, origin))
    return out


def depth_frame(seed, width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5, nboxes=12):
    out = np.empty((width * height, 3), np.float32)
    origin = np.empty(3, np.float64)
    lib().bxs_depth_frame(seed, width, height, fx, fy, cx, cy, nboxes, _p(out), _p(origin))
    return out, origin
