"""The oracle: ctypes access to the CPU restatement (oracle/liboracle.so) and,
where it was built, to the reference's own VoxelGrid (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and the CPU baseline.  The GPU
library never loads it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_grid.so")

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_D = ctypes.c_double
_I = ctypes.c_int
_lib = None
_ref = None


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(LIB_PATH)
        sig = {
            "bxo_grid_create": (_P, [_D, _U32, _U32, _U32]),
            "bxo_grid_destroy": (None, [_P]),
            "bxo_grid_clear_memory": (None, [_P]),
            "bxo_grid_clear_cells": (None, [_P]),
            "bxo_grid_release_unused": (None, [_P]),
            "bxo_set_values": (None, [_P, _P, _P, _U64, _P]),
            "bxo_set_cells_on": (None, [_P, _P, _P, _U64, _P]),
            "bxo_set_cells_off": (None, [_P, _P, _U64, _P]),
            "bxo_lookup": (None, [_P, _P, _U64, _P, _P]),
            "bxo_active_cells": (_U64, [_P]),
            "bxo_root_count": (_U64, [_P]),
            "bxo_leaf_count": (_U64, [_P]),
            "bxo_export": (_U64, [_P, _I, ctypes.c_int32, _P, _P, _U64]),
            "bxo_pos_to_coord": (None, [_P, _U64, _D, _P]),
            "bxo_digest": (None, [_P, _P]),
            "bxo_prob_create": (_P, [_D]),
            "bxo_prob_destroy": (None, [_P]),
            "bxo_prob_grid": (_P, [_P]),
            "bxo_prob_set_options": (None, [_P, _P]),
            "bxo_prob_update_count": (_U32, [_P]),
            "bxo_prob_set_update_count": (None, [_P, _U32]),
            "bxo_prob_stats": (None, [_P, _P]),
            "bxo_prob_add_hit": (None, [_P, _P]),
            "bxo_prob_add_miss": (None, [_P, _P]),
            "bxo_prob_update_free_cells": (None, [_P, _P]),
            "bxo_prob_insert_f32": (None, [_P, _P, _U64, _P, _D]),
            "bxo_prob_insert_tf": (None, [_P, _P, _U64, _U64, _P, _D]),
            "bxo_tf_ambiguity": (_U64, [_P, _U64, _U64, _P, _D, _D]),
            "bxo_prob_insert_f64": (None, [_P, _P, _U64, _P, _D]),
            "bxo_endpoints_f32": (_U64, [_P, _U64, _P, _D, _D, _P, _P]),
            "bxo_prob_classify": (None, [_P, _P, _U64, _P]),
            "bxo_compute_ray": (_U64, [_P, _P, _P, _U64]),
        }
        for k, (res, args) in sig.items():
            f = getattr(h, k)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's VoxelGrid, compiled from its own headers (dev container only)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            raise ImportError(f"{REF_PATH} missing: run `make -C oracle ref` where the reference is present")
        h = ctypes.CDLL(REF_PATH)
        sig = {
            "ref_grid_create": (_P, [_I, _D, _I, _I]),
            "ref_grid_destroy": (None, [_P]),
            "ref_set_values": (None, [_P, _P, _P, _U64, _P]),
            "ref_set_cells_on": (None, [_P, _P, _P, _U64, _P]),
            "ref_set_cells_off": (None, [_P, _P, _U64, _P]),
            "ref_lookup": (None, [_P, _P, _U64, _P, _P]),
            "ref_export": (_U64, [_P, _P, _P, _U64]),
            "ref_active_cells": (_U64, [_P]),
            "ref_serialize": (_U64, [_P, _P, _U64]),
            "ref_deserialize": (_I, [_P, _P, _U64]),
            "ref_clear": (None, [_P, _I]),
            "ref_release_unused": (None, [_P]),
            "ref_pos_to_coord": (None, [_D, _P, _U64, _P]),
            "ref_coord_hash": (None, [_P, _U64, _P]),
            "ref_prob_create": (_P, [_D]),
            "ref_prob_destroy": (None, [_P]),
            "ref_prob_insert_f32": (None, [_P, _P, _U64, _P, _D]),
            "ref_prob_export": (_U64, [_P, _P, _P, _U64]),
            "ref_prob_update_count": (_U32, [_P]),
        }
        for k, (res, args) in sig.items():
            f = getattr(h, k)
            f.restype = res
            f.argtypes = args
        _ref = h
    return _ref


def _c(a):
    return np.ascontiguousarray(a, np.int32)


def _sorted_cells(coords, vals):
    order = np.lexsort((coords[:, 2], coords[:, 1], coords[:, 0]))
    return coords[order], (vals[order] if vals is not None else None)


class Grid:
    """VoxelGrid<T> restatement (value_bytes = sizeof(T), 0 for binary)."""

    def __init__(self, voxel_size, inner_bits=2, leaf_bits=3, dtype=np.float32, _handle=None):
        self.dtype = np.dtype(dtype) if dtype is not None else None
        self.vb = self.dtype.itemsize if self.dtype is not None else 0
        self.voxel_size = voxel_size
        self._own = _handle is None
        self.h = _handle or lib().bxo_grid_create(voxel_size, inner_bits, leaf_bits, self.vb)

    def close(self):
        if self._own and self.h:
            lib().bxo_grid_destroy(self.h)
        self.h = None

    def set_values(self, coords, vals):
        c = _c(coords)
        v = np.ascontiguousarray(vals, self.dtype) if self.vb else None
        was = np.empty(len(c), np.uint8)
        lib().bxo_set_values(self.h, _p(c), _p(v), len(c), _p(was))
        return was.astype(bool)

    def set_cells_on(self, coords, defaults=None):
        c = _c(coords)
        v = np.ascontiguousarray(defaults, self.dtype) if (self.vb and defaults is not None) else (
            np.zeros(len(c), self.dtype) if self.vb else None)
        was = np.empty(len(c), np.uint8)
        lib().bxo_set_cells_on(self.h, _p(c), _p(v), len(c), _p(was))
        return was.astype(bool)

    def set_cells_off(self, coords):
        c = _c(coords)
        was = np.empty(len(c), np.uint8)
        lib().bxo_set_cells_off(self.h, _p(c), len(c), _p(was))
        return was.astype(bool)

    def lookup(self, coords):
        c = _c(coords)
        vals = np.empty(len(c), self.dtype) if self.vb else None
        found = np.empty(len(c), np.uint8)
        lib().bxo_lookup(self.h, _p(c), len(c), _p(vals), _p(found))
        return vals, found.astype(bool)

    def export(self, filter=0, threshold=0, sort=True):
        n = lib().bxo_export(self.h, filter, threshold, None, None, 0)
        coords = np.empty((n, 3), np.int32)
        vals = np.empty(n, self.dtype) if self.vb else None
        lib().bxo_export(self.h, filter, threshold, _p(coords), _p(vals), n)
        return _sorted_cells(coords, vals) if sort else (coords, vals)

    def active_cells_count(self):
        return lib().bxo_active_cells(self.h)

    def leaf_count(self):
        return lib().bxo_leaf_count(self.h)

    def root_count(self):
        return lib().bxo_root_count(self.h)

    def digest(self):
        out = np.zeros(4, np.uint64)
        lib().bxo_digest(self.h, _p(out))
        return tuple(int(x) for x in out)

    def clear(self, option=0):
        if option == 0:
            lib().bxo_grid_clear_memory(self.h)
        else:
            lib().bxo_grid_clear_cells(self.h)

    def release_unused(self):
        lib().bxo_grid_release_unused(self.h)


class ProbMap:
    """ProbabilisticMap restatement (sequential, as the reference)."""

    def __init__(self, resolution):
        self.h = lib().bxo_prob_create(resolution)
        self.grid = Grid(resolution, dtype=np.uint32, _handle=lib().bxo_prob_grid(self.h))
        self.resolution = resolution

    def close(self):
        if self.h:
            lib().bxo_prob_destroy(self.h)
            self.h = None

    def set_options(self, miss, hit, cmin, cmax, thr):
        o = np.array([miss, hit, cmin, cmax, thr], np.int32)
        lib().bxo_prob_set_options(self.h, _p(o))

    @property
    def update_count(self):
        return lib().bxo_prob_update_count(self.h)

    @update_count.setter
    def update_count(self, v):
        lib().bxo_prob_set_update_count(self.h, v)

    def insert_point_cloud(self, points, origin, max_range):
        o = np.ascontiguousarray(origin, np.float64)
        p = np.ascontiguousarray(points)
        if p.dtype == np.float64:
            lib().bxo_prob_insert_f64(self.h, _p(p), len(p), _p(o), float(max_range))
        else:
            p = np.ascontiguousarray(p, np.float32)
            lib().bxo_prob_insert_f32(self.h, _p(p), len(p), _p(o), float(max_range))

    def insert_point_cloud_tf(self, points, sensor_to_world, max_range):
        """The ingest pre-step (finite filter, float transform, origin from the
        matrix) then insertPointCloud; points (n, k >= 3) float32, x y z first."""
        p = np.ascontiguousarray(points, np.float32)
        t = np.ascontiguousarray(sensor_to_world, np.float32).reshape(16)
        lib().bxo_prob_insert_tf(self.h, _p(p), len(p), p.shape[1], _p(t), float(max_range))

    def add_hit_point(self, xyz):
        lib().bxo_prob_add_hit(self.h, _p(np.ascontiguousarray(xyz, np.float64)))

    def add_miss_point(self, xyz):
        lib().bxo_prob_add_miss(self.h, _p(np.ascontiguousarray(xyz, np.float64)))

    def stats(self):
        out = np.zeros(4, np.uint64)
        lib().bxo_prob_stats(self.h, _p(out))
        return {"endpoint_updates": int(out[0]), "rays": int(out[1]), "visits": int(out[2]),
                "ray_updates": int(out[3])}

    def cells(self):
        return self.grid.export()

    def classify(self, coords):
        c = _c(coords)
        out = np.empty(len(c), np.int8)
        lib().bxo_prob_classify(self.h, _p(c), len(c), _p(out))
        return out

    def occupied(self):
        return self.grid.export(1, 0)[0]

    def free(self):
        return self.grid.export(2, 0)[0]

    def digest(self):
        return self.grid.digest()


def endpoints(points, origin, max_range, resolution):
    """(coords (n,3), kind (n,) 1 hit / 0 miss, ambiguous count) of a scan."""
    p = np.ascontiguousarray(points, np.float32)
    o = np.ascontiguousarray(origin, np.float64)
    c = np.empty((len(p), 3), np.int32)
    k = np.empty(len(p), np.uint8)
    amb = lib().bxo_endpoints_f32(_p(p), len(p), _p(o), float(max_range), 1.0 / resolution, _p(c), _p(k))
    return c, k, int(amb)


def tf_ambiguity(points, sensor_to_world, max_range, resolution):
    """Points whose endpoint differs between the two float associations of
    the ingest transform."""
    p = np.ascontiguousarray(points, np.float32)
    t = np.ascontiguousarray(sensor_to_world, np.float32).reshape(16)
    return int(lib().bxo_tf_ambiguity(_p(p), len(p), p.shape[1], _p(t), float(max_range), 1.0 / resolution))


def compute_ray(o, e):
    o = np.ascontiguousarray(o, np.int32)
    e = np.ascontiguousarray(e, np.int32)
    n = lib().bxo_compute_ray(_p(o), _p(e), None, 0)
    out = np.empty((n, 3), np.int32)
    lib().bxo_compute_ray(_p(o), _p(e), _p(out), n)
    return out


class RefGrid:
    """The reference's VoxelGrid<T> (kind 0 float, 1 int32, 2 uint32, 3 int64,
    4 binary, 5 CellT-layout)."""

    KIND = {np.dtype(np.float32): 0, np.dtype(np.int32): 1, np.dtype(np.uint32): 2,
            np.dtype(np.int64): 3}

    def __init__(self, voxel_size, dtype=np.float32, inner_bits=2, leaf_bits=3):
        self.dtype = np.dtype(dtype) if dtype is not None else None
        kind = 4 if dtype is None else self.KIND[self.dtype]
        self.h = ref().ref_grid_create(kind, voxel_size, inner_bits, leaf_bits)

    def close(self):
        if self.h:
            ref().ref_grid_destroy(self.h)
            self.h = None

    def set_values(self, coords, vals):
        c = _c(coords)
        v = np.ascontiguousarray(vals, self.dtype) if self.dtype is not None else None
        was = np.empty(len(c), np.uint8)
        ref().ref_set_values(self.h, _p(c), _p(v), len(c), _p(was))
        return was.astype(bool)

    def set_cells_on(self, coords, defaults=None):
        c = _c(coords)
        v = None
        if self.dtype is not None:
            v = np.ascontiguousarray(defaults if defaults is not None else np.zeros(len(c)), self.dtype)
        was = np.empty(len(c), np.uint8)
        ref().ref_set_cells_on(self.h, _p(c), _p(v), len(c), _p(was))
        return was.astype(bool)

    def set_cells_off(self, coords):
        c = _c(coords)
        was = np.empty(len(c), np.uint8)
        ref().ref_set_cells_off(self.h, _p(c), len(c), _p(was))
        return was.astype(bool)

    def lookup(self, coords):
        c = _c(coords)
        vals = np.empty(len(c), self.dtype) if self.dtype is not None else None
        found = np.empty(len(c), np.uint8)
        ref().ref_lookup(self.h, _p(c), len(c), _p(vals), _p(found))
        return vals, found.astype(bool)

    def export(self, sort=True):
        n = ref().ref_export(self.h, None, None, 0)
        coords = np.empty((n, 3), np.int32)
        vals = np.empty(n, self.dtype) if self.dtype is not None else None
        ref().ref_export(self.h, _p(coords), _p(vals), n)
        return _sorted_cells(coords, vals) if sort else (coords, vals)

    def active_cells_count(self):
        return ref().ref_active_cells(self.h)

    def serialize(self):
        n = ref().ref_serialize(self.h, None, 0)
        buf = np.empty(n, np.uint8)
        ref().ref_serialize(self.h, _p(buf), n)
        return buf.tobytes()

    def clear(self, option=0):
        ref().ref_clear(self.h, option)

    def release_unused(self):
        ref().ref_release_unused(self.h)


class RefProb:
    """Scan integration restated over the reference's VoxelGrid<CellT>."""

    def __init__(self, resolution):
        self.h = ref().ref_prob_create(resolution)

    def close(self):
        if self.h:
            ref().ref_prob_destroy(self.h)
            self.h = None

    def insert_point_cloud(self, points, origin, max_range):
        p = np.ascontiguousarray(points, np.float32)
        o = np.ascontiguousarray(origin, np.float64)
        ref().ref_prob_insert_f32(self.h, _p(p), len(p), _p(o), float(max_range))

    def cells(self):
        n = ref().ref_prob_export(self.h, None, None, 0)
        coords = np.empty((n, 3), np.int32)
        words = np.empty(n, np.uint32)
        ref().ref_prob_export(self.h, _p(coords), _p(words), n)
        return _sorted_cells(coords, words)

    @property
    def update_count(self):
        return ref().ref_prob_update_count(self.h)
