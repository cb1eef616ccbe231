"""bonxai_amd -- Python host side of the MI355X sparse voxel map.

A thin ctypes layer over the C ABI in include/bonxai_gpu.h
(bonxai_amd/libbonxai_gpu.so).  It mirrors the reference's
VoxelGrid<T> / ProbabilisticMap interface with bulk (array) calls:

    reference (C++, per cell)                    here (per batch)
    VoxelGrid<T>::Accessor::setValue             VoxelGrid.set_values
    Accessor::setCellOn / setCellOff             VoxelGrid.set_cells_on / set_cells_off
    ConstAccessor::value / isCellOn              VoxelGrid.lookup
    VoxelGrid::forEachCell                       VoxelGrid.export (all on cells)
    VoxelGrid::activeCellsCount / memUsage       active_cells_count / mem_usage
    VoxelGrid::clear                             clear
    ProbabilisticMap::insertPointCloud           ProbabilisticMap.insert_point_cloud
    ProbabilisticMap::getOccupied/FreeVoxels     get_occupied_voxels / get_free_voxels
    ProbabilisticMap::isOccupied/isUnknown/...   classify / is_occupied / ...

Inputs may be numpy arrays (host; staged by the library) or DeviceArray
(device memory of the context).  There is no CPU fallback: without the built
HIP library, importing a class that needs it raises ImportError.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "BonxaiError", "Context", "DeviceArray", "VoxelGrid", "ProbabilisticMap", "ScanGraph",
    "insert_batch", "lib", "OK", "ERR_OUT_OF_POOL", "ERR_COORD_RANGE", "ERR_BUDGET",
    "ERR_CAPACITY", "ERR_INVALID_ARG", "ERR_MAP_FAILED", "ERR_NOT_SUPPORTED",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbonxai_gpu.so")
PROBE_LIB_PATH = os.path.join(_HERE, "libbonxai_probes.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_OUT_OF_POOL = 2
ERR_HIP = 3
ERR_COORD_RANGE = 4
ERR_BUDGET = 5
ERR_NOT_SUPPORTED = 6
ERR_CAPACITY = 7
ERR_MAP_FAILED = 8

MAP_PROBABILISTIC = 1
POINTS_F64 = 1
CLEAR_MEMORY = 0
SET_ALL_CELLS_OFF = 1
EXPORT_ALL, EXPORT_OCCUPIED, EXPORT_FREE = 0, 1, 2

# The C ABI, one entry per exported function: (name, restype, argtypes).
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_I32 = ctypes.c_int32
_INT = ctypes.c_int
_D = ctypes.c_double
_PU64 = ctypes.POINTER(ctypes.c_uint64)
_PU32 = ctypes.POINTER(ctypes.c_uint32)
_PD = ctypes.POINTER(ctypes.c_double)


class ProbOptions(ctypes.Structure):
    """ProbabilisticMap::Options (probabilistic_map.hpp:56-64)."""
    _fields_ = [("prob_miss_log", ctypes.c_int32), ("prob_hit_log", ctypes.c_int32),
                ("clamp_min_log", ctypes.c_int32), ("clamp_max_log", ctypes.c_int32),
                ("occupancy_threshold_log", ctypes.c_int32)]


class MapStats(ctypes.Structure):
    _fields_ = [("leaf_count", ctypes.c_uint64), ("leaf_capacity", ctypes.c_uint64),
                ("root_count", ctypes.c_uint64), ("root_capacity", ctypes.c_uint64),
                ("scans_applied", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("last_rays", ctypes.c_uint64), ("last_ray_updates", ctypes.c_uint64),
                ("update_count", ctypes.c_uint32), ("status", ctypes.c_uint32),
                ("last_ray_visits", ctypes.c_uint64), ("replays", ctypes.c_uint64)]


class ScanDesc(ctypes.Structure):
    _fields_ = [("map", ctypes.c_void_p), ("points_dev", ctypes.c_void_p), ("n", ctypes.c_uint64),
                ("flags", ctypes.c_uint32), ("stride", ctypes.c_uint32),
                ("origin", ctypes.c_double * 3), ("max_range", ctypes.c_double),
                ("sensor_to_world", ctypes.c_void_p)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("max_inflight", ctypes.c_uint32), ("max_batch_scans", ctypes.c_uint32),
                ("streams", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("kind", ctypes.c_uint32), ("count", ctypes.c_uint64),
                ("submit_ns", ctypes.c_uint64), ("admit_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "failed", "retried", "deferred",
                                               "waves", "wave_scans", "queued", "inflight")]


_PPTR = ctypes.POINTER(ctypes.c_void_p)

ABI = [
    ("bxg_status_string", ctypes.c_char_p, [_INT]),
    ("bxg_ctx_create", _INT, [_INT, _U64, _PPTR]),
    ("bxg_ctx_destroy", _INT, [_P]),
    ("bxg_ctx_sync", _INT, [_P]),
    ("bxg_ctx_last_error", ctypes.c_char_p, [_P]),
    ("bxg_ctx_stream", _INT, [_P, _PPTR]),
    ("bxg_ctx_memory", _INT, [_P, _PU64, _PU64]),
    ("bxg_ctx_device_info", _INT, [_P, ctypes.c_char_p, _U64, ctypes.POINTER(_INT), _PU64]),
    ("bxg_malloc", _INT, [_P, _U64, _PPTR]),
    ("bxg_free", _INT, [_P, _P]),
    ("bxg_memcpy_h2d", _INT, [_P, _P, _P, _U64]),
    ("bxg_memcpy_d2h", _INT, [_P, _P, _P, _U64]),
    ("bxg_memcpy_d2d", _INT, [_P, _P, _P, _U64]),
    ("bxg_event_record", _INT, [_P, _U32]),
    ("bxg_event_elapsed", _INT, [_P, _U32, _U32, ctypes.POINTER(ctypes.c_float)]),
    ("bxg_profile_enable", _INT, [_P, _INT]),
    ("bxg_profile_read", _INT, [_P, ctypes.c_char_p, _PU64, _PD]),
    ("bxg_profile_reset", _INT, [_P]),
    ("bxg_set_counting", _INT, [_P, _INT]),
    ("bxg_ctx_set_tuning", _INT, [_P, ctypes.c_char_p, ctypes.c_int64]),
    ("bxg_ctx_get_tuning", _INT, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    ("bxg_map_create", _INT, [_P, _D, _U32, _U32, _U32, _U32, _PPTR]),
    ("bxg_map_destroy", _INT, [_P]),
    ("bxg_map_reserve", _INT, [_P, _U64, _U64]),
    ("bxg_map_clear", _INT, [_P, _INT]),
    ("bxg_map_release_unused", _INT, [_P]),
    ("bxg_map_sync", _INT, [_P]),
    ("bxg_map_clear_error", _INT, [_P]),
    ("bxg_map_get_stats", _INT, [_P, ctypes.POINTER(MapStats)]),
    ("bxg_map_active_cells", _INT, [_P, _PU64]),
    ("bxg_map_mem_usage", _INT, [_P, _PU64]),
    ("bxg_map_digest", _INT, [_P, _PU64]),
    ("bxg_map_voxel_size", _D, [_P]),
    ("bxg_set_values", _INT, [_P, _P, _P, _U64, _P]),
    ("bxg_set_cells_on", _INT, [_P, _P, _P, _U64, _P]),
    ("bxg_set_cells_off", _INT, [_P, _P, _U64, _P]),
    ("bxg_lookup", _INT, [_P, _P, _U64, _P, _P]),
    ("bxg_export", _INT, [_P, _INT, _P, _P, _U64, _PU64]),
    ("bxg_set_values_host", _INT, [_P, _P, _P, _U64, _P]),
    ("bxg_set_cells_on_host", _INT, [_P, _P, _P, _U64, _P]),
    ("bxg_set_cells_off_host", _INT, [_P, _P, _U64, _P]),
    ("bxg_lookup_host", _INT, [_P, _P, _U64, _P, _P]),
    ("bxg_export_host", _INT, [_P, _INT, _P, _P, _U64, _PU64]),
    ("bxg_read_leaves_host", _INT, [_P, _P, _U64, _P, _P, _P]),
    ("bxg_write_leaves_host", _INT, [_P, _P, _U64, _P, _P]),
    ("bxg_leaf_origins_host", _INT, [_P, _P, _U64, _PU64]),
    ("bxg_set_options", _INT, [_P, ctypes.POINTER(ProbOptions)]),
    ("bxg_get_options", _INT, [_P, ctypes.POINTER(ProbOptions)]),
    ("bxg_get_update_count", _INT, [_P, _PU32]),
    ("bxg_set_update_count", _INT, [_P, _U32]),
    ("bxg_insert_point_cloud", _INT, [_P, _P, _U64, _U32, _PD, _D]),
    ("bxg_insert_point_cloud_host", _INT, [_P, _P, _U64, _U32, _PD, _D]),
    ("bxg_insert_batch", _INT, [_P, ctypes.POINTER(ScanDesc), _U32]),
    ("bxg_add_point_host", _INT, [_P, _PD, _INT]),
    ("bxg_insert_point_cloud_tf", _INT, [_P, _P, _U64, _U32, _P, ctypes.c_double]),
    ("bxg_insert_point_cloud_tf_host", _INT, [_P, _P, _U64, _U32, _P, ctypes.c_double]),
    ("bxg_insert_point_cloud_noalloc", _INT, [_P, _P, _U64, _U32, _PD, _D]),
    ("bxg_classify", _INT, [_P, _P, _U64, _P]),
    ("bxg_classify_host", _INT, [_P, _P, _U64, _P]),
    ("bxg_graph_create", _INT, [_P, _U32, _U64, _PPTR]),
    ("bxg_graph_set_scans", _INT, [_P, ctypes.POINTER(ScanDesc), _U32]),
    ("bxg_graph_launch", _INT, [_P]),
    ("bxg_graph_destroy", _INT, [_P]),
    ("bxg_server_create", _INT, [_P, ctypes.POINTER(ServerOptions), _PPTR]),
    ("bxg_server_destroy", _INT, [_P]),
    ("bxg_server_drain", _INT, [_P]),
    ("bxg_server_insert", _INT, [_P, _P, _P, _U64, _U32, _PD, ctypes.c_double, _PU64]),
    ("bxg_server_lookup", _INT, [_P, _P, _P, _U64, _P, _P, _PU64]),
    ("bxg_server_export", _INT, [_P, _P, _INT, _P, _P, _U64, _PU64]),
    ("bxg_server_wait", _INT, [_P, _U64, ctypes.POINTER(RequestInfo)]),
    ("bxg_server_query", _INT, [_P, _U64, ctypes.POINTER(RequestInfo)]),
    ("bxg_server_get_stats", _INT, [_P, ctypes.POINTER(ServerStats)]),
]

# measurement probes (include/bonxai_gpu_probes.h), a library of their own
PROBE_ABI = [
    ("bxg_fp64_probe", _INT, [_P, _P, _P, _P, _U64, _P]),
    ("bxg_copy_probe", _INT, [_P, _P, _P, _U64]),
    ("bxg_gather_probe", _INT, [_P, _P, _U64, _U64, _P]),
]

_lib = None
_probes = None


def probes():
    """The loaded probe library (ImportError if it has not been built)."""
    global _probes
    if _probes is None:
        lib()  # the probe library links against the map library
        if not os.path.exists(PROBE_LIB_PATH):
            raise ImportError(f"{PROBE_LIB_PATH} is missing: build it with __graft_entry__.build()")
        handle = ctypes.CDLL(PROBE_LIB_PATH)
        for name, res, args in PROBE_ABI:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _probes = handle
    return _probes


def lib():
    """The loaded HIP library (ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in ABI:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class BonxaiError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def _check(rc, ctx=None, what=""):
    if rc != OK:
        msg = lib().bxg_status_string(rc).decode()
        if ctx is not None:
            detail = lib().bxg_ctx_last_error(ctx._h)
            if detail:
                msg += ": " + detail.decode()
        raise BonxaiError(rc, (what + ": " if what else "") + msg)
    return rc


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceArray:
    """A device allocation of the context with a numpy dtype and shape."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        out = ctypes.c_void_p()
        _check(lib().bxg_malloc(ctx._h, max(self.nbytes, 1), ctypes.byref(out)), ctx, "malloc")
        self.ptr = out.value

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        _check(lib().bxg_memcpy_h2d(self.ctx._h, self.ptr, _ptr(a), a.nbytes), self.ctx, "h2d")

    def numpy(self, count=None):
        shape = self.shape if count is None else (count,) + self.shape[1:]
        out = np.empty(shape, self.dtype)
        if out.nbytes:
            _check(lib().bxg_memcpy_d2h(self.ctx._h, _ptr(out), self.ptr, out.nbytes), self.ctx, "d2h")
        return out

    def free(self):
        if self.ptr and not getattr(self, "_view", False):
            lib().bxg_free(self.ctx._h, self.ptr)
        self.ptr = None

    def view(self, count):
        """The first `count` rows, sharing this allocation (not owning it)."""
        if count > self.shape[0]:
            raise ValueError("view longer than the array")
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.dtype, v.ptr, v._view = self.ctx, self.dtype, self.ptr, True
        v.shape = (int(count),) + self.shape[1:]
        v.nbytes = int(np.prod(v.shape)) * self.dtype.itemsize
        return v

    def __len__(self):
        return self.shape[0]


class Context:
    """One GPU: stream, scratch, memory budget (bytes, 0 = unlimited).  With
    tuned=True the launch setting recorded for this device by
    bonxai_amd.autotune (bonxai_amd/tuning.json) is applied, if there is one."""

    def __init__(self, device=0, budget_bytes=0, tuned=False):
        h = ctypes.c_void_p()
        rc = lib().bxg_ctx_create(device, budget_bytes, ctypes.byref(h))
        if rc != OK:
            raise BonxaiError(rc, "context creation failed on device %d: %s"
                              % (device, lib().bxg_status_string(rc).decode()))
        self._h = h.value
        self.tuning = None
        if tuned:
            from . import autotune
            self.tuning = autotune.load(self)
            if self.tuning:
                autotune.apply(self, self.tuning)

    def close(self):
        if self._h:
            lib().bxg_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        _check(lib().bxg_ctx_sync(self._h), self, "sync")

    def last_error(self):
        """The message of the context's last failed call."""
        msg = lib().bxg_ctx_last_error(self._h)
        return msg.decode() if msg else ""

    def memory(self):
        used, budget = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().bxg_ctx_memory(self._h, ctypes.byref(used), ctypes.byref(budget)), self)
        return used.value, budget.value

    def device_info(self):
        name = ctypes.create_string_buffer(128)
        cu = ctypes.c_int()
        hbm = ctypes.c_uint64()
        _check(lib().bxg_ctx_device_info(self._h, name, 128, ctypes.byref(cu), ctypes.byref(hbm)), self)
        return {"arch": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value}

    def array(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        return DeviceArray.from_numpy(self, a)

    # timing on the context stream
    def event(self, slot):
        _check(lib().bxg_event_record(self._h, slot), self)

    def elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        _check(lib().bxg_event_elapsed(self._h, a, b, ctypes.byref(ms)), self)
        return ms.value

    def profile(self, on=True):
        _check(lib().bxg_profile_enable(self._h, 1 if on else 0), self)

    def profile_reset(self):
        _check(lib().bxg_profile_reset(self._h), self)

    def profile_read(self, kernel):
        n, ms = ctypes.c_uint64(), ctypes.c_double()
        _check(lib().bxg_profile_read(self._h, kernel.encode(), ctypes.byref(n), ctypes.byref(ms)), self)
        return n.value, ms.value

    def set_tuning(self, key, value):
        _check(lib().bxg_ctx_set_tuning(self._h, key.encode(), int(value)), self, "tuning")

    def get_tuning(self, key):
        v = ctypes.c_int64()
        _check(lib().bxg_ctx_get_tuning(self._h, key.encode(), ctypes.byref(v)), self, "tuning")
        return v.value

    def counting(self, on=True):
        _check(lib().bxg_set_counting(self._h, 1 if on else 0), self)

    def copy_probe(self, nbytes, reps=1):
        """Run the plain copy kernel over nbytes (calibration of byte counters)."""
        src = self.array(nbytes // 16, np.dtype((np.uint32, 4)))
        dst = self.array(nbytes // 16, np.dtype((np.uint32, 4)))
        try:
            for _ in range(reps):
                _check(probes().bxg_copy_probe(self._h, src.ptr, dst.ptr, nbytes), self, "copy probe")
            self.sync()
        finally:
            src.free()
            dst.free()

    def gather_probe(self, table_bytes, n, reps=1, table=None):
        """Run n scattered 4-byte reads over a table of table_bytes (a power of
        two); returns the words read, as a device array the caller frees."""
        own = table is None
        if not own and table.nbytes < table_bytes:
            raise ValueError("gather_probe: table_bytes exceeds the table's size")
        if own:
            table = self.array(table_bytes // 4, np.uint32)
        out = self.array(n, np.uint32)
        try:
            for _ in range(reps):
                _check(probes().bxg_gather_probe(self._h, table.ptr, table_bytes, n, out.ptr), self,
                       "gather probe")
            self.sync()
        except BaseException:
            out.free()
            raise
        finally:
            if own:
                table.free()
        return out

    def fp64_probe(self, a, b, c):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        c = np.ascontiguousarray(c, np.float64)
        n = a.size
        da, db, dc = self.to_device(a), self.to_device(b), self.to_device(c)
        out = self.array((n, 5), np.float64)
        try:
            _check(probes().bxg_fp64_probe(self._h, da.ptr, db.ptr, dc.ptr, n, out.ptr), self, "fp64 probe")
            return out.numpy()
        finally:
            for d in (da, db, dc, out):
                d.free()


def _coords(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("coordinates must be an (n, 3) int32 array")
    return a


class VoxelGrid:
    """Bonxai::VoxelGrid<T> (bonxai.hpp:114-333) resident on the GPU.

    dtype is the cell type T (any fixed-size numpy dtype; None for the
    BinaryVoxelGrid of bonxai.hpp:335)."""

    def __init__(self, ctx, voxel_size, inner_bits=2, leaf_bits=3, dtype=np.float32, _flags=0):
        self.ctx = ctx
        self.dtype = np.dtype(dtype) if dtype is not None else None
        self.value_bytes = self.dtype.itemsize if self.dtype is not None else 0
        self.inner_bits, self.leaf_bits = inner_bits, leaf_bits
        h = ctypes.c_void_p()
        _check(lib().bxg_map_create(ctx._h, voxel_size, inner_bits, leaf_bits, self.value_bytes,
                                    _flags, ctypes.byref(h)), ctx, "VoxelGrid")
        self._h = h.value
        self.voxel_size = voxel_size

    def close(self):
        if self._h:
            lib().bxg_map_destroy(self._h)
            self._h = None

    # ---- conversions (bonxai.hpp:404-417)
    def pos_to_coord(self, xyz):
        xyz = np.asarray(xyz, np.float64)
        return np.floor(xyz * (1.0 / self.voxel_size)).astype(np.int32)

    def coord_to_pos(self, c):
        return np.asarray(c, np.float64) * self.voxel_size

    # ---- bulk cell operations
    def _vals(self, vals, n):
        if self.dtype is None:
            return None
        if vals is None:
            return None
        v = np.ascontiguousarray(vals, dtype=self.dtype)
        if v.shape[0] != n:
            raise ValueError("one value per coordinate")
        return v

    def set_values(self, coords, vals, return_was_on=False):
        if isinstance(coords, DeviceArray):
            was = self.ctx.array(len(coords), np.uint8) if return_was_on else None
            _check(lib().bxg_set_values(self._h, coords.ptr, vals.ptr if vals is not None else None,
                                        len(coords), was.ptr if was else None), self.ctx, "set_values")
            return was
        c = _coords(coords)
        v = self._vals(vals, len(c))
        was = np.empty(len(c), np.uint8) if return_was_on else None
        _check(lib().bxg_set_values_host(self._h, _ptr(c), _ptr(v), len(c), _ptr(was)), self.ctx, "set_values")
        return was.astype(bool) if return_was_on else None

    def set_cells_on(self, coords, defaults=None, return_was_on=False):
        c = _coords(coords)
        v = self._vals(defaults, len(c))
        was = np.empty(len(c), np.uint8) if return_was_on else None
        _check(lib().bxg_set_cells_on_host(self._h, _ptr(c), _ptr(v), len(c), _ptr(was)), self.ctx, "set_cells_on")
        return was.astype(bool) if return_was_on else None

    def set_cells_off(self, coords, return_was_on=False):
        c = _coords(coords)
        was = np.empty(len(c), np.uint8) if return_was_on else None
        _check(lib().bxg_set_cells_off_host(self._h, _ptr(c), len(c), _ptr(was)), self.ctx, "set_cells_off")
        return was.astype(bool) if return_was_on else None

    def lookup(self, coords, out_vals=None, out_found=None):
        """Values (zero where absent) and found flags of a batch of coords."""
        if isinstance(coords, DeviceArray):
            _check(lib().bxg_lookup(self._h, coords.ptr, len(coords),
                                    out_vals.ptr if out_vals is not None else None,
                                    out_found.ptr if out_found is not None else None), self.ctx, "lookup")
            return out_vals, out_found
        c = _coords(coords)
        vals = np.empty(len(c), self.dtype) if self.dtype is not None else None
        found = np.empty(len(c), np.uint8)
        _check(lib().bxg_lookup_host(self._h, _ptr(c), len(c), _ptr(vals), _ptr(found)), self.ctx, "lookup")
        return vals, found.astype(bool)

    def export(self, filter=EXPORT_ALL):
        """forEachCell as arrays: (coords (n,3) int32, values (n,) dtype)."""
        cnt = ctypes.c_uint64()
        rc = lib().bxg_export_host(self._h, filter, None, None, 0, ctypes.byref(cnt))
        if rc not in (OK, ERR_CAPACITY):
            _check(rc, self.ctx, "export")
        n = cnt.value
        coords = np.empty((n, 3), np.int32)
        vals = np.empty(n, self.dtype) if self.dtype is not None else None
        _check(lib().bxg_export_host(self._h, filter, _ptr(coords), _ptr(vals), n, ctypes.byref(cnt)),
               self.ctx, "export")
        return coords, vals

    def active_cells_count(self):
        out = ctypes.c_uint64()
        _check(lib().bxg_map_active_cells(self._h, ctypes.byref(out)), self.ctx)
        return out.value

    def mem_usage(self):
        out = ctypes.c_uint64()
        _check(lib().bxg_map_mem_usage(self._h, ctypes.byref(out)), self.ctx)
        return out.value

    def digest(self):
        out = (ctypes.c_uint64 * 4)()
        _check(lib().bxg_map_digest(self._h, out), self.ctx, "digest")
        return tuple(int(x) for x in out)

    def clear(self, option=CLEAR_MEMORY):
        _check(lib().bxg_map_clear(self._h, option), self.ctx, "clear")

    def release_unused_memory(self):
        """releaseUnusedMemory (bonxai.hpp:367-387): leaves with every cell off
        and roots left without leaves are freed; the pools shrink."""
        _check(lib().bxg_map_release_unused(self._h), self.ctx, "release_unused_memory")

    def reserve(self, leaves, roots=0):
        _check(lib().bxg_map_reserve(self._h, leaves, roots), self.ctx, "reserve")

    def stats(self):
        s = MapStats()
        _check(lib().bxg_map_get_stats(self._h, ctypes.byref(s)), self.ctx, "stats")
        return {k: getattr(s, k) for k, _ in MapStats._fields_}

    def sync(self):
        return lib().bxg_map_sync(self._h)

    def clear_error(self):
        _check(lib().bxg_map_clear_error(self._h), self.ctx)

    def read_leaves(self, origins):
        o = _coords(origins)
        lw = max((1 << (3 * self.leaf_bits)) // 64, 1)
        ls = 1 << (3 * self.leaf_bits)
        masks = np.empty((len(o), lw), np.uint64)
        cells = np.empty((len(o), ls * self.value_bytes), np.uint8)
        exists = np.empty(len(o), np.uint8)
        _check(lib().bxg_read_leaves_host(self._h, _ptr(o), len(o), _ptr(masks), _ptr(cells), _ptr(exists)),
               self.ctx, "read_leaves")
        return masks, cells, exists.astype(bool)

    def write_leaves(self, origins, masks, cells):
        o = _coords(origins)
        masks = np.ascontiguousarray(masks, np.uint64)
        cells = np.ascontiguousarray(cells, np.uint8)
        _check(lib().bxg_write_leaves_host(self._h, _ptr(o), len(o), _ptr(masks), _ptr(cells)),
               self.ctx, "write_leaves")


def _as_points(points):
    if isinstance(points, DeviceArray):
        return points, (POINTS_F64 if points.dtype == np.float64 else 0)
    p = np.ascontiguousarray(points)
    if p.dtype not in (np.float32, np.float64):
        p = p.astype(np.float32)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("points must be an (n, 3) float array")
    return p, (POINTS_F64 if p.dtype == np.float64 else 0)


class ProbabilisticMap(VoxelGrid):
    """Bonxai::ProbabilisticMap (probabilistic_map.hpp:27-137) on the GPU."""

    UNKNOWN_PROBABILITY = 0  # logods(0.5f), probabilistic_map.cpp:8

    def __init__(self, ctx, resolution):
        super().__init__(ctx, resolution, 2, 3, np.uint32, _flags=MAP_PROBABILISTIC)

    @staticmethod
    def logods(prob):
        """int32(1e6 * log(p / (1 - p))) with p a float32 widened to double
        (probabilistic_map.hpp:34-36)."""
        p = float(np.float32(prob))
        return int(1e6 * np.log(p / (1.0 - p)))

    def options(self):
        o = ProbOptions()
        _check(lib().bxg_get_options(self._h, ctypes.byref(o)), self.ctx)
        return {k: getattr(o, k) for k, _ in ProbOptions._fields_}

    def set_options(self, **kw):
        o = ProbOptions(**{**self.options(), **kw})
        _check(lib().bxg_set_options(self._h, ctypes.byref(o)), self.ctx)

    @property
    def update_count(self):
        out = ctypes.c_uint32()
        _check(lib().bxg_get_update_count(self._h, ctypes.byref(out)), self.ctx)
        return out.value

    @update_count.setter
    def update_count(self, v):
        _check(lib().bxg_set_update_count(self._h, v), self.ctx)

    def insert_point_cloud(self, points, origin, max_range, sync=True):
        """insertPointCloud(points, origin, max_range).  Host arrays block
        (and grow the pools if needed); DeviceArray points only enqueue."""
        p, flags = _as_points(points)
        o = (ctypes.c_double * 3)(*[float(x) for x in origin])
        if isinstance(p, DeviceArray):
            _check(lib().bxg_insert_point_cloud(self._h, p.ptr, len(p), flags, o, float(max_range)),
                   self.ctx, "insert_point_cloud")
            if sync:
                _check(self.sync(), self.ctx, "insert_point_cloud")
            return
        _check(lib().bxg_insert_point_cloud_host(self._h, _ptr(p), len(p), flags, o, float(max_range)),
               self.ctx, "insert_point_cloud")

    def insert_point_cloud_noalloc(self, points, origin, max_range):
        """insertPointCloud through the allocation-free entry point: device
        points (at most 524288), the library's static scratch, the pools as
        reserved; blocks.  Returns the status code (ERR_OUT_OF_POOL when the
        pools cannot hold the scan: no cell changed)."""
        if not isinstance(points, DeviceArray):
            raise TypeError("the allocation-free insert takes device points")
        p, flags = _as_points(points)
        o = (ctypes.c_double * 3)(*[float(x) for x in origin])
        return lib().bxg_insert_point_cloud_noalloc(self._h, p.ptr, len(p), flags, o, float(max_range))

    def insert_point_cloud_tf(self, points, sensor_to_world, max_range, sync=True):
        """The reference server's ingest pre-step fused into the scan
        (bonxai_server.cpp:147-179): float points in the sensor frame, (n, k)
        with x y z first (k = 4 is pcl::PointXYZ's layout), non-finite points
        dropped, the rest moved by the 4x4 float sensor-to-world matrix, origin
        = the matrix's translation."""
        t = np.ascontiguousarray(sensor_to_world, np.float32).reshape(16)
        if isinstance(points, DeviceArray):
            if points.dtype != np.float32:
                raise ValueError("the ingest transform takes float points")
            stride = int(np.prod(points.shape[1:])) * 4
            _check(lib().bxg_insert_point_cloud_tf(self._h, points.ptr, len(points), stride, _ptr(t),
                                                   float(max_range)), self.ctx, "insert_point_cloud_tf")
            if sync:
                _check(self.sync(), self.ctx, "insert_point_cloud_tf")
            return
        p = np.ascontiguousarray(points, np.float32)
        if p.ndim != 2 or p.shape[1] < 3:
            raise ValueError("points must be (n, k >= 3) float32")
        _check(lib().bxg_insert_point_cloud_tf_host(self._h, _ptr(p), len(p), p.shape[1] * 4, _ptr(t),
                                                    float(max_range)), self.ctx, "insert_point_cloud_tf")

    def add_hit_point(self, xyz):
        p = (ctypes.c_double * 3)(*[float(x) for x in xyz])
        _check(lib().bxg_add_point_host(self._h, p, 1), self.ctx, "add_hit_point")

    def add_miss_point(self, xyz):
        p = (ctypes.c_double * 3)(*[float(x) for x in xyz])
        _check(lib().bxg_add_point_host(self._h, p, 0), self.ctx, "add_miss_point")

    def cells(self):
        """All allocated on cells: coords (n,3) and raw CellT words
        ((log-odds << 4) | (update_id & 15))."""
        return self.export(EXPORT_ALL)

    def get_occupied_voxels(self):
        return self.export(EXPORT_OCCUPIED)[0]

    def get_free_voxels(self):
        return self.export(EXPORT_FREE)[0]

    def classify(self, coords):
        """1 occupied, -1 free, 0 unknown (probabilistic_map.cpp:56-75)."""
        c = _coords(coords)
        out = np.empty(len(c), np.int8)
        _check(lib().bxg_classify_host(self._h, _ptr(c), len(c), _ptr(out)), self.ctx, "classify")
        return out

    def is_occupied(self, coords):
        return self.classify(coords) == 1

    def is_free(self, coords):
        return self.classify(coords) == -1

    def is_unknown(self, coords):
        return self.classify(coords) == 0


def _scan_descs(scans):
    """scans: (map, DeviceArray points, origin, max_range) or (map, DeviceArray
    float points (n, k >= 3), None, max_range, sensor_to_world 4x4) tuples."""
    arr = (ScanDesc * len(scans))()
    keep = []
    for i, sc in enumerate(scans):
        m, pts, origin, max_range = sc[:4]
        if not isinstance(pts, DeviceArray):
            raise TypeError("batched scans take DeviceArray points")
        arr[i].map = m._h
        arr[i].points_dev = pts.ptr
        arr[i].n = len(pts)
        arr[i].flags = POINTS_F64 if pts.dtype == np.float64 else 0
        arr[i].stride = int(np.prod(pts.shape[1:])) * pts.dtype.itemsize
        if len(sc) > 4 and sc[4] is not None:
            t = np.ascontiguousarray(sc[4], np.float32).reshape(16)
            arr[i].sensor_to_world = t.ctypes.data
            keep.append(t)
        else:
            for a in range(3):
                arr[i].origin[a] = float(origin[a])
        arr[i].max_range = float(max_range)
        keep.append(pts)
    return arr, keep


def insert_batch(ctx, scans):
    """Many scans, one per map, in one pass: scans = [(map, DeviceArray
    points, origin, max_range), ...] (enqueued; call sync on the maps)."""
    arr, _ = _scan_descs(scans)
    _check(lib().bxg_insert_batch(ctx._h, arr, len(scans)), ctx, "insert_batch")


class ScanGraph:
    """A captured HIP graph of the batched scan pipeline (one launch per
    frame of up to max_scans scans / max_points points)."""

    def __init__(self, ctx, max_scans, max_points):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().bxg_graph_create(ctx._h, max_scans, max_points, ctypes.byref(h)), ctx, "graph")
        self._h = h.value

    @staticmethod
    def prepare(scans):
        """The descriptor array of a frame, for set_scans (build it once for a
        frame that is loaded many times)."""
        return _scan_descs(scans)

    def set_scans(self, scans):
        """Load a frame: a list of scans as for insert_batch, or the result of
        prepare().  Scans still pending from the previous frame are not waited
        for."""
        if isinstance(scans, tuple):
            arr, keep = scans
        else:
            arr, keep = _scan_descs(scans)
        self._keep = keep
        _check(lib().bxg_graph_set_scans(self._h, arr, len(arr)), self.ctx, "graph frame")

    def launch(self):
        _check(lib().bxg_graph_launch(self._h), self.ctx, "graph launch")

    def close(self):
        if self._h:
            lib().bxg_graph_destroy(self._h)
            self._h = None


class Server:
    """The serving engine (bxg_server_*): maps stay resident, requests are
    admitted continuously by in-flight slots and pool memory, requests of one
    map run in submission order, requests of different maps overlap.  Inputs
    and outputs are DeviceArrays that must live until the request is waited
    for."""

    KINDS = ("insert", "lookup", "export")

    def __init__(self, ctx, max_inflight=32, max_batch_scans=16, streams=4):
        self.ctx = ctx
        o = ServerOptions(max_inflight, max_batch_scans, streams, 0)
        h = ctypes.c_void_p()
        _check(lib().bxg_server_create(ctx._h, ctypes.byref(o), ctypes.byref(h)), ctx, "server")
        self._h = h.value

    def insert(self, m, points, origin, max_range):
        if not isinstance(points, DeviceArray):
            raise TypeError("server requests take DeviceArray inputs")
        flags = POINTS_F64 if points.dtype == np.float64 else 0
        o = (ctypes.c_double * 3)(*[float(x) for x in origin])
        rid = ctypes.c_uint64()
        _check(lib().bxg_server_insert(self._h, m._h, points.ptr, points.shape[0], flags, o,
                                       float(max_range), ctypes.byref(rid)), self.ctx, "server insert")
        return rid.value

    def lookup(self, m, coords, vals=None, found=None):
        rid = ctypes.c_uint64()
        _check(lib().bxg_server_lookup(self._h, m._h, coords.ptr, coords.shape[0],
                                       vals.ptr if vals is not None else None,
                                       found.ptr if found is not None else None, ctypes.byref(rid)),
               self.ctx, "server lookup")
        return rid.value

    def export(self, m, coords, vals=None, filter=EXPORT_ALL):
        rid = ctypes.c_uint64()
        _check(lib().bxg_server_export(self._h, m._h, filter, coords.ptr,
                                       vals.ptr if vals is not None else None, coords.shape[0],
                                       ctypes.byref(rid)), self.ctx, "server export")
        return rid.value

    @staticmethod
    def _info(i):
        return {"status": i.status, "kind": Server.KINDS[i.kind], "count": i.count,
                "submit_ns": i.submit_ns, "admit_ns": i.admit_ns, "done_ns": i.done_ns,
                "latency_ms": (i.done_ns - i.submit_ns) / 1e6 if i.done_ns else None}

    def wait(self, rid):
        """Block until request rid is done; its info dict (status 0 = ok)."""
        i = RequestInfo()
        lib().bxg_server_wait(self._h, rid, ctypes.byref(i))
        return self._info(i)

    def query(self, rid):
        i = RequestInfo()
        _check(lib().bxg_server_query(self._h, rid, ctypes.byref(i)), self.ctx, "server query")
        return self._info(i)

    def drain(self):
        _check(lib().bxg_server_drain(self._h), self.ctx, "server drain")

    def stats(self):
        s = ServerStats()
        _check(lib().bxg_server_get_stats(self._h, ctypes.byref(s)), self.ctx, "server stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib().bxg_server_destroy(self._h)
            self._h = None
