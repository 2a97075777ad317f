"""ctypes binding of libwtet_amd.so, mirroring the reference's ``wtet._core``
PyO3 module (src/python.rs:13-402): ``delaunay``, ``Config``,
``DelaunayResult``, ``check_quality``, ``initialize_gpu``, ``gpu_info``,
``encode_adjacency``, ``decode_adjacency``, ``INVALID``, ``__version__``.

The HIP library is required: importing works without it, but every call
that needs the device raises ``RuntimeError`` naming the missing file.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libwtet_amd.so")

INVALID = 0xFFFFFFFF
PRESIZE_FEATURES = 522
PRESIZE_OUTPUTS = 3
PRESIZE_PARAMS = 84896
__version__ = "0.1.0"

_STATUS = {
    0: "ok",
    1: "invalid argument",
    2: "too few points",
    3: "device error",
    4: "out of memory",
    5: "not converged",
    6: "internal error",
}


class _Config(ctypes.Structure):
    _fields_ = [
        ("insertion_rule", ctypes.c_int32),
        ("enable_flipping", ctypes.c_int32),
        ("enable_sorting", ctypes.c_int32),
        ("enable_hilbert_sorting", ctypes.c_int32),
        ("enable_splaying", ctypes.c_int32),
        ("max_insert_iterations", ctypes.c_uint32),
        ("max_flip_iterations", ctypes.c_uint32),
        ("predicate_mode", ctypes.c_int32),
        ("collect_stats", ctypes.c_int32),
        ("presize_mode", ctypes.c_int32),
    ]


class _ResultOwner:
    """Holds a wtet_result until no array over its memory is left."""

    def __init__(self, res):
        self.res = res

    def __del__(self):
        try:
            lib().wtet_result_free(ctypes.byref(self.res))
        except Exception:  # interpreter shutdown: the process frees it
            pass


class _BatchOwner:
    """Holds a wtet_batch_result until no array over its memory is left."""

    def __init__(self, res):
        self.res = res

    def __del__(self):
        try:
            lib().wtet_batch_result_free(ctypes.byref(self.res))
        except Exception:  # interpreter shutdown: the process frees it
            pass


def _owned_rows(ptr, m: int, owner) -> np.ndarray:
    """(m, 4) uint32 view of library memory that keeps `owner` alive."""
    buf = (ctypes.c_uint32 * (4 * m)).from_address(ctypes.addressof(ptr.contents))
    buf._owner = owner
    return np.frombuffer(buf, dtype=np.uint32).reshape(m, 4)


class _Result(ctypes.Structure):
    _fields_ = [
        ("num_tets", ctypes.c_uint32),
        ("tets", ctypes.POINTER(ctypes.c_uint32)),
        ("adjacency", ctypes.POINTER(ctypes.c_uint32)),
        ("num_failed", ctypes.c_uint32),
        ("failed_verts", ctypes.POINTER(ctypes.c_uint32)),
    ]


class _DevResult(ctypes.Structure):
    _fields_ = [
        ("num_tets", ctypes.c_uint32),
        ("d_tets", ctypes.c_void_p),
        ("d_adjacency", ctypes.c_void_p),
        ("num_failed", ctypes.c_uint32),
    ]


class _BatchResult(ctypes.Structure):
    _fields_ = [
        ("num_clouds", ctypes.c_uint32),
        ("num_tets", ctypes.c_uint64),
        ("tets", ctypes.POINTER(ctypes.c_uint32)),
        ("adjacency", ctypes.POINTER(ctypes.c_uint32)),
        ("tet_offsets", ctypes.POINTER(ctypes.c_uint64)),
        ("num_failed", ctypes.c_uint64),
        ("failed_verts", ctypes.POINTER(ctypes.c_uint32)),
        ("failed_offsets", ctypes.POINTER(ctypes.c_uint64)),
    ]


class _BatchDevResult(ctypes.Structure):
    _fields_ = [
        ("num_clouds", ctypes.c_uint32),
        ("rows", ctypes.c_uint64),
        ("d_tets", ctypes.c_void_p),
        ("d_adjacency", ctypes.c_void_p),
        ("tet_start", ctypes.POINTER(ctypes.c_uint64)),
        ("tet_count", ctypes.POINTER(ctypes.c_uint32)),
        ("num_failed", ctypes.c_uint64),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [
        ("pred", ctypes.c_uint64 * 8),
        ("sum_uninserted", ctypes.c_uint64),
        ("sum_active", ctypes.c_uint64),
        ("sum_flips", ctypes.c_uint64),
        ("num_points", ctypes.c_uint64),
        ("num_tets_out", ctypes.c_uint64),
        ("tet_slots", ctypes.c_uint64),
        ("rounds", ctypes.c_uint32),
        ("flip_iterations", ctypes.c_uint32),
        ("violations_gpu", ctypes.c_uint32),
        ("repair_regions", ctypes.c_uint32),
        ("walk_failures", ctypes.c_uint32),
        ("grow_events", ctypes.c_uint32),
        ("duplicates", ctypes.c_uint32),
        ("ms_device", ctypes.c_double),
        ("ms_repair", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("peak_active", ctypes.c_uint64),
        ("presize_slots", ctypes.c_uint64),
        ("tet_capacity", ctypes.c_uint64),
        ("presize_pred", ctypes.c_float * 3),
        ("repair_rows_staged", ctypes.c_uint64),
        ("repair_full_downloads", ctypes.c_uint32),
        ("ms_repair_staging", ctypes.c_double),
        ("ms_repair_compute", ctypes.c_double),
        ("repair_fetches", ctypes.c_uint32),
        ("repair_retries", ctypes.c_uint32),
        ("repair_max_region", ctypes.c_uint32),
        ("points_left", ctypes.c_uint32),
        ("flips_23", ctypes.c_uint64),
        ("flips_32", ctypes.c_uint64),
        ("small_checked", ctypes.c_uint64),
        ("small_flips", ctypes.c_uint64),
    ]


class _StageInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ("round", "uninserted", "winners", "tet_slots", "active",
                                              "checked", "flips", "violations", "deferred")]


_lib = None
_lib_lock = threading.Lock()


def lib():
    """The loaded HIP library (raises RuntimeError when it is not built)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"wtet_amd HIP library not found at {LIB_PATH}; build it with "
                    "`make -C wtet_amd/csrc` (hipcc, gfx950)")
            L = ctypes.CDLL(LIB_PATH)
            P = ctypes.c_void_p
            u32 = ctypes.c_uint32
            L.wtet_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
            L.wtet_destroy.argtypes = [P]
            L.wtet_last_error.argtypes = [P]
            L.wtet_last_error.restype = ctypes.c_char_p
            L.wtet_config_default.argtypes = [ctypes.POINTER(_Config)]
            L.wtet_delaunay.argtypes = [P, P, u32, ctypes.POINTER(_Config), ctypes.POINTER(_Result)]
            L.wtet_delaunay_device.argtypes = [P, P, u32, ctypes.POINTER(_Config),
                                               ctypes.POINTER(_Result)]
            L.wtet_result_free.argtypes = [ctypes.POINTER(_Result)]
            L.wtet_delaunay_resident.argtypes = [P, P, u32, ctypes.POINTER(_Config),
                                                 ctypes.POINTER(_DevResult)]
            L.wtet_delaunay_batch.argtypes = [P, P, P, u32, ctypes.POINTER(_Config),
                                              ctypes.POINTER(_BatchResult)]
            L.wtet_delaunay_batch_device.argtypes = [P, P, P, u32, ctypes.POINTER(_Config),
                                                     ctypes.POINTER(_BatchResult)]
            L.wtet_delaunay_batch_resident.argtypes = [P, P, P, u32, ctypes.POINTER(_Config),
                                                       ctypes.POINTER(_BatchDevResult)]
            L.wtet_batch_result_free.argtypes = [ctypes.POINTER(_BatchResult)]
            L.wtet_get_stats.argtypes = [P, ctypes.POINTER(_Stats)]
            L.wtet_presize_features.argtypes = [P, P, P, u32, P]
            L.wtet_presize_features_noalloc.argtypes = [P, P, P, u32, P]
            L.wtet_device_sort_pairs.argtypes = [P, P, P, P, P, u32, u32, ctypes.c_int]
            L.wtet_set_audit_records.argtypes = [P, u32]
            L.wtet_presize_predict.argtypes = [P, P, P, u32, P]
            L.wtet_set_presize_model.argtypes = [P, P]
            L.wtet_mlp_loss_grad.argtypes = [P, P, P, P, u32, P, P]
            L.wtet_mlp_train.argtypes = [P, P, P, P, u32, u32, u32, ctypes.c_float,
                                         ctypes.c_uint64, P]
            L.wtet_check_quality.argtypes = [P, u32, P, P, u32, P, P]
            L.wtet_encode_adjacency.argtypes = [u32, u32]
            L.wtet_encode_adjacency.restype = u32
            L.wtet_decode_adjacency.argtypes = [u32, P, P]
            L.wtet_version.restype = ctypes.c_char_p
            L.wtet_device_info.argtypes = [P, ctypes.c_char_p, u32]
            L.wtet_set_profiling.argtypes = [P, ctypes.c_int]
            L.wtet_set_insertion_ring.argtypes = [P, ctypes.c_int]
            L.wtet_set_insertion_warmup.argtypes = [P, u32]
            L.wtet_set_repair_rings.argtypes = [P, ctypes.c_int]
            L.wtet_set_repair_region_cap.argtypes = [P, u32]
            L.wtet_set_flip_graphs.argtypes = [P, ctypes.c_int]
            L.wtet_set_flip_items.argtypes = [P, u32]
            L.wtet_get_flip_items.argtypes = [P, P, P]
            L.wtet_autotune.argtypes = [P, P, P, u32, ctypes.POINTER(_Config), P, P]
            L.wtet_device_upload.argtypes = [P, P, ctypes.c_size_t, ctypes.POINTER(P)]
            L.wtet_device_download.argtypes = [P, P, P, ctypes.c_size_t]
            L.wtet_device_free.argtypes = [P, P]
            L.wtet_device_copy.argtypes = [P, P, P, ctypes.c_size_t]
            L.wtet_kernel_times.argtypes = [P, u32, ctypes.POINTER(ctypes.c_char_p),
                                            ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_double)]
            L.wtet_initial_stride.argtypes = [u32]
            L.wtet_initial_stride.restype = u32
            L.wtet_stratified_ranks.argtypes = [u32, u32, P]
            L.wtet_hilbert_index.argtypes = [P, u32, P]
            L.wtet_predicates_host.argtypes = [ctypes.c_int, P, P, u32, ctypes.c_int, P]
            L.wtet_predicates_device.argtypes = [P, ctypes.c_int, P, P, u32, ctypes.c_int, P, P]
            L.wtet_repair.argtypes = [P, u32, ctypes.POINTER(ctypes.POINTER(u32)),
                                      ctypes.POINTER(ctypes.POINTER(u32)), ctypes.POINTER(u32),
                                      ctypes.c_int, P]
            SI = ctypes.POINTER(_StageInfo)
            L.wtet_stage_init.argtypes = [P, P, u32, ctypes.POINTER(_Config), SI]
            for name in ("vote", "split", "split_points", "relocate", "gather"):
                getattr(L, "wtet_stage_" + name).argtypes = [P, SI]
            L.wtet_stage_flip.argtypes = [P, ctypes.c_int, SI]
            L.wtet_stage_readback.argtypes = [P, ctypes.POINTER(_Result)]
            L.wtet_stage_locations.argtypes = [P, u32, P, P, ctypes.POINTER(u32)]
            L.wtet_stage_finish.argtypes = [P, ctypes.POINTER(_Result)]
            L.wtet_stage_end.argtypes = [P]
            _lib = L
        return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One HIP device context (stream + persistent device buffers)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        st = L.wtet_create(device, ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"wtet_create(device={device}) failed: {_STATUS.get(st, st)}")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().wtet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self) -> str:
        return lib().wtet_last_error(self._h).decode()

    def stats(self) -> dict:
        s = _Stats()
        lib().wtet_get_stats(self._h, ctypes.byref(s))
        d = {k: getattr(s, k) for k, _ in _Stats._fields_ if k not in ("pred", "presize_pred")}
        d["pred"] = list(s.pred)
        d["presize_pred"] = list(s.presize_pred)
        return d

    # -- learned presizing ---------------------------------------------------
    def presize_features(self, ptr: int, offsets) -> np.ndarray:
        """(C, 522) features of the clouds of device points at ``ptr``."""
        off = _check_offsets(offsets)
        C = len(off) - 1
        out = np.empty((C, PRESIZE_FEATURES), np.float32)
        st = lib().wtet_presize_features(self._h, ctypes.c_void_p(ptr), _p(off), C, _p(out))
        if st != 0:
            raise RuntimeError(f"wtet_presize_features failed: {self.error()}")
        return out

    def presize_features_noalloc(self, ptr: int, offsets) -> np.ndarray:
        """presize_features without any device allocation: the library's
        static per-GPU scratch array (at most NOALLOC_MAX_CLOUDS clouds per
        call; calls that overlap on one GPU race)."""
        off = _check_offsets(offsets)
        C = len(off) - 1
        out = np.empty((C, PRESIZE_FEATURES), np.float32)
        st = lib().wtet_presize_features_noalloc(self._h, ctypes.c_void_p(ptr), _p(off), C, _p(out))
        if st == 1:
            raise ValueError(f"wtet_presize_features_noalloc: {self.error()}")
        if st != 0:
            raise RuntimeError(f"wtet_presize_features_noalloc failed: {self.error()}")
        return out

    def presize_predict(self, params: np.ndarray, feat: np.ndarray) -> np.ndarray:
        p = _params(params)
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, PRESIZE_FEATURES)
        out = np.empty((len(f), PRESIZE_OUTPUTS), np.float32)
        if lib().wtet_presize_predict(self._h, _p(p), _p(f), len(f), _p(out)) != 0:
            raise RuntimeError(f"wtet_presize_predict failed: {self.error()}")
        return out

    def set_presize_model(self, params: "np.ndarray | None"):
        if params is None:
            lib().wtet_set_presize_model(self._h, None)
            return
        p = _params(params)
        if lib().wtet_set_presize_model(self._h, _p(p)) != 0:
            raise ValueError(f"presize model rejected: {self.error()}")
        self._presize_params = p

    def mlp_loss_grad(self, params, feat, targets):
        p = _params(params)
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, PRESIZE_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, PRESIZE_OUTPUTS)
        loss = np.zeros(1, np.float32)
        grad = np.zeros(PRESIZE_PARAMS, np.float32)
        if lib().wtet_mlp_loss_grad(self._h, _p(p), _p(f), _p(t), len(f), _p(loss), _p(grad)) != 0:
            raise RuntimeError(f"wtet_mlp_loss_grad failed: {self.error()}")
        return float(loss[0]), grad

    def mlp_train(self, params, feat, targets, epochs: int, batch: int, lr: float, seed: int = 0):
        p = _params(params).copy()
        f = np.ascontiguousarray(feat, np.float32).reshape(-1, PRESIZE_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, PRESIZE_OUTPUTS)
        curve = np.zeros(max(epochs, 1), np.float32)
        st = lib().wtet_mlp_train(self._h, _p(p), _p(f), _p(t), len(f), epochs, batch,
                                  ctypes.c_float(lr), ctypes.c_uint64(seed), _p(curve))
        if st != 0:
            raise RuntimeError(f"wtet_mlp_train failed: {self.error()}")
        return p, curve[:epochs]

    def upload(self, arr: np.ndarray) -> int:
        """Copy a host array to a new device allocation; returns its address."""
        a = np.ascontiguousarray(arr)
        d = ctypes.c_void_p()
        if lib().wtet_device_upload(self._h, _p(a), a.nbytes, ctypes.byref(d)) != 0:
            raise RuntimeError(f"device upload failed: {self.error()}")
        return d.value

    def download(self, ptr: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        if lib().wtet_device_download(self._h, ctypes.c_void_p(ptr), _p(out), out.nbytes) != 0:
            raise RuntimeError(f"device download failed: {self.error()}")
        return out

    def copy(self, dst: int, src: int, nbytes: int):
        """Device-to-device copy by the library's streaming kernel (synchronous)."""
        if lib().wtet_device_copy(self._h, ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes) != 0:
            raise RuntimeError(f"device copy failed: {self.error()}")

    def sort_pairs(self, d_keys: int, d_vals: int, d_keys_out: int, d_vals_out: int, n: int,
                   key_bits: int, end_bit: int):
        """The library's stable radix sort on device arrays (synchronous)."""
        st = lib().wtet_device_sort_pairs(self._h, ctypes.c_void_p(d_keys), ctypes.c_void_p(d_vals),
                                          ctypes.c_void_p(d_keys_out), ctypes.c_void_p(d_vals_out),
                                          int(n), int(key_bits), int(end_bit))
        if st == 1:
            raise ValueError(self.error())
        if st != 0:
            raise RuntimeError(f"device sort failed: {self.error()}")

    def free(self, ptr: int):
        lib().wtet_device_free(self._h, ctypes.c_void_p(ptr))

    def set_audit_records(self, records: int):
        """Violation records the first audit of a call keeps (default 2^20);
        more violations run the audit again into a larger buffer."""
        if lib().wtet_set_audit_records(self._h, int(records)) != 0:
            raise ValueError("records must be at least 64")

    def set_insertion_ring(self, ring: int):
        """Insertion independence radius in face-steps (0, 1 or 2)."""
        if lib().wtet_set_insertion_ring(self._h, int(ring)) != 0:
            raise ValueError("ring must be 0, 1 or 2")

    def set_insertion_warmup(self, rounds: "int | None"):
        """The first `rounds` rounds insert one point per tet (ring 0); None:
        chosen from the largest cloud (the default)."""
        v = 0xFFFFFFFF if rounds is None else int(rounds)
        if lib().wtet_set_insertion_warmup(self._h, v) != 0:
            raise ValueError(self.error())

    def set_repair_rings(self, rings: int):
        """Rows within `rings` face-steps of each stuck face go to the host repair
        (three more when the stuck faces number over one per 2,000 points)."""
        if lib().wtet_set_repair_rings(self._h, int(rings)) != 0:
            raise ValueError("rings must be in 1..16")

    def set_repair_region_cap(self, tets: int):
        """A repair region past `tets` tets reruns the device phase with the
        insertion ring one wider (0: no limit; default 16,384)."""
        if lib().wtet_set_repair_region_cap(self._h, int(tets)) != 0:
            raise ValueError(self.error())

    def set_flip_graphs(self, on: bool):
        """Replay the captured graph of a round's flip iterations (default) or
        launch them directly."""
        lib().wtet_set_flip_graphs(self._h, int(bool(on)))

    def set_flip_items(self, items: int):
        """List entries per thread that the persistent flip grids are sized for."""
        if lib().wtet_set_flip_items(self._h, int(items)) != 0:
            raise ValueError(f"invalid items per thread {items}")

    def flip_items(self):
        """(entries per thread, graph captures so far)"""
        b = np.zeros(2, np.uint32)
        lib().wtet_get_flip_items(self._h, _p(b[:1]), _p(b[1:]))
        return int(b[0]), int(b[1])

    def autotune(self, ptr: int, offsets, config: "Config | None" = None):
        """Time 1, 2, 4 and 8 list entries per thread on these device points
        (median of three calls each, after two untimed ones, the second of
        which captures the graphs);
        keeps the fastest.  Returns (entries per thread, median device ms)."""
        off = _check_offsets(offsets)
        self._prepare(config)
        c = config._c() if config is not None else Config()._c()
        g = np.zeros(1, np.uint32)
        ms = np.zeros(1, np.float64)
        st = lib().wtet_autotune(self._h, ctypes.c_void_p(ptr), _p(off), len(off) - 1, ctypes.byref(c),
                                 _p(g), _p(ms))
        if st != 0:
            raise RuntimeError(f"wtet_autotune failed: {self.error()}")
        return int(g[0]), float(ms[0])

    def set_profiling(self, on: bool = True):
        lib().wtet_set_profiling(self._h, int(on))

    def kernel_times(self) -> dict:
        """{kernel group: (launches, total ms)} accumulated since set_profiling."""
        out = {}
        k = 0
        while True:
            nm = ctypes.c_char_p()
            cnt = ctypes.c_uint32()
            ms = ctypes.c_double()
            if lib().wtet_kernel_times(self._h, k, ctypes.byref(nm), ctypes.byref(cnt), ctypes.byref(ms)) != 0:
                return out
            out[nm.value.decode()] = (cnt.value, ms.value)
            k += 1

    def info(self) -> str:
        buf = ctypes.create_string_buffer(512)
        lib().wtet_device_info(self._h, buf, 512)
        return buf.value.decode()

    def _prepare(self, cfg):
        # presize="learned" uses the shipped model unless one was set
        if cfg is not None and cfg.presize == "learned" and getattr(self, "_presize_params", None) is None:
            from . import presize
            self.set_presize_model(presize.load_model())

    def _run(self, fn, ptr, n, cfg):
        self._prepare(cfg)
        res = _Result()
        c = cfg._c() if cfg is not None else Config()._c()
        st = fn(self._h, ptr, n, ctypes.byref(c), ctypes.byref(res))
        if st != 0:
            msg = self.error()
            if st in (1, 2):
                raise ValueError(msg)
            raise RuntimeError(f"wtet_delaunay failed ({_STATUS.get(st, st)}): {msg}")
        return self._take(res)

    def _take(self, res) -> "DelaunayResult":
        """Numpy arrays over a wtet_result: tets and adjacency use the
        library's arrays in place (no copy of the rows); the result is freed
        when the last array over them is gone."""
        m = res.num_tets
        if not m:
            lib().wtet_result_free(ctypes.byref(res))
            z = np.zeros((0, 4), np.uint32)
            return DelaunayResult(z, z.copy(), np.zeros((0,), np.uint32), self.stats())
        try:
            k = res.num_failed
            failed = np.ctypeslib.as_array(res.failed_verts, shape=(k,)).copy() if k else \
                np.zeros((0,), np.uint32)
        except BaseException:
            lib().wtet_result_free(ctypes.byref(res))
            raise
        owner = _ResultOwner(res)
        tets = _owned_rows(res.tets, m, owner)
        adj = _owned_rows(res.adjacency, m, owner)
        return DelaunayResult(tets, adj, failed, self.stats())

    # -- stage-level calls (one step of the device loop at a time) ---------------

    def _stage(self, name, fn, *args) -> dict:
        info = _StageInfo()
        st = fn(self._h, *args, ctypes.byref(info))
        if st != 0:
            msg = self.error()
            if st in (1, 2):
                raise ValueError(msg)
            raise RuntimeError(f"wtet_stage_{name} failed ({_STATUS.get(st, st)}): {msg}")
        return {k: getattr(info, k) for k, _ in _StageInfo._fields_}

    def stage_init(self, points, config: "Config | None" = None) -> dict:
        """Starts a stage session: points uploaded, normalized and ordered, five
        start tets, every point uninserted (GpuState::new + dispatch_init)."""
        pts = _validate(points)
        self._prepare(config)
        c = config._c() if config is not None else Config()._c()
        return self._stage("init", lib().wtet_stage_init, _p(pts), len(pts), ctypes.byref(c))

    def stage_vote(self) -> dict:
        return self._stage("vote", lib().wtet_stage_vote)

    def stage_split(self) -> dict:
        return self._stage("split", lib().wtet_stage_split)

    def stage_split_points(self) -> dict:
        return self._stage("split_points", lib().wtet_stage_split_points)

    def stage_flip(self, predicate_mode: int = -1) -> dict:
        """One flip iteration; predicate_mode -1 the session's, 0 staged, 1 exact."""
        return self._stage("flip", lib().wtet_stage_flip, int(predicate_mode))

    def stage_relocate(self) -> dict:
        return self._stage("relocate", lib().wtet_stage_relocate)

    def stage_gather(self) -> dict:
        return self._stage("gather", lib().wtet_stage_gather)

    def stage_readback(self) -> "DelaunayResult":
        """The current mesh (live tets in slot order, no repair)."""
        res = _Result()
        st = lib().wtet_stage_readback(self._h, ctypes.byref(res))
        if st != 0:
            raise (ValueError if st in (1, 2) else RuntimeError)(self.error())
        return self._take(res)

    def stage_locations(self):
        """(point_ids, rows): uninserted input ids and the readback rows of
        the tets holding them."""
        cnt = ctypes.c_uint32(0)
        st = lib().wtet_stage_locations(self._h, 0, None, None, ctypes.byref(cnt))
        if st != 0:
            raise (ValueError if st in (1, 2) else RuntimeError)(self.error())
        ids = np.zeros(cnt.value, np.uint32)
        rows = np.zeros(cnt.value, np.uint32)
        if cnt.value:
            st = lib().wtet_stage_locations(self._h, cnt.value, _p(ids), _p(rows), ctypes.byref(cnt))
            if st != 0:
                raise (ValueError if st in (1, 2) else RuntimeError)(self.error())
        return ids, rows

    def stage_finish(self) -> "DelaunayResult":
        """Remaining flips, gather, readback and host repair; ends the session."""
        res = _Result()
        st = lib().wtet_stage_finish(self._h, ctypes.byref(res))
        if st != 0:
            raise (ValueError if st in (1, 2) else RuntimeError)(self.error())
        return self._take(res)

    def stage_end(self):
        lib().wtet_stage_end(self._h)

    def delaunay(self, points: np.ndarray, config: "Config | None" = None) -> "DelaunayResult":
        pts = _validate(points)
        return self._run(lib().wtet_delaunay, _p(pts), len(pts), config)

    def delaunay_device(self, ptr: int, n: int, config: "Config | None" = None) -> "DelaunayResult":
        """Points already in device memory (n*3 f32 at ``ptr``), e.g. a torch
        tensor's ``data_ptr()``; the caller synchronizes its own stream first."""
        if n < 4:
            raise ValueError(f"Need at least 4 points for 3D Delaunay, got {n}")
        return self._run(lib().wtet_delaunay_device, ctypes.c_void_p(ptr), n, config)

    def delaunay_resident(self, ptr: int, n: int, config: "Config | None" = None):
        """Points and result in device memory.  Returns (num_tets, d_tets,
        d_adjacency) device addresses owned by the context (valid until the
        next call)."""
        if n < 4:
            raise ValueError(f"Need at least 4 points for 3D Delaunay, got {n}")
        self._prepare(config)
        res = _DevResult()
        c = config._c() if config is not None else Config()._c()
        st = lib().wtet_delaunay_resident(self._h, ctypes.c_void_p(ptr), n, ctypes.byref(c),
                                          ctypes.byref(res))
        if st != 0:
            raise RuntimeError(f"wtet_delaunay_resident failed ({_STATUS.get(st, st)}): {self.error()}")
        return res.num_tets, res.d_tets, res.d_adjacency

    def _run_batch(self, fn, ptr, offsets, cfg):
        self._prepare(cfg)
        res = _BatchResult()
        c = cfg._c() if cfg is not None else Config()._c()
        off = np.ascontiguousarray(offsets, np.uint64)
        st = fn(self._h, ptr, _p(off), len(off) - 1, ctypes.byref(c), ctypes.byref(res))
        if st != 0:
            msg = self.error()
            if st in (1, 2):
                raise ValueError(msg)
            raise RuntimeError(f"wtet_delaunay_batch failed ({_STATUS.get(st, st)}): {msg}")
        # tets and adjacency: the library's arrays in place (no copy of the
        # rows), freed when the last array over them is gone
        m, C, k = res.num_tets, res.num_clouds, res.num_failed
        try:
            arr = lambda p, n, dt=np.uint32: (np.ctypeslib.as_array(p, shape=(n,)).copy() if n  # noqa: E731
                                              else np.zeros((0,), dt))
            toff = arr(res.tet_offsets, C + 1, np.uint64)
            failed = arr(res.failed_verts, k)
            foff = arr(res.failed_offsets, C + 1, np.uint64)
        except BaseException:
            lib().wtet_batch_result_free(ctypes.byref(res))
            raise
        if m:
            owner = _BatchOwner(res)
            tets = _owned_rows(res.tets, m, owner)
            adj = _owned_rows(res.adjacency, m, owner)
        else:
            lib().wtet_batch_result_free(ctypes.byref(res))
            tets = np.zeros((0, 4), np.uint32)
            adj = np.zeros((0, 4), np.uint32)
        return BatchResult(tets, adj, toff, failed, foff, self.stats())

    def delaunay_batch(self, clouds, offsets=None, config: "Config | None" = None) -> "BatchResult":
        """Many independent clouds in one pass: ``clouds`` is a list of (n_c, 3)
        arrays, or one (N, 3) array with ``offsets`` (C+1, offsets[0] = 0).
        Cloud c's result equals ``delaunay(cloud c)``."""
        pts, off = _batch_input(clouds, offsets)
        return self._run_batch(lib().wtet_delaunay_batch, _p(pts), off, config)

    def delaunay_batch_device(self, ptr: int, offsets, config: "Config | None" = None):
        """Points already in device memory (offsets[-1]*3 f32 at ``ptr``)."""
        off = _check_offsets(offsets)
        return self._run_batch(lib().wtet_delaunay_batch_device, ctypes.c_void_p(ptr), off, config)

    def delaunay_batch_resident(self, ptr: int, offsets, config: "Config | None" = None):
        """Points and result in device memory.  Returns (d_tets, d_adjacency,
        tet_start, tet_count): device addresses owned by the context and the
        per-cloud rows (valid until the next call)."""
        off = _check_offsets(offsets)
        self._prepare(config)
        res = _BatchDevResult()
        c = config._c() if config is not None else Config()._c()
        st = lib().wtet_delaunay_batch_resident(self._h, ctypes.c_void_p(ptr), _p(off), len(off) - 1,
                                                ctypes.byref(c), ctypes.byref(res))
        if st != 0:
            msg = self.error()
            if st in (1, 2):
                raise ValueError(msg)
            raise RuntimeError(f"wtet_delaunay_batch_resident failed ({_STATUS.get(st, st)}): {msg}")
        C = res.num_clouds
        start = np.ctypeslib.as_array(res.tet_start, shape=(C,)).copy()
        count = np.ctypeslib.as_array(res.tet_count, shape=(C,)).copy()
        return res.d_tets, res.d_adjacency, start, count

    def predicates(self, kind: int, pts: np.ndarray, ids: np.ndarray | None, mode: int = 0):
        """Device predicate batch; returns (signs, stage counters)."""
        pts = np.ascontiguousarray(pts, np.float32)
        per = 4 if kind in (0, 2) else 5
        n = pts.size // (3 * per)
        out = np.empty(n, np.int32)
        st8 = np.zeros(8, np.uint64)
        idp = None
        if ids is not None:
            ids = np.ascontiguousarray(ids, np.uint32)
            idp = _p(ids)
        s = lib().wtet_predicates_device(self._h, kind, _p(pts), idp, n, mode, _p(out), _p(st8))
        if s != 0:
            raise RuntimeError(f"device predicates failed: {self.error()}")
        return out, st8


def initial_stride(n: int) -> int:
    """The reference's first stratified stride for n points."""
    return int(lib().wtet_initial_stride(int(n)))


def stratified_ranks(n: int, stride: int = 0) -> np.ndarray:
    """Position of each of n sorted points in the stratified insertion order."""
    out = np.zeros(max(n, 1), np.uint32)
    if lib().wtet_stratified_ranks(int(n), int(stride), _p(out)) != 0:
        raise ValueError(f"invalid n {n} or stride {stride}")
    return out[:n]


def hilbert_index(cells) -> np.ndarray:
    """3D Hilbert index of (k, 3) integer cells in 0..255."""
    c = np.ascontiguousarray(cells, np.uint32).reshape(-1, 3)
    out = np.zeros(max(len(c), 1), np.uint32)
    lib().wtet_hilbert_index(_p(c), len(c), _p(out))
    return out[:len(c)]


def predicates_host(kind: int, pts: np.ndarray, ids: np.ndarray | None, mode: int = 0):
    pts = np.ascontiguousarray(pts, np.float32)
    per = 4 if kind in (0, 2) else 5
    n = pts.size // (3 * per)
    out = np.empty(n, np.int32)
    idp = None
    if ids is not None:
        ids = np.ascontiguousarray(ids, np.uint32)
        idp = _p(ids)
    if lib().wtet_predicates_host(kind, _p(pts), idp, n, mode, _p(out)) != 0:
        raise ValueError("invalid predicate batch")
    return out


def repair(points_normalized: np.ndarray, tets: np.ndarray, adjacency: np.ndarray | None = None,
           mode: int = 0):
    """Host phase 2 on a valid triangulation (ids: points, super n..n+3,
    infinity n+4): returns (tets, adjacency) of the exact Delaunay mesh."""
    L = lib()
    pts = np.ascontiguousarray(points_normalized, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
    tp = (ctypes.c_uint32 * t.size).from_buffer_copy(t.tobytes())
    ap = None
    if adjacency is not None:
        a = np.ascontiguousarray(adjacency, np.uint32).reshape(-1, 4)
        ap = (ctypes.c_uint32 * a.size).from_buffer_copy(a.tobytes())
    tptr = ctypes.cast(tp, ctypes.POINTER(ctypes.c_uint32))
    aptr = ctypes.cast(ap, ctypes.POINTER(ctypes.c_uint32)) if ap is not None else \
        ctypes.POINTER(ctypes.c_uint32)()
    m = ctypes.c_uint32(len(t))
    regions = np.zeros(2, np.uint32)
    st = L.wtet_repair(_p(pts), len(pts), ctypes.byref(tptr), ctypes.byref(aptr), ctypes.byref(m),
                       mode, _p(regions))
    if st != 0:
        raise RuntimeError(f"wtet_repair failed ({_STATUS.get(st, st)})")
    k = m.value
    out_t = np.ctypeslib.as_array(tptr, shape=(k * 4,)).copy().reshape(k, 4)
    out_a = np.ctypeslib.as_array(aptr, shape=(k * 4,)).copy().reshape(k, 4)
    L.wtet_result_free(ctypes.byref(_Result(k, tptr, aptr, 0, ctypes.POINTER(ctypes.c_uint32)())))
    return out_t, out_a, int(regions[0])


def _validate(points) -> np.ndarray:
    a = np.asarray(points)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"Points must have shape (N, 3), got {list(a.shape)}")
    if a.shape[0] < 4:
        raise ValueError(f"Need at least 4 points for 3D Delaunay, got {a.shape[0]}")
    return np.ascontiguousarray(a, dtype=np.float32)


def _params(params) -> np.ndarray:
    p = np.ascontiguousarray(params, np.float32).ravel()
    if p.shape != (PRESIZE_PARAMS,):
        raise ValueError(f"presize parameters must hold {PRESIZE_PARAMS} floats, got {p.size}")
    return p


def _check_offsets(offsets) -> np.ndarray:
    off = np.ascontiguousarray(offsets, np.uint64)
    if off.ndim != 1 or len(off) < 2 or off[0] != 0:
        raise ValueError("offsets must be a 1-D array of C+1 >= 2 entries starting at 0")
    d = np.diff(off.astype(np.int64))
    if (d < 0).any():
        raise ValueError("offsets must be nondecreasing")
    if (d < 4).any():
        c = int(np.argmax(d < 4))
        raise ValueError(f"Need at least 4 points for 3D Delaunay, got {int(d[c])} in cloud {c}")
    return off


def _batch_input(clouds, offsets):
    if offsets is None:
        arrs = [_validate(c) for c in clouds]
        off = np.zeros(len(arrs) + 1, np.uint64)
        off[1:] = np.cumsum([len(a) for a in arrs])
        pts = np.ascontiguousarray(np.concatenate(arrs) if arrs else np.zeros((0, 3), np.float32))
        return pts, _check_offsets(off)
    a = np.asarray(clouds)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"Points must have shape (N, 3), got {list(a.shape)}")
    off = _check_offsets(offsets)
    if int(off[-1]) != a.shape[0]:
        raise ValueError(f"offsets[-1] = {int(off[-1])} but {a.shape[0]} points were given")
    return np.ascontiguousarray(a, np.float32), off


_PRESIZE = {"fixed": 0, "learned": 1, "grow": 2}


class Config:
    """Triangulation options (src/python.rs:53-120; GDelConfig src/types.rs:32-61).

    ``predicate_mode`` ("filtered" | "exact"), ``collect_stats`` and
    ``presize`` ("fixed": the reference's max(8(N+5)+8, 12N, 64) tet slots,
    src/gpu/mod.rs:101-108; "learned": the presizing network's prediction;
    "grow": grow-and-retry from 2 per point) are additions of this library.
    ``max_insert_iterations`` caps the insertion rounds as in
    src/phase1.rs:26; the points they leave are inserted by the host phase
    (when ``enable_splaying``) and listed in ``failed_verts``."""

    def __init__(self, insertion_rule: str = "circumcenter", enable_flipping: bool = True,
                 enable_sorting: bool = False, enable_hilbert_sorting: bool = False,
                 enable_splaying: bool = True, max_insert_iterations: int = 100,
                 max_flip_iterations: int = 10, predicate_mode: str = "filtered",
                 collect_stats: bool = False, presize: str = "fixed"):
        rule = str(insertion_rule).lower()
        if rule not in ("circumcenter", "centroid"):
            raise ValueError(
                f"Invalid insertion_rule '{insertion_rule}'. Must be 'circumcenter' or 'centroid'.")
        if predicate_mode not in ("filtered", "exact"):
            raise ValueError(f"Invalid predicate_mode '{predicate_mode}'. Must be 'filtered' or 'exact'.")
        self.insertion_rule = rule
        self.enable_flipping = bool(enable_flipping)
        self.enable_sorting = bool(enable_sorting)
        self.enable_hilbert_sorting = bool(enable_hilbert_sorting)
        self.enable_splaying = bool(enable_splaying)
        self.max_insert_iterations = int(max_insert_iterations)
        self.max_flip_iterations = int(max_flip_iterations)
        if presize not in _PRESIZE:
            raise ValueError(f"Invalid presize '{presize}'. Must be 'fixed', 'learned' or 'grow'.")
        self.predicate_mode = predicate_mode
        self.collect_stats = bool(collect_stats)
        self.presize = presize

    def _c(self) -> _Config:
        return _Config(0 if self.insertion_rule == "circumcenter" else 1, int(self.enable_flipping),
                       int(self.enable_sorting), int(self.enable_hilbert_sorting),
                       int(self.enable_splaying), self.max_insert_iterations,
                       self.max_flip_iterations, 1 if self.predicate_mode == "exact" else 0,
                       int(self.collect_stats), _PRESIZE[self.presize])

    def __repr__(self) -> str:
        b = lambda v: "true" if v else "false"  # noqa: E731 (Rust bool formatting)
        return (f"Config(insertion_rule='{self.insertion_rule}', flipping={b(self.enable_flipping)}, "
                f"sorting={b(self.enable_sorting)}, hilbert={b(self.enable_hilbert_sorting)}, "
                f"splaying={b(self.enable_splaying)}, max_insert_iter={self.max_insert_iterations}, "
                f"max_flip_iter={self.max_flip_iterations})")


class DelaunayResult:
    """src/python.rs:126-160: ``tets`` (M,4) uint32, ``adjacency`` (M,4)
    uint32 packed ``(tet << 5) | face``, ``failed_verts`` (K,) uint32."""

    def __init__(self, tets, adjacency, failed_verts, stats=None):
        self.tets = tets
        self.adjacency = adjacency
        self.failed_verts = failed_verts
        self.stats = stats or {}

    @property
    def num_tets(self) -> int:
        return int(self.tets.shape[0])

    @property
    def num_failed(self) -> int:
        return int(self.failed_verts.shape[0])

    def __repr__(self) -> str:
        return f"DelaunayResult(num_tets={self.num_tets}, failed_verts={self.num_failed})"


class BatchResult:
    """Result of a batch: all clouds' rows in cloud order.  ``result[c]`` is
    cloud c's DelaunayResult (views; cloud-local vertex ids and adjacency)."""

    def __init__(self, tets, adjacency, tet_offsets, failed_verts, failed_offsets, stats=None):
        self.tets = tets
        self.adjacency = adjacency
        self.tet_offsets = tet_offsets
        self.failed_verts = failed_verts
        self.failed_offsets = failed_offsets
        self.stats = stats or {}

    @property
    def num_clouds(self) -> int:
        return int(len(self.tet_offsets) - 1)

    def __len__(self) -> int:
        return self.num_clouds

    def __getitem__(self, c: int) -> DelaunayResult:
        if c < 0:
            c += self.num_clouds
        if not 0 <= c < self.num_clouds:
            raise IndexError(c)
        a, b = int(self.tet_offsets[c]), int(self.tet_offsets[c + 1])
        fa, fb = int(self.failed_offsets[c]), int(self.failed_offsets[c + 1])
        return DelaunayResult(self.tets[a:b], self.adjacency[a:b], self.failed_verts[fa:fb])

    def __repr__(self) -> str:
        return f"BatchResult(num_clouds={self.num_clouds}, num_tets={int(self.tets.shape[0])})"


_default = None
_default_lock = threading.Lock()


def _ctx() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(0)
        return _default


def delaunay(points, config: Config | None = None) -> DelaunayResult:
    """3D Delaunay triangulation of (N, 3) float32 points (src/python.rs:166-250)."""
    pts = _validate(points)
    return _ctx().delaunay(pts, config)


def delaunay_batch(clouds, offsets=None, config: Config | None = None) -> BatchResult:
    """Triangulate many independent clouds in one pass (a list of (n_c, 3)
    arrays, or one (N, 3) array with C+1 offsets)."""
    return _ctx().delaunay_batch(clouds, offsets, config)


def check_quality(points, result: DelaunayResult):
    """(violations, interior_tets) among all-real tet pairs (src/python.rs:282-339)."""
    a = np.asarray(points)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"Points must have shape (N, 3), got {list(a.shape)}")
    pts = np.ascontiguousarray(a, np.float32)
    t = np.ascontiguousarray(result.tets, np.uint32)
    adj = np.ascontiguousarray(result.adjacency, np.uint32)
    v = np.zeros(1, np.uint64)
    i = np.zeros(1, np.uint64)
    st = lib().wtet_check_quality(_p(pts), len(pts), _p(t), _p(adj), len(t), _p(v), _p(i))
    if st != 0:
        raise RuntimeError("check_quality: malformed adjacency")
    return int(v[0]), int(i[0])


def initialize_gpu() -> str:
    """Create the default device context; returns a description (src/python.rs:344-353)."""
    c = _ctx()
    return "GPU initialized: " + c.info().replace("\n", "; ")


def gpu_info() -> str:
    """src/python.rs:358-371."""
    if _default is None:
        return "GPU not yet initialized. Call initialize_gpu() or run delaunay() to initialize."
    return _default.info()


def encode_adjacency(tet_idx: int, face_idx: int) -> int:
    return ((int(tet_idx) << 5) | (int(face_idx) & 3)) & 0xFFFFFFFF


def decode_adjacency(packed: int):
    p = int(packed) & 0xFFFFFFFF
    return (p >> 5, p & 3)
