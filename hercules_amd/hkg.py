"""ctypes binding of libhkg.so (the C ABI declared in include/hkg.h).

This module is plumbing: it loads the in-tree HIP library and raises if it
is missing -- there is no CPU fallback.  Names follow the reference's query
API (code/hercules/include/hercules_query_engine.h): an index directory is
opened (hercules_index_read), then exact k-NN queries are answered.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhkg.so")

HKG_PATH_AUTO, HKG_PATH_INDEX, HKG_PATH_BATCHED, HKG_PATH_SCAN, HKG_PATH_BF16 = range(5)
HKG_OK, HKG_EINVAL, HKG_ENOMEM, HKG_EIO, HKG_EFORMAT, HKG_EDEVICE, HKG_EBUSY, HKG_ENOTREADY, HKG_EINTERNAL = range(9)
PATHS = {"auto": HKG_PATH_AUTO, "index": HKG_PATH_INDEX, "batched": HKG_PATH_BATCHED,
         "scan": HKG_PATH_SCAN, "bf16": HKG_PATH_BF16}
STATUS = {0: "ok", 1: "invalid argument", 2: "out of memory", 3: "i/o error", 4: "malformed index",
          5: "device error", 6: "busy", 7: "not ready", 8: "internal check failed"}
FLT_MAX = np.float32(np.finfo(np.float32).max)
NO_ID = np.uint64(0xFFFFFFFFFFFFFFFF)


class HkgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Options(C.Structure):
    _fields_ = [("device", C.c_int), ("keep_bf16", C.c_int), ("verbose", C.c_int)]


class QueryOpts(C.Structure):
    _fields_ = [("path", C.c_int), ("batch", C.c_uint32), ("profile", C.c_uint32), ("cand_cap", C.c_uint32),
                ("bf16_bound_pass", C.c_uint32), ("epsilon", C.c_float), ("delta", C.c_float),
                ("nprobes", C.c_uint32), ("approx_leaves", C.c_uint32), ("eapca_threshold", C.c_float),
                ("sax_threshold", C.c_float), ("batched_filter", C.c_uint32), ("bf16_query_tiles", C.c_uint32)]


BATCHED_FILTERS = {"auto": 0, "items": 1, "leaf": 2}


def query_opts(path: str = "auto", batch: int = 0, profile: bool = False, cand_cap: int = 0,
               bf16_bound_pass: int = 0, batched_filter: str = "auto", **modes) -> QueryOpts:
    """hkg_default_query_opts, then the given settings.  ``batched_filter``
    (the batched path's filter): "auto", "items" (per (query, leaf) work
    items) or "leaf" (leaf-ordered, one query per lane).  ``modes``:
    epsilon, delta, nprobes, approx_leaves, eapca_threshold, sax_threshold
    (the reference's --epsilon, --delta, --nprobes, --approx-stop-condition,
    --eapca-threshold, --sax-threshold), and bf16_query_tiles (the BF16
    filter's 32-query tiles per wave; 0: the tuned or default shape)."""
    qo = QueryOpts()
    lib().hkg_default_query_opts(C.byref(qo))
    qo.path, qo.batch, qo.profile, qo.cand_cap, qo.bf16_bound_pass = (PATHS[path], batch, int(profile), cand_cap,
                                                                       bf16_bound_pass)
    qo.batched_filter = BATCHED_FILTERS[batched_filter]
    for key, v in modes.items():
        if key not in ("epsilon", "delta", "nprobes", "approx_leaves", "eapca_threshold", "sax_threshold",
                       "bf16_query_tiles"):
            raise TypeError(f"unknown query option {key!r}")
        setattr(qo, key, v)
    return qo


STAGES = ["prep", "eapca", "select_leaf", "approx", "collect", "refine", "topk", "scan", "bound"]


class QueryStats(C.Structure):
    _fields_ = [("queries", C.c_uint64), ("leaves_total", C.c_uint64), ("leaves_after_eapca", C.c_uint64),
                ("series_total", C.c_uint64), ("series_lb_checked", C.c_uint64),
                ("series_refined", C.c_uint64), ("bytes_algorithmic", C.c_uint64), ("device_ms", C.c_double),
                ("overflow_retries", C.c_uint32), ("launches", C.c_uint32), ("bf16_pairs", C.c_uint64),
                ("stage_ms", C.c_double * len(STAGES)),
                ("stage_bytes", C.c_uint64 * len(STAGES)), ("stage_launches", C.c_uint64 * len(STAGES)),
                ("mode_leaves_first", C.c_uint64), ("mode_candidates", C.c_uint64),
                ("mode_candidates_scanned", C.c_uint64), ("mode_sequential", C.c_uint64)]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("stage_")}
        for f in ("stage_ms", "stage_bytes", "stage_launches"):
            d[f] = {STAGES[i]: getattr(self, f)[i] for i in range(len(STAGES))}
        return d


class Tuning(C.Structure):
    _fields_ = [("path", C.c_uint32), ("length", C.c_uint32), ("batch", C.c_uint32),
                ("filter_blocks_per_cu", C.c_uint32), ("candidates", C.c_uint32), ("scan_blocks_per_cu", C.c_uint32),
                ("batched_filter", C.c_uint32), ("bf16_query_tiles", C.c_uint32),
                ("ms_per_query", C.c_double), ("ms_per_query_default", C.c_double)]


class IndexInfo(C.Structure):
    _fields_ = [("n_series", C.c_uint64), ("length", C.c_uint32), ("paa_segments", C.c_uint32),
                ("sax_bit_cardinality", C.c_uint32), ("max_leaf_size", C.c_uint32), ("n_leaves", C.c_uint64),
                ("n_nodes", C.c_uint64), ("device_bytes", C.c_uint64), ("has_tree", C.c_int),
                ("has_ids", C.c_int)]


class BuildOpts(C.Structure):
    _fields_ = [("leaf_size", C.c_uint32), ("init_segments", C.c_uint32), ("paa_segments", C.c_uint32),
                ("device", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("scratch_bytes", C.c_uint64), ("queue_capacity", C.c_uint32),
                ("path", C.c_uint32), ("max_inflight", C.c_uint32), ("max_k", C.c_uint32),
                ("bf16_bound_pass", C.c_uint32), ("reserved", C.c_uint32)]


# every entry point of include/hkg.h with its ctypes signature
_P = C.c_void_p
SIGNATURES = {
    "hkg_default_options": (None, [_P]),
    "hkg_default_query_opts": (None, [_P]),
    "hkg_open_dir": (C.c_int, [C.c_char_p, _P, C.POINTER(_P)]),
    "hkg_open_raw": (C.c_int, [_P, C.c_uint64, C.c_uint32, _P, C.POINTER(_P)]),
    "hkg_set_ids": (C.c_int, [_P, _P, C.c_uint64]),
    "hkg_info": (C.c_int, [_P, _P]),
    "hkg_inspect_dir": (C.c_int, [C.c_char_p, _P]),
    "hkg_close": (None, [_P]),
    "hkg_knn": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P]),
    "hkg_knn_device": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, _P, _P, _P, _P]),
    "hkg_scan_knn_device_noalloc": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, _P, _P]),
    "hkg_autotune": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, _P]),
    "hkg_default_build_opts": (None, [_P]),
    "hkg_build": (C.c_int, [_P, C.c_uint64, C.c_uint32, _P, C.POINTER(_P)]),
    "hkg_build_device": (C.c_int, [_P, C.c_uint64, C.c_uint32, _P, C.POINTER(_P)]),
    "hkg_write_dir": (C.c_int, [_P, C.c_char_p]),
    "hkg_gen_series_device": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, _P, _P]),
    "hkg_gen_noisy_queries_device": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                                               C.c_uint32, _P, _P, _P]),
    "hkg_default_server_opts": (None, [_P]),
    "hkg_server_start": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "hkg_submit": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "hkg_poll": (C.c_int, [_P, C.c_uint64, _P, _P, C.POINTER(C.c_double)]),
    "hkg_wait": (C.c_int, [_P, C.c_uint64, _P, _P, C.POINTER(C.c_double)]),
    "hkg_server_stop": (None, [_P]),
    "hkg_status_str": (C.c_char_p, [C.c_int]),
    "hkg_last_error": (C.c_char_p, []),
    "hkg_breakpoints256": (None, [_P]),
    "hkg_device_alloc": (_P, [C.c_uint64]),
    "hkg_device_free": (None, [_P]),
    "hkg_memcpy": (C.c_int, [_P, _P, C.c_uint64, C.c_int]),
    "hkg_device_sync": (C.c_int, []),
}

_lib = None


def lib() -> C.CDLL:
    """Load libhkg.so (built in-tree by hercules_amd.build); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m hercules_amd.build` "
                              "(there is no CPU fallback for the query path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise HkgError(rc, lib().hkg_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def inspect_dir(index_dir: str) -> dict:
    """Parse and validate an index directory on the host (no device)."""
    inf = IndexInfo()
    _check(lib().hkg_inspect_dir(os.fsencode(index_dir), C.byref(inf)))
    return {f: getattr(inf, f) for f, _ in IndexInfo._fields_}


def breakpoints256() -> np.ndarray:
    out = np.zeros(255, np.float32)
    lib().hkg_breakpoints256(_ptr(out))
    return out


@dataclass
class KnnResult:
    ids: np.ndarray      # (nq, kmax) uint64, NO_ID where unfilled
    d2: np.ndarray       # (nq, kmax) float32 squared distances, FLT_MAX where unfilled
    stats: dict


class DeviceArray:
    """A raw device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().hkg_device_alloc(self.nbytes)
        if not self.ptr:
            raise HkgError(2, f"device allocation of {nbytes} bytes failed")

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().hkg_memcpy(self.ptr, _ptr(a), a.nbytes, 1))

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        _check(lib().hkg_memcpy(_ptr(a), self.ptr, a.nbytes, 2))
        return a

    @classmethod
    def view(cls, ptr: int, nbytes: int) -> "DeviceArray":
        """A view of device memory that this object never frees."""
        v = cls.__new__(cls)
        v.ptr, v.nbytes, v.owned = ptr, int(nbytes), False
        return v

    def close(self) -> None:
        if self.ptr and getattr(self, "owned", True):
            lib().hkg_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HerculesIndex:
    """A Hercules index resident in HBM.

    ``HerculesIndex.open(dir)`` replaces hercules_index_read + cache_sax_file
    (code/hercules/src/hercules_index.c:1793-1905, 2342-2351);
    ``knn`` replaces the exact query functions
    (code/hercules/include/hercules_query_engine.h:203-207, 256-261, 284-289).
    """

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    # ---- construction ----------------------------------------------------
    @classmethod
    def open(cls, index_dir: str, device: int = 0, keep_bf16: bool = False, verbose: bool = False):
        o = Options(device, int(keep_bf16), int(verbose))
        h = C.c_void_p()
        _check(lib().hkg_open_dir(os.fsencode(index_dir), C.byref(o), C.byref(h)))
        return cls(h.value)

    @classmethod
    def open_raw(cls, X: np.ndarray, device: int = 0):
        X = np.ascontiguousarray(X, np.float32)
        o = Options(device, 0, 0)
        h = C.c_void_p()
        _check(lib().hkg_open_raw(_ptr(X), X.shape[0], X.shape[1], C.byref(o), C.byref(h)))
        return cls(h.value)

    @classmethod
    def build(cls, X: np.ndarray, leaf_size: int, device: int = 0):
        X = np.ascontiguousarray(X, np.float32)
        bo = BuildOpts()
        lib().hkg_default_build_opts(C.byref(bo))
        bo.leaf_size = leaf_size
        bo.device = device
        h = C.c_void_p()
        _check(lib().hkg_build(_ptr(X), X.shape[0], X.shape[1], C.byref(bo), C.byref(h)))
        return cls(h.value)

    @classmethod
    def build_device(cls, dX: DeviceArray, n: int, L: int, leaf_size: int, device: int = 0):
        """Build from rows already in HBM; the library takes ownership of dX."""
        bo = BuildOpts()
        lib().hkg_default_build_opts(C.byref(bo))
        bo.leaf_size = leaf_size
        bo.device = device
        h = C.c_void_p()
        ptr = dX.ptr
        dX.ptr = None  # ownership moves to the library
        _check(lib().hkg_build_device(ptr, n, L, C.byref(bo), C.byref(h)))
        return cls(h.value)

    def write(self, index_dir: str) -> None:
        os.makedirs(index_dir, exist_ok=True)
        _check(lib().hkg_write_dir(self._h, os.fsencode(index_dir)))

    def set_ids(self, ids: np.ndarray) -> None:
        ids = np.ascontiguousarray(ids, np.uint64)
        _check(lib().hkg_set_ids(self._h, _ptr(ids), ids.shape[0]))

    def info(self) -> dict:
        inf = IndexInfo()
        _check(lib().hkg_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in IndexInfo._fields_}

    def close(self) -> None:
        if self._h and self._h.value:
            lib().hkg_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- queries -------------------------------------------------------------
    def knn(self, Q: np.ndarray, k, path: str = "auto", batch: int = 0, profile: bool = False,
            cand_cap: int = 0, bf16_bound_pass: int = 0, **modes) -> KnnResult:
        """k-NN of each row of Q.  ``k`` is an int or a per-query array.
        Exact unless ``modes`` asks for the reference's epsilon or nprobes
        search (see query_opts)."""
        Q = np.ascontiguousarray(Q, np.float32)
        if Q.ndim == 1:
            Q = Q[None, :]
        nq = Q.shape[0]
        if np.isscalar(k):
            kmax = int(k)
            kv = None
        else:
            kv = np.ascontiguousarray(k, np.uint32)
            kmax = int(kv.max()) if nq else 1
        ids = np.empty((nq, kmax), np.uint64)
        d2 = np.empty((nq, kmax), np.float32)
        qo = query_opts(path, batch, profile, cand_cap, bf16_bound_pass, **modes)
        st = QueryStats()
        _check(lib().hkg_knn(self._h, _ptr(Q), nq, _ptr(kv) if kv is not None else None, kmax, C.byref(qo),
                             _ptr(ids), _ptr(d2), C.byref(st)))
        return KnnResult(ids, d2, st.as_dict())

    def autotune(self, Q: np.ndarray, k, path: str = "batched") -> dict:
        """Time the launch shapes of ``path`` on the sample queries Q (k an
        int or per-query array) and keep the fastest for later calls that
        leave ``batch`` at 0."""
        Q = np.ascontiguousarray(Q, np.float32)
        if Q.ndim == 1:
            Q = Q[None, :]
        nq = Q.shape[0]
        if np.isscalar(k):
            kmax, kv = int(k), None
        else:
            kv = np.ascontiguousarray(k, np.uint32)
            kmax = int(kv.max()) if nq else 1
        t = Tuning()
        _check(lib().hkg_autotune(self._h, _ptr(Q), nq, _ptr(kv) if kv is not None else None, kmax, PATHS[path],
                                  C.byref(t)))
        names = {v: n for n, v in PATHS.items()}
        return {"path": names[t.path], "length": t.length, "batch": t.batch,
                "filter_blocks_per_cu": t.filter_blocks_per_cu, "scan_blocks_per_cu": t.scan_blocks_per_cu,
                "batched_filter": {v: n for n, v in BATCHED_FILTERS.items()}.get(t.batched_filter, "auto"),
                "candidates": t.candidates, "bf16_query_tiles": t.bf16_query_tiles,
                "ms_per_query": t.ms_per_query, "ms_per_query_default": t.ms_per_query_default}

    def knn_device(self, dQ: DeviceArray, nq: int, dk: DeviceArray, kmax: int, dids: DeviceArray,
                   dd2: DeviceArray, path: str = "auto", batch: int = 0, profile: bool = False,
                   bf16_bound_pass: int = 0, **modes) -> dict:
        qo = query_opts(path, batch, profile, 0, bf16_bound_pass, **modes)
        st = QueryStats()
        _check(lib().hkg_knn_device(self._h, dQ.ptr, nq, dk.ptr, kmax, C.byref(qo), dids.ptr, dd2.ptr, None,
                                    C.byref(st)))
        return st.as_dict()

    def scan_knn_device_noalloc(self, dQ: DeviceArray, nq: int, dk: DeviceArray, kmax: int, dids: DeviceArray,
                                dd2: DeviceArray) -> None:
        """The full scan through hkg_scan_knn_device_noalloc: no device
        allocation (scratch in the library's own device array)."""
        _check(lib().hkg_scan_knn_device_noalloc(self._h, dQ.ptr, nq, dk.ptr, kmax, dids.ptr, dd2.ptr, None))


def gen_series_device(kind: int, seed: int, first: int, n: int, L: int, out: DeviceArray) -> None:
    _check(lib().hkg_gen_series_device(kind, seed, first, n, L, out.ptr, None))
    _check(lib().hkg_device_sync())


def gen_series(kind: int, seed: int, first: int, n: int, L: int) -> np.ndarray:
    d = DeviceArray(n * L * 4)
    try:
        gen_series_device(kind, seed, first, n, L, d)
        return d.download(np.empty((n, L), np.float32))
    finally:
        d.close()


def gen_noisy_queries(kind: int, data_seed: int, N: int, query_seed: int, sigmas, L: int):
    sig = np.ascontiguousarray(sigmas, np.float32)
    nq = sig.shape[0]
    d = DeviceArray(max(nq, 1) * L * 4)
    src = np.empty(nq, np.uint64)
    try:
        _check(lib().hkg_gen_noisy_queries_device(kind, data_seed, N, query_seed, nq, L, _ptr(sig), d.ptr,
                                                  _ptr(src)))
        return d.download(np.empty((nq, L), np.float32)), src
    finally:
        d.close()


class Server:
    """Continuous-batching request loop over a resident index."""

    def __init__(self, index: HerculesIndex, max_batch: int = 0, scratch_bytes: int = 0, queue_capacity: int = 0,
                 path: str = "batched", max_inflight: int = 0, max_k: int = 0, bf16_bound_pass: int = 0):
        so = ServerOpts()
        lib().hkg_default_server_opts(C.byref(so))
        if max_batch:
            so.max_batch = max_batch
        if scratch_bytes:
            so.scratch_bytes = scratch_bytes
        if queue_capacity:
            so.queue_capacity = queue_capacity
        so.path = PATHS[path]
        if max_inflight:
            so.max_inflight = max_inflight
        if max_k:
            so.max_k = max_k
        so.bf16_bound_pass = bf16_bound_pass
        h = C.c_void_p()
        _check(lib().hkg_server_start(index._h, C.byref(so), C.byref(h)))
        self._h = h
        self.index = index

    def submit(self, q: np.ndarray, k: int) -> int:
        q = np.ascontiguousarray(q, np.float32)
        t = C.c_uint64()
        _check(lib().hkg_submit(self._h, _ptr(q), k, C.byref(t)))
        return t.value

    def poll(self, ticket: int, k: int):
        ids = np.empty(k, np.uint64)
        d2 = np.empty(k, np.float32)
        lat = C.c_double()
        rc = lib().hkg_poll(self._h, ticket, _ptr(ids), _ptr(d2), C.byref(lat))
        if rc == HKG_ENOTREADY:
            return None
        _check(rc)
        return ids, d2, lat.value

    def wait(self, ticket: int, k: int):
        ids = np.empty(k, np.uint64)
        d2 = np.empty(k, np.float32)
        lat = C.c_double()
        _check(lib().hkg_wait(self._h, ticket, _ptr(ids), _ptr(d2), C.byref(lat)))
        return ids, d2, lat.value

    def stop(self) -> None:
        if self._h and self._h.value:
            lib().hkg_server_stop(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
