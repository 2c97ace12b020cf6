"""ctypes bindings to the checkers (test infrastructure only).

* ``Oracle``  -- oracle/liboracle.so, the C restatement of the reference's
  arithmetic and the workload generator specification.
* ``RefLib``  -- oracle/_ref/libherc_ref.so, the reference's own functions
  compiled from its sources (present only where the reference was built).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


class Oracle:
    def __init__(self) -> None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            build_oracle()
        L = C.CDLL(path)
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_draw.restype = C.c_uint64
        L.orc_draw.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_log.restype = C.c_double
        L.orc_log.argtypes = [C.c_double]
        L.orc_gen_series.restype = None
        L.orc_gen_series.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, _f32p, C.c_int]
        L.orc_gen_noisy_queries.restype = None
        L.orc_gen_noisy_queries.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                                            C.c_uint32, _f32p, _f32p, _u64p]
        L.orc_ed2.restype = C.c_float
        L.orc_ed2.argtypes = [_f32p, _f32p, C.c_int, C.c_float]
        L.orc_ed2_f64.restype = C.c_double
        L.orc_ed2_f64.argtypes = [_f32p, _f32p, C.c_int]
        L.orc_paa.restype = None
        L.orc_paa.argtypes = [_f32p, C.c_int, C.c_int, _f32p]
        L.orc_breakpoints256.restype = None
        L.orc_breakpoints256.argtypes = [_f32p]
        L.orc_sax_word.restype = None
        L.orc_sax_word.argtypes = [_f32p, C.c_int, _f32p, _u8p]
        L.orc_lb_sax.restype = C.c_float
        L.orc_lb_sax.argtypes = [_f32p, _u8p, C.c_int, _f32p]
        L.orc_seg_mean_std.restype = None
        L.orc_seg_mean_std.argtypes = [_f32p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_lb_eapca.restype = C.c_float
        L.orc_lb_eapca.argtypes = [_f32p, C.c_int, _i16p, _f32p, C.c_int]
        L.orc_knn_bruteforce.restype = C.c_int
        L.orc_knn_bruteforce.argtypes = [_f32p, C.c_uint64, C.c_int, _f32p, C.c_uint32, C.c_uint32,
                                         _u64p, _f32p, C.c_int]
        _i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        _u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        L.orc_knn_mode.restype = C.c_int
        L.orc_knn_mode.argtypes = [_i64p, _f32p, C.c_uint32, _i16p, _f32p, _f32p, _u8p, C.c_int, _f32p,
                                   C.c_uint32, C.c_uint32, C.c_float, C.c_int, C.c_uint32, C.c_float, C.c_float,
                                   _u64p, _f32p, _u32p, C.c_int]
        self.lib = L
        bp = np.zeros(255, np.float32)
        L.orc_breakpoints256(bp)
        self.bp = bp

    # generator -------------------------------------------------------------
    def series(self, kind: int, seed: int, first: int, n: int, L: int, threads: int = 8) -> np.ndarray:
        out = np.empty((n, L), np.float32)
        self.lib.orc_gen_series(kind, seed, first, n, L, out, threads)
        return out

    def noisy_queries(self, kind: int, data_seed: int, N: int, query_seed: int, sigmas, L: int):
        sig = np.ascontiguousarray(sigmas, np.float32)
        nq = sig.shape[0]
        out = np.empty((nq, L), np.float32)
        src = np.empty(nq, np.uint64)
        self.lib.orc_gen_noisy_queries(kind, data_seed, N, query_seed, nq, L, sig, out, src)
        return out, src

    # arithmetic --------------------------------------------------------------
    def ed2(self, q, x, bound=np.float32(np.finfo(np.float32).max)) -> np.float32:
        return np.float32(self.lib.orc_ed2(q, x, q.shape[0], bound))

    def paa(self, x, w=16) -> np.ndarray:
        out = np.empty(w, np.float32)
        self.lib.orc_paa(x, x.shape[0], w, out)
        return out

    def sax(self, paa) -> np.ndarray:
        out = np.empty(paa.shape[0], np.uint8)
        self.lib.orc_sax_word(paa, paa.shape[0], self.bp, out)
        return out

    def lb_sax(self, paa, sax, L) -> np.float32:
        return np.float32(self.lib.orc_lb_sax(paa, sax, L, self.bp))

    def seg_mean_std(self, x, s, e):
        mu, sd = C.c_float(), C.c_float()
        self.lib.orc_seg_mean_std(x, s, e, C.byref(mu), C.byref(sd))
        return np.float32(mu.value), np.float32(sd.value)

    def lb_eapca(self, q, seg_end, ind, use_fma=1) -> np.float32:
        return np.float32(self.lib.orc_lb_eapca(q, seg_end.shape[0], seg_end, ind, use_fma))

    def knn(self, X, Q, k, threads=8):
        nq = Q.shape[0]
        ids = np.empty((nq, k), np.uint64)
        d2 = np.empty((nq, k), np.float32)
        rc = self.lib.orc_knn_bruteforce(X, X.shape[0], X.shape[1], Q, nq, k, ids, d2, threads)
        assert rc == 0
        return ids, d2

    def knn_mode(self, tables: dict, Xleaf, sax, Q, k, epsilon=0.0, approx_budget=80, nprobes=0,
                 eapca_th=0.25, sax_th=0.9, threads=8):
        """The reference's epsilon / ng search over an index tree (tables from
        root_idx.node_tables); Xleaf and sax (n x 16 u8) in leaves_raw.idx
        order.  Returns ids (rows of Xleaf), d2 and per query {leaves scanned
        first, candidate leaves, candidate branch, candidate leaves scanned}."""
        nq = Q.shape[0]
        ids = np.empty((nq, k), np.uint64)
        d2 = np.empty((nq, k), np.float32)
        info = np.zeros((nq, 4), np.uint32)
        nd = tables["nodes"]
        rc = self.lib.orc_knn_mode(nd, tables["split_value"], nd.shape[0], tables["seg_end"], tables["ind"],
                                   Xleaf, np.ascontiguousarray(sax, np.uint8).reshape(-1), Xleaf.shape[1], Q, nq, k,
                                   float(epsilon), int(approx_budget), int(nprobes), float(eapca_th), float(sax_th),
                                   ids, d2, info, threads)
        assert rc == 0
        return ids, d2, info


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libherc_ref.so"))


class RefLib:
    """The reference's own functions, compiled from its sources."""

    def __init__(self) -> None:
        L = C.CDLL(os.path.join(REF_DIR, "libherc_ref.so"))
        L.ts_euclidean_distance_SIMD.restype = C.c_float
        L.ts_euclidean_distance_SIMD.argtypes = [_f32p, _f32p, C.c_int, C.c_float]
        L.paa_from_ts.restype = C.c_int
        L.paa_from_ts.argtypes = [_f32p, _f32p, C.c_int, C.c_int]
        L.sax_from_ts.restype = C.c_int
        L.sax_from_ts.argtypes = [_f32p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.minidist_paa_to_isax_raw_SIMD.restype = C.c_float
        L.minidist_paa_to_isax_raw_SIMD.argtypes = [_f32p, _u8p, _u8p, C.c_ubyte, C.c_int, C.c_int,
                                                    C.c_int, C.c_int, C.c_float]
        L.calc_mean_stdev_SIMD.restype = None
        L.calc_mean_stdev_SIMD.argtypes = [_f32p, C.c_int, C.c_int, C.POINTER(C.c_float),
                                           C.POINTER(C.c_float)]
        self.lib = L

    def ed2(self, q, x, bound=np.float32(np.finfo(np.float32).max)):
        return np.float32(self.lib.ts_euclidean_distance_SIMD(q, x, q.shape[0], bound))

    def paa(self, x, w=16):
        out = np.zeros(w, np.float32)
        self.lib.paa_from_ts(x, out, w, x.shape[0] // w)
        return out

    def sax(self, x, w=16):
        out = np.zeros(w, np.uint8)
        self.lib.sax_from_ts(x, out, x.shape[0] // w, w, 256, 8)
        return out

    def lb_sax(self, paa, sax, L):
        card = np.full(16, 8, np.uint8)
        return np.float32(self.lib.minidist_paa_to_isax_raw_SIMD(
            paa, sax, card, 8, 256, 16, -2000000, 2000000, float(L // 16)))

    def seg_mean_std(self, x, s, e):
        mu, sd = C.c_float(), C.c_float()
        self.lib.calc_mean_stdev_SIMD(x, s, e, C.byref(mu), C.byref(sd))
        return np.float32(mu.value), np.float32(sd.value)

    def breakpoints256(self) -> np.ndarray:
        """Recover the reference's card-256 breakpoints through its own
        sax_from_ts: with one value per segment the symbol is the number of
        breakpoints strictly below the value, so breakpoint j is the largest
        float whose symbol is <= j (bisection over float bit patterns)."""
        def sym(p: np.float32) -> int:
            x = np.full(16, p, np.float32)
            out = np.zeros(16, np.uint8)
            self.lib.sax_from_ts(x, out, 1, 16, 256, 8)
            return int(out[0])

        def f2o(f):  # order-preserving int of a float
            i = int(np.float32(f).view(np.int32))
            return i if i >= 0 else -(i & 0x7FFFFFFF)

        def o2f(o):
            i = o if o >= 0 else (-o) | (-0x80000000)
            return np.int32(i).view(np.float32) if o >= 0 else np.array([i], np.int64).astype(np.uint32).view(np.float32)[0]

        bps = np.empty(255, np.float32)
        for j in range(255):
            lo, hi = f2o(np.float32(-10)), f2o(np.float32(10))  # sym(lo) <= j < sym(hi)
            while hi - lo > 1:
                mid = (lo + hi) // 2
                if sym(o2f(mid)) <= j:
                    lo = mid
                else:
                    hi = mid
            bps[j] = o2f(lo)
        return bps
