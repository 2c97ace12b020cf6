"""The launch-shape autotuner on an MI355X: for each path and series length
it returns one of its candidate shapes, the answers after tuning equal
those before it bit for bit, and later calls that leave the batch at 0 run
in the tuned batch size."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CANDIDATE_BATCHES = {"batched": {256, 512, 1024, 2048}, "bf16": {512, 768, 1024, 1536, 2048, 3072, 4096},
                     "scan": {1, 4, 16, 64, 256}, "index": {1}}
# BF16 filter shapes: 32-query MFMA tiles per wave, and the batches timed with each
BF16_SHAPES = {2: {512, 1024, 2048, 4096}, 3: {768, 1536, 3072}, 4: {1024, 2048, 4096}}


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


@pytest.mark.parametrize("kind,L", [(0, 256), (1, 96)])
def test_autotune_paths(hkg, oracle, kind, L):
    n = 300_000
    dX = hkg.DeviceArray(n * L * 4)
    hkg.gen_series_device(kind, 11, 0, n, L, dX)
    rng = np.random.default_rng(12)
    nq = 1200
    sig = rng.uniform(0.0, 0.2, nq).astype(np.float32)
    Q, _ = hkg.gen_noisy_queries(kind, 11, n, 13, sig, L)
    kv = rng.integers(1, 65, nq).astype(np.uint32)
    with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=3000) as ix:
        for path, m in [("batched", nq), ("bf16", nq), ("scan", 256), ("index", 48)]:
            print(f"autotune L={L} path={path} on {m} queries", flush=True)
            before = ix.knn(Q[:m], kv[:m], path=path)
            t = ix.autotune(Q[:m], kv[:m], path=path)
            print(f"  -> {t}", flush=True)
            assert t["path"] == path and t["length"] == L
            assert t["batch"] in CANDIDATE_BATCHES[path]
            if path == "batched":
                # the per-item filter with its grid, or the leaf-ordered one (no grid)
                assert t["batched_filter"] in ("items", "leaf") and t["scan_blocks_per_cu"] == 0
                assert t["filter_blocks_per_cu"] in ((2, 4, 8) if t["batched_filter"] == "items" else (0,))
            elif path == "index":
                assert t["filter_blocks_per_cu"] in (2, 4, 8) and t["scan_blocks_per_cu"] == 0
            elif path == "scan":
                assert t["scan_blocks_per_cu"] in (3, 6, 8) and t["filter_blocks_per_cu"] == 0
            else:
                assert t["filter_blocks_per_cu"] == 0 and t["scan_blocks_per_cu"] == 0
                # a query-tile shape compiled for this length, at one of its batches
                assert t["bf16_query_tiles"] in ((2, 3) if L > 128 else (2, 3, 4))
                assert t["batch"] in BF16_SHAPES[t["bf16_query_tiles"]]
            if path != "bf16":
                assert t["bf16_query_tiles"] == 0
            assert t["candidates"] >= 1 and t["ms_per_query"] > 0 and t["ms_per_query_default"] > 0
            # the largest batch within 1% of the fastest shape is kept
            assert t["ms_per_query"] <= t["ms_per_query_default"] * 1.01
            after = ix.knn(Q[:m], kv[:m], path=path)
            assert np.array_equal(after.d2.view(np.uint32), before.d2.view(np.uint32)), path
            assert np.array_equal(after.ids, before.ids), path
            assert after.stats["launches"] == -(-m // t["batch"]), (path, after.stats["launches"])
        # the exact answers, checked by the oracle on a few queries
        sel = np.arange(0, nq, 150)
        r = ix.knn(Q[sel], 10, path="batched")
        for j, qi in enumerate(sel):
            for rank in range(10):
                row = oracle.series(kind, 11, int(r.ids[j, rank]), 1, L, threads=1)[0]
                assert oracle.ed2(Q[qi], row) == r.d2[j, rank]


def test_autotune_rejects_bad_arguments(hkg):
    X = np.ascontiguousarray(np.random.default_rng(0).standard_normal((4000, 64)), np.float32)
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    with hkg.HerculesIndex.build_device(dX, 4000, 64, leaf_size=500) as ix:
        t = hkg.Tuning()
        assert hkg.lib().hkg_autotune(ix._h, X.ctypes.data, 0, None, 5, 2, C.byref(t)) == hkg.HKG_EINVAL
        assert hkg.lib().hkg_autotune(ix._h, X.ctypes.data, 4, None, 5, 9, C.byref(t)) == hkg.HKG_EINVAL
        assert hkg.lib().hkg_autotune(ix._h, X.ctypes.data, 4, None, 0, 2, C.byref(t)) == hkg.HKG_EINVAL
        assert hkg.lib().hkg_autotune(ix._h, X.ctypes.data, 4, None, 5, 2, None) == hkg.HKG_EINVAL
