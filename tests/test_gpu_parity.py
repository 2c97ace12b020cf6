"""GPU parity: the HIP paths against the oracle (bit-exact d2, equal ids)
and against the reference CLI's own answers.  Run on an MI355X."""
from __future__ import annotations

import os

import numpy as np
import pytest

import ref_runner as rr

pytestmark = pytest.mark.gpu

L256 = 256


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


def row_map(index_dir: str, X: np.ndarray) -> np.ndarray:
    """Original ordinal of every leaf-order row of leaves_raw.idx, by exact
    row match (the index stores no ids)."""
    L = X.shape[1]
    raw = np.fromfile(os.path.join(index_dir, "leaves_raw.idx"), np.float32).reshape(-1, L)
    lookup = {X[i].tobytes(): i for i in range(X.shape[0])}
    return np.array([lookup[r.tobytes()] for r in raw], np.uint64)


def assert_same(res_ids, res_d2, ids, d2, what: str) -> None:
    assert res_d2.dtype == np.float32
    bad = np.nonzero((res_d2.view(np.uint32) != d2.view(np.uint32)).any(1))[0]
    assert bad.size == 0, f"{what}: d2 differs for queries {bad[:10]}: {res_d2[bad[0]][:5]} vs {d2[bad[0]][:5]}"
    bad = np.nonzero((res_ids != ids).any(1))[0]
    assert bad.size == 0, f"{what}: ids differ for queries {bad[:10]}: {res_ids[bad[0]][:5]} vs {ids[bad[0]][:5]}"


# ---- generator ---------------------------------------------------------------
@pytest.mark.parametrize("kind,L", [(0, 256), (1, 96), (0, 96)])
def test_generator_bit_exact(hkg, oracle, kind, L):
    _say(f"generator kind={kind} L={L}")
    for first, n in [(0, 3000), (99_990_000, 500)]:
        got = hkg.gen_series(kind, 7, first, n, L)
        exp = oracle.series(kind, 7, first, n, L)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), (kind, L, first)


def test_noisy_queries_bit_exact(hkg, oracle):
    _say("noisy queries")
    sig = np.linspace(0.0, 0.2, 37).astype(np.float32)
    got, src = hkg.gen_noisy_queries(0, 1, 1_000_000, 3, sig, 256)
    exp, esrc = oracle.noisy_queries(0, 1, 1_000_000, 3, sig, 256)
    assert np.array_equal(src, esrc)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


# ---- full scan (PSCAN) ---------------------------------------------------------
@pytest.fixture(scope="module")
def small_data(oracle):
    X = oracle.series(0, 11, 0, 20000, L256)
    Q = oracle.series(0, 12, 0, 40, L256)
    # a few hard/easy queries: noisy copies of dataset rows
    Qn, _ = oracle.noisy_queries(0, 11, 20000, 13, np.array([0.0, 0.01, 0.1, 0.2] * 5, np.float32), L256)
    return X, np.concatenate([Q, Qn])


@pytest.mark.parametrize("k", [1, 7, 100])
def test_scan_matches_oracle(hkg, oracle, small_data, k):
    X, Q = small_data
    _say(f"scan k={k}")
    ids, d2 = oracle.knn(X, Q, k)
    with hkg.HerculesIndex.open_raw(X) as ix:
        r = ix.knn(Q, k, path="scan")
    assert_same(r.ids, r.d2, ids, d2, f"scan k={k}")


def test_scan_ties_and_padding(hkg, oracle):
    _say("scan ties")
    X = oracle.series(0, 21, 0, 3000, L256)
    X[2500] = X[17]          # exact duplicate: same d2 for two ids
    X[1200] = X[17]
    Q = np.ascontiguousarray(X[[17, 5, 2500]] + np.float32(0.001))
    ids, d2 = oracle.knn(X, Q, 5)
    assert ids[0, 0] == 17   # smallest id among equal distances
    with hkg.HerculesIndex.open_raw(X) as ix:
        r = ix.knn(Q, 5, path="scan")
        assert_same(r.ids, r.d2, ids, d2, "ties")
        # k larger than the number of distinct distances: padded ranks
        small = np.ascontiguousarray(X[:3])
    with hkg.HerculesIndex.open_raw(np.repeat(small, 2, axis=0)) as ix2:
        r2 = ix2.knn(small[:1], 8, path="scan")
    assert (r2.d2[0, 3:] == np.finfo(np.float32).max).all()
    assert (r2.ids[0, 3:] == np.uint64(2**64 - 1)).all()
    assert r2.ids[0, 0] == 0 and set(r2.ids[0, :3].tolist()) == {0, 2, 4}


# ---- index paths on a reference-built index ------------------------------------
@pytest.fixture(scope="module")
def ref_index(tmp_path_factory, small_data):
    if not rr.available():
        pytest.skip("reference binaries (oracle/_ref) not built")
    X, Q = small_data
    d = tmp_path_factory.mktemp("tiny")
    rr.write_f32(str(d / "data.bin"), X)
    rr.write_f32(str(d / "q.bin"), Q)
    rr.build_index(str(d / "data.bin"), X.shape[0], L256, 500, str(d / "idx"), threads=8)
    return str(d / "idx"), str(d / "q.bin")


# the batched path's two filters: per (query, leaf) item and leaf-ordered
PATHS = {"index": {"path": "index"}, "batched": {"path": "batched"},
         "batched-items": {"path": "batched", "batched_filter": "items"},
         "batched-leaf": {"path": "batched", "batched_filter": "leaf"}, "scan": {"path": "scan"}}


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("k", [1, 10])
def test_index_paths_match_oracle(hkg, oracle, small_data, ref_index, path, k):
    X, Q = small_data
    idx, _ = ref_index
    _say(f"index path={path} k={k}")
    ids, d2 = oracle.knn(X, Q, k)
    with hkg.HerculesIndex.open(idx) as ix:
        ix.set_ids(row_map(idx, X))
        r = ix.knn(Q, k, **PATHS[path])
    assert_same(r.ids, r.d2, ids, d2, f"{path} k={k}")


def test_index_matches_reference_cli(hkg, small_data, ref_index):
    """The reference's own multithreaded SIMS answers (exact d2 printed in
    hex by the oracle-side interposer) equal the GPU's."""
    X, Q = small_data
    idx, qpath = ref_index
    _say("reference cli")
    d2_ref, _, _, _ = rr.query(idx, qpath, Q.shape[0], L256, 10, threads=8, sims=True)
    with hkg.HerculesIndex.open(idx) as ix:
        r = ix.knn(Q, 10, path="batched")
    assert np.array_equal(r.d2.view(np.uint32), d2_ref.view(np.uint32))


def test_mixed_k_and_determinism(hkg, oracle, small_data, ref_index):
    X, Q = small_data
    idx, _ = ref_index
    _say("mixed k")
    kv = np.array([1, 3, 10, 100] * (Q.shape[0] // 4 + 1), np.uint32)[: Q.shape[0]]
    ids, d2 = oracle.knn(X, Q, 100)
    with hkg.HerculesIndex.open(idx) as ix:
        ix.set_ids(row_map(idx, X))
        for filt in ("items", "leaf"):
            r1 = ix.knn(Q, kv, path="batched", batched_filter=filt)
            r2 = ix.knn(Q, kv, path="batched", batched_filter=filt)
            assert np.array_equal(r1.ids, r2.ids) and np.array_equal(r1.d2, r2.d2)
            for i, k in enumerate(kv):
                assert np.array_equal(r1.d2[i, :k].view(np.uint32), d2[i, :k].view(np.uint32)), (filt, i)
                assert np.array_equal(r1.ids[i, :k], ids[i, :k]), (filt, i)
                assert (r1.d2[i, k:] == np.finfo(np.float32).max).all()


def test_overflow_rerun_stays_exact(hkg, oracle, small_data, ref_index):
    """A tiny candidate buffer forces overflows; the seeded re-runs (and the
    full-scan last resort) must still give the exact answers, on the
    leaf-ordered filter of a whole batch and on the per-item filter of small
    batches."""
    X, Q = small_data
    idx, _ = ref_index
    _say("overflow re-run")
    ids, d2 = oracle.knn(X, Q, 50)
    with hkg.HerculesIndex.open(idx) as ix:
        ix.set_ids(row_map(idx, X))
        r = ix.knn(Q, 50, path="batched", cand_cap=64, batched_filter="leaf")
        assert r.stats["overflow_retries"] > 0
        assert_same(r.ids, r.d2, ids, d2, "overflow")
        r = ix.knn(Q, 50, path="batched", batch=8, cand_cap=64, batched_filter="items")
        assert r.stats["overflow_retries"] > 0
        assert_same(r.ids, r.d2, ids, d2, "overflow, batches of 8")


def test_profile_stages(hkg, small_data, ref_index):
    X, Q = small_data
    idx, _ = ref_index
    with hkg.HerculesIndex.open(idx) as ix:
        r = ix.knn(Q, 10, path="batched", profile=True)
        s = ix.knn(Q, 10, path="scan", profile=True)
    assert r.stats["stage_ms"]["refine"] > 0 and r.stats["stage_launches"]["refine"] == 1
    assert r.stats["stage_bytes"]["refine"] > 0
    assert s.stats["stage_bytes"]["scan"] == ((Q.shape[0] + 3) // 4) * X.nbytes


# ---- BF16 certified-bound path ----------------------------------------------------
@pytest.mark.parametrize("bound_pass", [0, 1])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_bf16_path_matches_oracle(hkg, oracle, small_data, ref_index, k, bound_pass):
    X, Q = small_data
    idx, _ = ref_index
    _say(f"bf16 k={k} bound_pass={bound_pass}")
    ids, d2 = oracle.knn(X, Q, k)
    with hkg.HerculesIndex.open(idx, keep_bf16=True) as ix:
        ix.set_ids(row_map(idx, X))
        r = ix.knn(Q, k, path="bf16", bf16_bound_pass=bound_pass, profile=True)
        s = ix.knn(Q, k, path="batched")
    assert_same(r.ids, r.d2, ids, d2, f"bf16 k={k}")
    # pairs of the row tiles the two filter passes visited (128-row tiles;
    # skipped tiles excluded)
    assert 0 < r.stats["bf16_pairs"] <= 2 * Q.shape[0] * ((X.shape[0] + 127) // 128) * 128
    if bound_pass == 1:
        assert r.stats["stage_launches"]["bound"] == 1 and r.stats["stage_bytes"]["bound"] > 0
    assert r.stats["overflow_retries"] == 0
    # the certified bound prunes far more than the iSAX bound
    assert r.stats["series_refined"] <= s.stats["series_refined"] + Q.shape[0] * k
    # (query, leaf) pairs within the k-th bound after the approximate stage:
    # at least the best leaf of every query, at most every pair
    nleaf = ix_leaves(idx)
    for st in (r.stats, s.stats):
        assert Q.shape[0] < st["leaves_after_eapca"] <= Q.shape[0] * nleaf


@pytest.mark.parametrize("path", ["scan", "index", "batched", "bf16"])
def test_ids_beyond_32_bits(hkg, oracle, small_data, ref_index, path):
    """Public ids of 2^32 and above (the reference's positions are 64-bit):
    an id map of original ordinal * 2^33 + 7 (the same order) returns the
    oracle's ids mapped the same way, with the same d2."""
    X, Q = small_data
    idx, _ = ref_index
    k = 10
    _say(f"64-bit ids path={path}")
    ids, d2 = oracle.knn(X, Q, k)
    big = lambda v: v.astype(np.uint64) * np.uint64(1 << 33) + np.uint64(7)
    with hkg.HerculesIndex.open(idx, keep_bf16=path == "bf16") as ix:
        ix.set_ids(big(row_map(idx, X)))
        r = ix.knn(Q, k, path=path)
    assert int(r.ids.max()) >= 1 << 33
    assert_same(r.ids, r.d2, big(ids), d2, f"64-bit ids {path}")


def ix_leaves(idx: str) -> int:
    from hercules_amd import hkg
    return int(hkg.inspect_dir(idx)["n_leaves"])


def test_device_entry_rejects_bad_k_and_path(hkg, small_data, ref_index):
    """hkg_knn_device checks every per-query k on the device (values
    outside [1, kmax] are clamped for the kernels and reported as
    HKG_EINVAL), and an unknown path before any launch."""
    import ctypes as C
    X, Q = small_data
    idx, _ = ref_index
    _say("bad k / bad path")
    nq, kmax = 8, 5
    dQ = hkg.DeviceArray(nq * L256 * 4)
    dQ.upload(np.ascontiguousarray(Q[:nq]))
    dids, dd2 = hkg.DeviceArray(nq * kmax * 8), hkg.DeviceArray(nq * kmax * 4)
    with hkg.HerculesIndex.open(idx) as ix:
        for path in ("batched", "bf16", "scan", "index"):
            for bad in (0, kmax + 1):
                kv = np.full(nq, 3, np.uint32)
                kv[5] = bad
                dk = hkg.DeviceArray(kv.nbytes)
                dk.upload(kv)
                with pytest.raises(hkg.HkgError) as e:
                    ix.knn_device(dQ, nq, dk, kmax, dids, dd2, path=path)
                assert e.value.status == hkg.HKG_EINVAL and "outside [1, kmax" in str(e.value)
        # good k after the bad calls: the answers are the host entry's
        kv = np.full(nq, 3, np.uint32)
        dk = hkg.DeviceArray(kv.nbytes)
        dk.upload(kv)
        ix.knn_device(dQ, nq, dk, kmax, dids, dd2, path="batched")
        r = ix.knn(Q[:nq], kv, path="batched")
        got = dd2.download(np.empty((nq, kmax), np.float32))
        assert np.array_equal(got[:, :3].view(np.uint32), r.d2[:, :3].view(np.uint32))
        qo = hkg.query_opts("batched")
        qo.path = 9
        rc = hkg.lib().hkg_knn_device(ix._h, dQ.ptr, nq, dk.ptr, kmax, C.byref(qo), dids.ptr, dd2.ptr, None, None)
        assert rc == hkg.HKG_EINVAL
        qo = hkg.query_opts("batched")
        qo.batched_filter = 3
        rc = hkg.lib().hkg_knn_device(ix._h, dQ.ptr, nq, dk.ptr, kmax, C.byref(qo), dids.ptr, dd2.ptr, None, None)
        assert rc == hkg.HKG_EINVAL and b"batched_filter" in hkg.lib().hkg_last_error()


def test_bf16_embeddings_l96(hkg, oracle, tmp_path):
    """L = 96 embedding-like series (N(0,1) -> L2 -> z-normalised)."""
    if not rr.available():
        pytest.skip("reference binaries not built")
    _say("bf16 L=96")
    X = oracle.series(1, 21, 0, 20000, 96)
    Qn, _ = oracle.noisy_queries(1, 21, 20000, 23, np.array([0.05, 0.1, 0.2, 0.4] * 5, np.float32), 96)
    Q = np.ascontiguousarray(np.concatenate([oracle.series(1, 22, 0, 20, 96), Qn]))
    rr.write_f32(str(tmp_path / "d.bin"), X)
    rr.build_index(str(tmp_path / "d.bin"), X.shape[0], 96, 400, str(tmp_path / "idx"), threads=8)
    ids, d2 = oracle.knn(X, Q, 10)
    with hkg.HerculesIndex.open(str(tmp_path / "idx")) as ix:
        ix.set_ids(row_map(str(tmp_path / "idx"), X))
        for path in ("bf16", "batched", "index"):
            r = ix.knn(Q, 10, path=path)
            assert_same(r.ids, r.d2, ids, d2, f"L96 {path}")


@pytest.mark.parametrize("kind,L", [(0, 256), (1, 96), (0, 128)])
def test_bf16_query_tile_shapes_agree(hkg, oracle, kind, L):
    """Every query-tile shape of the BF16 filter (2 or 3 32-query MFMA tiles
    per wave, 4 for L <= 128) gives the oracle's answers, bit for bit, with
    batches that fill whole block groups and batches that leave the last
    group partial; a shape not compiled for the length is refused."""
    import ctypes as C
    n = 40000
    _say(f"bf16 query tiles kind={kind} L={L}")
    X = oracle.series(kind, 31, 0, n, L)
    rng = np.random.default_rng(32)
    nq = 900
    Q, _ = oracle.noisy_queries(kind, 31, n, 33, rng.uniform(0.0, 0.2, nq).astype(np.float32), L)
    kv = rng.integers(1, 40, nq).astype(np.uint32)
    ids, d2 = oracle.knn(X, Q, 40)
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=700) as ix:
        for qt in ((2, 3) if L > 128 else (2, 3, 4)):
            for batch in (0, 128 * qt, 333):
                r = ix.knn(Q, kv, path="bf16", batch=batch, bf16_query_tiles=qt)
                for i in range(nq):
                    k = int(kv[i])
                    assert np.array_equal(r.d2[i, :k].view(np.uint32), d2[i, :k].view(np.uint32)), (qt, batch, i)
                    assert np.array_equal(r.ids[i, :k], ids[i, :k]), (qt, batch, i)
        qo = hkg.query_opts("bf16", 0, False, 0, 0, bf16_query_tiles=4 if L > 128 else 5)
        out_i = np.empty((4, 5), np.uint64)
        out_d = np.empty((4, 5), np.float32)
        rc = hkg.lib().hkg_knn(ix._h, Q[:4].ctypes.data, 4, None, 5, C.byref(qo), out_i.ctypes.data,
                               out_d.ctypes.data, None)
        assert rc == hkg.HKG_EINVAL and b"bf16_query_tiles" in hkg.lib().hkg_last_error()


def test_bf16_overflow_rerun_stays_exact(hkg, oracle, small_data, ref_index):
    """A tiny survivor list forces the seeded BF16 re-runs."""
    X, Q = small_data
    idx, _ = ref_index
    _say("bf16 overflow re-run")
    ids, d2 = oracle.knn(X, Q, 100)
    with hkg.HerculesIndex.open(idx) as ix:
        ix.set_ids(row_map(idx, X))
        r = ix.knn(Q, 100, path="bf16", cand_cap=128)
    assert r.stats["overflow_retries"] > 0
    assert_same(r.ids, r.d2, ids, d2, "bf16 overflow")
