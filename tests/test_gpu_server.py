"""Serving loop and captured-graph replays on an MI355X.

The server answers each request exactly as the same query alone: a query's
result does not depend on the batch it was admitted into.  Graph replays of
the batched path equal eager runs, with eager work between the replays."""
from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 256


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


@pytest.fixture(scope="module")
def setup(hkg, oracle):
    n = 30000
    X = oracle.series(0, 31, 0, n, L)
    sig = np.linspace(0.0, 0.2, 96).astype(np.float32)
    Qn, _ = oracle.noisy_queries(0, 31, n, 32, sig, L)
    Q = np.ascontiguousarray(np.concatenate([Qn, oracle.series(0, 33, 0, 32, L)]))
    kv = np.array([1, 2, 5, 10, 37, 100, 3, 64] * (Q.shape[0] // 8), np.uint32)
    ids, d2 = oracle.knn(X, Q, int(kv.max()))
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    ix = hkg.HerculesIndex.build_device(dX, n, L, leaf_size=600)
    yield ix, Q, kv, ids, d2
    ix.close()


def _check(i, k, rid, rd2, ids, d2, what):
    assert np.array_equal(rd2.view(np.uint32), d2[i, :k].view(np.uint32)), f"{what}: query {i} d2"
    assert np.array_equal(rid, ids[i, :k]), f"{what}: query {i} ids"


@pytest.mark.parametrize("path", ["batched", "bf16"])
def test_server_same_alone_and_in_the_mix(hkg, setup, path):
    ix, Q, kv, ids, d2 = setup
    _say(f"server {path}: alone")
    srv = hkg.Server(ix, max_batch=64, path=path, max_inflight=2, max_k=100)
    try:
        alone = []
        for i in range(0, Q.shape[0], 7):
            t = srv.submit(Q[i], int(kv[i]))
            rid, rd2, lat = srv.wait(t, int(kv[i]))
            assert lat > 0
            _check(i, int(kv[i]), rid, rd2, ids, d2, f"{path} alone")
            alone.append((i, rid, rd2))
        _say(f"server {path}: mixed, 4 client threads")
        out = {}
        errs = []

        def client(rows):
            try:
                tickets = [(i, srv.submit(Q[i], int(kv[i]))) for i in rows]
                for i, t in tickets:
                    out[i] = srv.wait(t, int(kv[i]))
            except Exception as e:  # surfaced below
                errs.append(e)

        rows = np.arange(Q.shape[0])
        ths = [threading.Thread(target=client, args=(rows[j::4],)) for j in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert not errs, errs
        assert len(out) == Q.shape[0]
        for i, (rid, rd2, lat) in out.items():
            _check(i, int(kv[i]), rid, rd2, ids, d2, f"{path} mixed")
        for i, rid, rd2 in alone:
            assert np.array_equal(out[i][0], rid) and np.array_equal(out[i][1].view(np.uint32), rd2.view(np.uint32))
    finally:
        srv.stop()


def test_server_trace_and_top_k_class(hkg, setup, monkeypatch, capfd):
    """HKG_SERVER_TRACE prints one line per batch; a batch of small-k
    requests that fits the scratch share at max_k runs at max_k (the sets
    prepared at start), and its answers are those of each query alone."""
    ix, Q, kv, ids, d2 = setup
    _say("server trace, k class")
    monkeypatch.setenv("HKG_SERVER_TRACE", "1")
    srv = hkg.Server(ix, max_batch=64, path="bf16", max_inflight=1, max_k=100)
    try:
        rows = list(range(0, 40))
        tickets = [(i, srv.submit(Q[i], 2)) for i in rows]
        for i, t in tickets:
            rid, rd2, _ = srv.wait(t, 2)
            _check(i, 2, rid, rd2, ids, d2, "trace, k = 2")
    finally:
        srv.stop()
    err = capfd.readouterr().err
    lines = [l for l in err.splitlines() if l.startswith("hkg server: worker")]
    assert lines, err[-2000:]
    assert all(" k class 100:" in l for l in lines), lines


def test_server_rejects_bad_requests(hkg, setup):
    ix, Q, kv, _, _ = setup
    srv = hkg.Server(ix, max_batch=16, max_k=10)
    try:
        with pytest.raises(hkg.HkgError):
            srv.submit(Q[0], 0)
        with pytest.raises(hkg.HkgError):
            srv.submit(Q[0], 11)
        with pytest.raises(hkg.HkgError):
            srv.wait(123456789, 1)
        t = srv.submit(Q[0], 3)
        rid, rd2, _ = srv.wait(t, 3)
        assert rid.shape == (3,)
    finally:
        srv.stop()


@pytest.mark.parametrize("path,filt", [("batched", "items"), ("batched", "leaf"), ("bf16", "auto")])
def test_graph_replays_equal_eager_runs(hkg, setup, path, filt):
    """batch == number of queries: the batched paths (FP32 with either
    filter, and BF16) capture their stage sequence once and replay it, with
    the stage stamps recorded by the graph when profiling; every replay, with
    eager work between, equals the eager result (a batch larger than the
    queries runs eagerly) and the oracle."""
    ix, Q, kv, ids, d2 = setup
    n = 64
    Qg, kg = np.ascontiguousarray(Q[:n]), np.ascontiguousarray(kv[:n])
    eager = ix.knn(Qg, kg, path=path, batch=n + 1, profile=True, batched_filter=filt)
    for rep in range(4):
        prof = rep % 2 == 1
        _say(f"{path} ({filt}): graph replay {rep} (profile {prof})")
        r = ix.knn(Qg, kg, path=path, batch=n, profile=prof, batched_filter=filt)
        assert np.array_equal(r.ids, eager.ids) and np.array_equal(r.d2.view(np.uint32), eager.d2.view(np.uint32))
        for i in range(n):
            _check(i, int(kg[i]), r.ids[i, :kg[i]], r.d2[i, :kg[i]], ids, d2, f"{path} replay {rep}")
        if prof:
            # the graph's own events time its stages
            assert r.stats["stage_ms"]["refine"] > 0 and r.stats["stage_launches"]["refine"] == 1
            assert r.stats["stage_ms"]["topk"] > 0
        # eager work of other shapes between replays
        ix.knn(Q[n:n + 17], 5, path="scan")
        ix.knn(Q[n:n + 9], 7, path=path, profile=True, batched_filter=filt)


@pytest.mark.parametrize("path", ["batched", "bf16"])
def test_buffer_sets_evicted_and_rebuilt_keep_answers(hkg, setup, path):
    """More buffer-set shapes than the index keeps alive (each distinct
    (batch, k) leases its own set; the least recently used idle one is
    dropped): every call, on a fresh or a re-allocated set, returns the
    oracle's answers."""
    ix, Q, kv, ids, d2 = setup
    _say(f"buffer-set eviction path={path}")
    shapes = [(b, k) for b in (16, 48, 80) for k in (1, 2, 5, 10, 37, 64)]  # 18 shapes > 16 kept
    for rep in range(2):
        for b, k in shapes:
            r = ix.knn(Q[:b], k, path=path, batch=b)
            assert np.array_equal(r.d2.view(np.uint32), d2[:b, :k].view(np.uint32)), (rep, b, k)
            assert np.array_equal(r.ids, ids[:b, :k]), (rep, b, k)
