"""GPU parity at the workloads' full sizes, where the oracle cannot scan the
dataset in seconds:

- 25M x 256 random walks, 1,000 1-NN queries: dataset series plus noise of
  sigma 0.01, 0.05 and 0.1 (333 / 333 / 334), renormalised;
- 50M x 256 random walks, 10,000 queries with sigma ~ U[0.01, 0.2] and k
  drawn from {1, 10, 100}, on the batch sizes and kernel shapes the autotuner
  picks for the BF16 and the batched path;
- 100M x 96 embeddings (N(0,1) -> L2 -> z-normalised), 10,000 10-NN
  queries of the same kind (sigma ~ U[0.01, 0.2]);
- 100M x 256 random walks (the serving workload), 4,096 queries with
  sigma ~ U[0, 0.2] and k ~ U{1..100}, plus a Poisson burst of them through
  the serving loop.

The checks are properties that hold at any size:
- the BF16 path answers every query as the batched path does, bit for bit
  (on the embeddings, whose iSAX bound prunes little, on a sample), and
  both equal the plain per-query index path and the full scan (PSCAN) on a
  sample;
- every returned id's squared distance, recomputed by the oracle from the
  generator's copy of that row (the generator is a function of (seed, id)),
  is the returned d2, no rank repeats a distance, and the first neighbour
  is never farther than the row the query was made from;
- dataset rows asked as queries come back first at distance 0 under their
  own id;
- FP64 distances of the returned pairs stay within FP32 rounding of d2 (the
  error per result), and the bounds prune most series;
- epsilon answers stay within (1 + epsilon) of the exact ones rank by rank,
  and no-guarantee answers are never below them;
- served requests (continuous batching, mixed k) get the answers of the
  same queries asked alone.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLT_MAX_BITS = np.float32(np.finfo(np.float32).max).view(np.uint32)


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


WORKLOADS = {
    "25Mx256_k1": dict(kind=0, n=25_000_000, L=256, leaf=10000, nq=1000, qseed=3, batched_all=True,
                       sigma=lambda rng, nq: np.repeat(np.float32([0.01, 0.05, 0.1]), [333, 333, 334]),
                       k=lambda rng, nq: np.ones(nq, np.uint32)),
    "50Mx256_kmix": dict(kind=0, n=50_000_000, L=256, leaf=10000, nq=10000, qseed=4, batched_all=True, tune=True,
                         sigma=lambda rng, nq: rng.uniform(0.01, 0.2, nq).astype(np.float32),
                         k=lambda rng, nq: rng.choice(np.uint32([1, 10, 100]), nq)),
    "100Mx96_k10": dict(kind=1, n=100_000_000, L=96, leaf=10000, nq=10000, qseed=6, batched_all=False,
                        sigma=lambda rng, nq: rng.uniform(0.01, 0.2, nq).astype(np.float32),
                        k=lambda rng, nq: np.full(nq, 10, np.uint32)),
    "100Mx256_kU100": dict(kind=0, n=100_000_000, L=256, leaf=10000, nq=4096, qseed=7, batched_all=True,
                           serve=512,
                           sigma=lambda rng, nq: rng.uniform(0.0, 0.2, nq).astype(np.float32),
                           k=lambda rng, nq: rng.integers(1, 101, nq).astype(np.uint32)),
}


def _same(a, b, sel, what):
    bad = np.nonzero((a.d2[sel].view(np.uint32) != b.d2[sel].view(np.uint32)).any(1) |
                     (a.ids[sel] != b.ids[sel]).any(1))[0]
    assert bad.size == 0, f"{what}: {bad.size} queries differ, first {np.asarray(sel)[bad[:5]]}"


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_properties(hkg, oracle, name):
    w = WORKLOADS[name]
    kind, n, L = w["kind"], w["n"], w["L"]
    rng = np.random.default_rng(w["qseed"])
    sig = w["sigma"](rng, w["nq"])
    kq = w["k"](rng, w["nq"])
    kmax = int(kq.max())
    t0 = time.perf_counter()
    _say(f"sizes {name}: generating {n} x {L} (kind {kind}, seed 1) on the device")
    dX = hkg.DeviceArray(n * L * 4)
    hkg.gen_series_device(kind, 1, 0, n, L, dX)
    Q, src = hkg.gen_noisy_queries(kind, 1, n, w["qseed"], sig, L)
    with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=w["leaf"]) as ix:
        _say(f"sizes {name}: built, {ix.info()['n_leaves']} leaves, {time.perf_counter() - t0:.1f} s")

        if w.get("tune"):
            # the batched paths in the shapes the autotuner picks on this index
            # (later calls leave the batch at 0: tuned batch, BF16 query tiles,
            # FP32 filter; full batches replay their captured graph)
            for path in ("bf16", "batched"):
                t = ix.autotune(Q[:2048], kq[:2048], path=path)
                _say(f"sizes {name}: tuned {path}: {t}")
                assert t["candidates"] > 1 and t["ms_per_query"] <= t["ms_per_query_default"] * 1.01
        # every query on the BF16 path, and on the batched path where its iSAX
        # bound prunes (random walks; on the embeddings it would scan nearly
        # every series per query, so there it runs on the sample)
        rf = ix.knn(Q, kq, path="bf16", profile=True)
        _say(f"sizes {name}: bf16 path on {w['nq']} queries, {time.perf_counter() - t0:.1f} s")
        sample = np.unique(np.concatenate([np.arange(0, w["nq"], max(1, w["nq"] // 64)),
                                           np.nonzero(kq == kmax)[0][:16], np.argsort(sig)[-16:]]))
        Qs = np.ascontiguousarray(Q[sample])
        if w["batched_all"]:
            rb = ix.knn(Q, kq, path="batched")
            _say(f"sizes {name}: batched path on {w['nq']} queries, {time.perf_counter() - t0:.1f} s")
            _same(rf, rb, slice(None), f"{name} bf16 vs batched")
            others = []
        else:
            others = [("batched", ix.knn(Qs, kq[sample], path="batched"))]
        # the plain per-query path and the full scan on the sample
        others += [("index", ix.knn(Qs, kq[sample], path="index")), ("scan", ix.knn(Qs, kq[sample], path="scan"))]
        _say(f"sizes {name}: other paths on {sample.size} queries, {time.perf_counter() - t0:.1f} s")
        for what, r in others:
            bad = np.nonzero((r.d2.view(np.uint32) != rf.d2[sample].view(np.uint32)).any(1) |
                             (r.ids != rf.ids[sample]).any(1))[0]
            assert bad.size == 0, f"{name} {what} vs bf16: queries {sample[bad[:5]]}"

        # ids and distances agree; ranks are distinct and increasing; unfilled
        # ranks are marked; the first neighbour is never farther than the
        # query's source row; FP64 error per result
        worst = 0.0
        for qi in sample[:40]:
            k = int(kq[qi])
            d2 = rf.d2[qi]
            assert (d2[k:].view(np.uint32) == FLT_MAX_BITS).all() and (rf.ids[qi, k:] == np.uint64(2**64 - 1)).all()
            assert (np.diff(d2[:k]) > 0).all(), (name, qi)
            srow = oracle.series(kind, 1, int(src[qi]), 1, L, threads=1)[0]
            assert d2[0] <= oracle.ed2(Q[qi], srow), (name, qi)
            for r in range(k):
                row = oracle.series(kind, 1, int(rf.ids[qi, r]), 1, L, threads=1)[0]
                assert oracle.ed2(Q[qi], row) == d2[r], (name, qi, r)
                e64 = oracle.lib.orc_ed2_f64(Q[qi], row, L)
                worst = max(worst, abs(float(d2[r]) - e64) / max(e64, 1e-30))
        _say(f"sizes {name}: worst FP64 relative error of a returned d2: {worst:.3g}")
        assert worst < (L + 4) * 2.0 ** -24  # L terms summed in FP32, plus the difference and square

        # rows asked as queries are found at distance 0 under their own id
        probes = np.unique(np.concatenate([np.linspace(0, n - 1, 24).astype(np.int64), src[:8].astype(np.int64)]))
        P = np.concatenate([oracle.series(kind, 1, int(p), 1, L, threads=1) for p in probes])
        rp = ix.knn(P, 1, path="bf16")
        assert (rp.d2[:, 0] == 0.0).all() and (rp.ids[:, 0] == probes.astype(np.uint64)).all(), name

        st = rf.stats
        pruned = 1.0 - st["series_refined"] / max(1, st["series_total"])
        _say(f"sizes {name}: bf16 pairs {st['bf16_pairs']}, series refined {st['series_refined']} of "
             f"{st['series_total']} ({pruned:.6f} pruned)")
        assert pruned > 0.99

        # the modes on a sample of the queries
        m = np.ascontiguousarray(Q[sample[:64]])
        km = kq[sample[:64]]
        ex = rf.d2[sample[:64]].astype(np.float64)
        re = ix.knn(m, km, epsilon=1.0)
        rn = ix.knn(m, km, nprobes=10)
        _say(f"sizes {name}: epsilon and ng modes on 64 queries, {time.perf_counter() - t0:.1f} s")
        for i, k in enumerate(km):
            k = int(k)
            assert (re.d2[i, :k].astype(np.float64) <= 2.0 * ex[i, :k] * (1 + 1e-6)).all(), (name, i)
            assert (rn.d2[i, :k] >= rf.d2[sample[i], :k]).all(), (name, i)

        if w.get("serve"):
            _serve_burst(hkg, ix, Q, kq, rf, w["serve"], name)


def _serve_burst(hkg, ix, Q, kq, rf, nreq: int, name: str) -> None:
    """Requests arrive as a Poisson stream (about 4 batches' worth in flight)
    while batches run: each served answer equals the direct batch's and, on
    a sample, the same query asked alone."""
    rng = np.random.default_rng(21)
    rows = rng.integers(0, Q.shape[0], nreq)
    gaps = rng.exponential(1.0 / 20000.0, nreq)
    srv = hkg.Server(ix, max_batch=256, path="bf16", max_inflight=2, max_k=int(kq.max()), scratch_bytes=16 << 30)
    try:
        tickets = []
        t = time.perf_counter()
        for i in range(nreq):
            while time.perf_counter() - t < gaps[i]:
                pass
            t = time.perf_counter()
            tickets.append(srv.submit(Q[rows[i]], int(kq[rows[i]])))
        out = [srv.wait(tk, int(kq[rows[i]])) for i, tk in enumerate(tickets)]
    finally:
        srv.stop()
    lat = np.array([o[2] for o in out])
    _say(f"sizes {name}: {nreq} served, latency p50 {np.percentile(lat, 50) / 1e3:.1f} ms, "
         f"p99 {np.percentile(lat, 99) / 1e3:.1f} ms")
    for i, (ids, d2, _) in enumerate(out):
        q, k = int(rows[i]), int(kq[rows[i]])
        assert np.array_equal(ids, rf.ids[q, :k]) and np.array_equal(d2.view(np.uint32), rf.d2[q, :k].view(np.uint32)), \
            (name, "served", i, q)
    for i in range(0, nreq, nreq // 8):
        q, k = int(rows[i]), int(kq[rows[i]])
        r = ix.knn(Q[q][None, :], k, path="bf16")
        assert np.array_equal(r.ids[0], out[i][0]) and np.array_equal(r.d2[0].view(np.uint32), out[i][1].view(np.uint32))
