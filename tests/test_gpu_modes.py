"""The epsilon and no-guarantee (nprobes) modes on an MI355X.

On the reference's own indexes (committed root.idx, regenerated leaf files)
the device answers equal the reference CLI's answers bit for bit, and the
traversal counts equal the oracle's; on an index the device builder made,
the answers equal the oracle run over the same tree and rows."""
from __future__ import annotations

import numpy as np
import pytest

import golden_data as gd
from root_idx import node_tables, parse_root_idx

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


def _mode_kw(c: dict) -> dict:
    if c.get("nprobes"):
        return {"nprobes": c["nprobes"]}
    return {"epsilon": c["epsilon"], "approx_leaves": c["budget"], "eapca_threshold": c["eapca"],
            "sax_threshold": c["sax"]}


def _oracle_kw(c: dict) -> dict:
    if c.get("nprobes"):
        return {"nprobes": c["nprobes"]}
    return {"epsilon": c["epsilon"], "approx_budget": c["budget"], "eapca_th": c["eapca"], "sax_th": c["sax"]}


def _check_stats(st, info, what):
    assert st["mode_leaves_first"] == int(info[:, 0].sum()), what
    assert st["mode_candidates"] == int(info[:, 1].sum()), what
    assert st["mode_candidates_scanned"] == int(info[:, 3].sum()), what
    assert st["mode_sequential"] == int((info[:, 2] == 1).sum()), what


@pytest.mark.parametrize("name", ["tiny", "emb"])
def test_modes_reproduce_reference_cli(hkg, oracle, name, tmp_path):
    d = tmp_path / "idx"
    raw, sims = gd.reference_index(oracle, name, str(d))
    tabs = node_tables(parse_root_idx(str(d / "root.idx")))
    Q = gd.queries(oracle, gd.meta(name))
    ans = gd.mode_answers(name)
    with hkg.HerculesIndex.open(str(d)) as ix:
        for i, c in enumerate(gd.mode_configs(name)):
            _say(f"modes {name} {c}")
            r = ix.knn(Q, c["k"], **_mode_kw(c))
            assert np.array_equal(r.d2.view(np.uint32), ans[f"m{i}"].view(np.uint32)), (name, c)
            ids, d2, info = oracle.knn_mode(tabs, raw, sims, Q, c["k"], **_oracle_kw(c))
            assert np.array_equal(r.ids, ids), (name, c)
            _check_stats(r.stats, info, (name, c))
        # a candidate buffer of k entries: queries that overflow it are
        # answered again, with the same answers
        for i, c in enumerate(gd.mode_configs(name)):
            _say(f"modes {name} {c} with a {c['k']}-entry candidate buffer")
            r = ix.knn(Q, c["k"], cand_cap=c["k"], **_mode_kw(c))
            assert np.array_equal(r.d2.view(np.uint32), ans[f"m{i}"].view(np.uint32)), (name, c)
            if name == "tiny" and i == 0:
                assert r.stats["overflow_retries"] > 0


def test_modes_with_per_query_k(hkg, oracle, tmp_path):
    """k differs per query: each query's answer is the oracle's for its k."""
    d = tmp_path / "idx"
    raw, sims = gd.reference_index(oracle, "tiny", str(d))
    tabs = node_tables(parse_root_idx(str(d / "root.idx")))
    Q = gd.queries(oracle, gd.meta("tiny"))
    kv = np.array([1, 3, 10, 64, 100, 7][: 6] * (Q.shape[0] // 6), np.uint32)
    with hkg.HerculesIndex.open(str(d)) as ix:
        for kw, okw in [({"epsilon": 3.0, "approx_leaves": 1}, {"epsilon": 3.0, "approx_budget": 1}),
                        ({"epsilon": 0.5, "approx_leaves": 2, "eapca_threshold": 0.99},
                         {"epsilon": 0.5, "approx_budget": 2, "eapca_th": 0.99}),
                        ({"nprobes": 4}, {"nprobes": 4})]:
            print(f"modes per-query k {kw}", flush=True)
            r = ix.knn(Q, kv, **kw)
            for k in np.unique(kv):
                sel = np.nonzero(kv == k)[0]
                ids, d2, _ = oracle.knn_mode(tabs, raw, sims, np.ascontiguousarray(Q[sel]), int(k), **okw)
                assert np.array_equal(r.d2[sel, :k].view(np.uint32), d2.view(np.uint32)), (kw, k)
                assert np.array_equal(r.ids[sel, :k], ids), (kw, k)
                if k < r.d2.shape[1]:
                    assert (r.ids[sel, k:] == np.uint64(2**64 - 1)).all()


def test_modes_on_device_built_index(hkg, oracle, tmp_path):
    n, L = 60000, 256
    X = oracle.series(0, 41, 0, n, L)
    sig = np.array([0.0, 0.02, 0.1, 0.3] * 6, np.float32)
    Qn, _ = oracle.noisy_queries(0, 41, n, 42, sig, L)
    Q = np.ascontiguousarray(np.concatenate([Qn, oracle.series(0, 43, 0, 8, L)]))
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    d = tmp_path / "built"
    # the built index answers with the dataset's ids; the same tree written
    # out and opened again answers with leaf-file rows, the oracle's ids
    with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=700) as built:
        built.write(str(d))
        tabs = node_tables(parse_root_idx(str(d / "root.idx")))
        raw = np.fromfile(d / "leaves_raw.idx", np.float32).reshape(n, L)
        sims = np.fromfile(d / "leaves_sims.idx", np.uint8).reshape(n, 16)
        ix = hkg.HerculesIndex.open(str(d))
        exact = ix.knn(Q, 20, path="batched")
        cfgs = [dict(k=20, epsilon=0.0, budget=1, eapca=0.25, sax=0.9),
                dict(k=20, epsilon=2.0, budget=1, eapca=0.25, sax=0.9),
                dict(k=20, epsilon=8.0, budget=4, eapca=0.9, sax=0.9),
                dict(k=5, epsilon=0.7, budget=2, eapca=0.25, sax=0.9999),
                dict(k=64, epsilon=1.5, budget=80, eapca=0.25, sax=0.9),
                dict(k=20, nprobes=1), dict(k=7, nprobes=6), dict(k=64, nprobes=40)]
        for c in cfgs:
            _say(f"modes built {c}")
            k = c["k"]
            r = ix.knn(Q, k, **_mode_kw(c)) if (c.get("nprobes") or c["epsilon"]) else \
                ix.knn(Q, k, epsilon=1e-30, approx_leaves=c["budget"])
            ids, d2, info = oracle.knn_mode(tabs, raw, sims, Q, k, **_oracle_kw(c))
            assert np.array_equal(r.d2.view(np.uint32), d2.view(np.uint32)), c
            assert np.array_equal(r.ids, ids), c
            _check_stats(r.stats, info, c)
            rb = built.knn(Q, k, **_mode_kw(c)) if (c.get("nprobes") or c["epsilon"]) else \
                built.knn(Q, k, epsilon=1e-30, approx_leaves=c["budget"])
            assert np.array_equal(rb.d2.view(np.uint32), d2.view(np.uint32)), c
            kk = min(k, 20)
            if not c.get("nprobes"):
                bound = (1.0 + max(c["epsilon"], 1e-30)) * exact.d2[:, :kk].astype(np.float64) * (1 + 1e-6)
                assert (r.d2[:, :kk] <= bound).all(), c
            else:
                assert (r.d2[:, :kk] >= exact.d2[:, :kk]).all(), c
        # epsilon 0 and no probes is the exact search of any path
        r0 = ix.knn(Q, 20, epsilon=0.0, nprobes=0)
        assert np.array_equal(r0.d2.view(np.uint32), exact.d2.view(np.uint32))
        ix.close()


def test_modes_edge_settings(hkg, oracle, tmp_path):
    """Short series (L = 64), k = 256, more probes and a larger budget than
    the tree has leaves, a tiny epsilon, duplicated rows (ties)."""
    n, L = 6000, 64
    X = oracle.series(1, 51, 0, n, L)
    X[100:110] = X[5]  # ten copies of one row: equal distances, smallest id wins
    Q = np.ascontiguousarray(np.concatenate([X[5:6], oracle.series(1, 52, 0, 11, L)]))
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    d = tmp_path / "edge"
    with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=150) as built:
        built.write(str(d))
    tabs = node_tables(parse_root_idx(str(d / "root.idx")))
    raw = np.fromfile(d / "leaves_raw.idx", np.float32).reshape(n, L)
    sims = np.fromfile(d / "leaves_sims.idx", np.uint8).reshape(n, 16)
    nleaf = int((tabs["nodes"][:, 0] == 1).sum())
    with hkg.HerculesIndex.open(str(d)) as ix:
        for c in [dict(k=256, nprobes=nleaf + 50), dict(k=256, epsilon=1e-6, budget=100000, eapca=0.25, sax=0.9),
                  dict(k=1, epsilon=0.3, budget=0, eapca=0.25, sax=0.9), dict(k=33, nprobes=1)]:
            print(f"modes edge {c}", flush=True)
            r = ix.knn(Q, c["k"], **_mode_kw(c))
            ids, d2, info = oracle.knn_mode(tabs, raw, sims, Q, c["k"], **_oracle_kw(c))
            assert np.array_equal(r.d2.view(np.uint32), d2.view(np.uint32)), c
            assert np.array_equal(r.ids, ids), c
            _check_stats(r.stats, info, c)


def test_modes_reject_bad_settings(hkg, oracle):
    X = oracle.series(0, 5, 0, 2000, 64)
    with hkg.HerculesIndex.open_raw(X) as ix:
        with pytest.raises(hkg.HkgError):
            ix.knn(X[:2], 3, epsilon=1.0)  # no tree
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    with hkg.HerculesIndex.build_device(dX, 2000, 64, leaf_size=200) as ix:
        with pytest.raises(hkg.HkgError):
            ix.knn(X[:2], 3, epsilon=-1.0)
        with pytest.raises(hkg.HkgError):
            ix.knn(X[:2], 3, epsilon=float("inf"))
