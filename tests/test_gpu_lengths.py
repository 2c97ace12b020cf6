"""Series longer than the headline's, up to the paper's 16384 (its series-length
sweep, experiments/workloads/fig8_workload.sh:86-145), on an MI355X: an index
the reference itself builds (oracle/_ref/hercules), read by hkg_open_dir,
answered by the plain per-query path, the batched path and the full scan
(one query and several per pass), each equal to the oracle's brute force
bit for bit and to the reference CLI's own exact answers; the epsilon search
equals the reference CLI's epsilon answers."""
from __future__ import annotations

import os

import numpy as np
import pytest

import ref_runner as rr
from test_gpu_parity import assert_same, row_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


@pytest.mark.parametrize("L,n,leaf", [(512, 20000, 500), (2048, 8000, 400), (4096, 6000, 300), (16384, 2000, 200)])
def test_long_series(hkg, oracle, tmp_path, L, n, leaf):
    if not rr.available():
        pytest.skip("reference binaries (oracle/_ref) not built")
    threads = max(2, min(16, os.cpu_count() or 2))
    print(f"lengths: L={L} n={n} leaf={leaf}", flush=True)
    X = oracle.series(0, 41, 0, n, L)
    Qn, _ = oracle.noisy_queries(0, 41, n, 42, np.array([0.0, 0.01, 0.05, 0.1, 0.2, 0.3] * 2, np.float32), L)
    Q = np.ascontiguousarray(np.concatenate([oracle.series(0, 43, 0, 8, L), Qn]))
    rr.write_f32(str(tmp_path / "data.bin"), X)
    rr.write_f32(str(tmp_path / "q.bin"), Q)
    idx = str(tmp_path / "idx")
    rr.build_index(str(tmp_path / "data.bin"), n, L, leaf, idx, threads=threads)
    k = 10
    ids, d2 = oracle.knn(X, Q, k)
    ref_d2, _, _, _ = rr.query(idx, str(tmp_path / "q.bin"), Q.shape[0], L, k, threads=threads, sims=True)
    assert np.array_equal(ref_d2.view(np.uint32), d2.view(np.uint32))  # the reference CLI is exact here
    eps_d2, _, _, _ = rr.query(idx, str(tmp_path / "q.bin"), Q.shape[0], L, k, threads=threads, sims=True,
                               extra=["--epsilon", 1.0, "--approx-stop-condition", 80])
    with hkg.HerculesIndex.open(idx) as ix:
        ix.set_ids(row_map(idx, X))
        for path, batch, filt in (("index", 0, "auto"), ("batched", 0, "items"), ("batched", 0, "leaf"),
                                  ("scan", 1, "auto"), ("scan", 0, "auto")):
            print(f"lengths: L={L} path={path} batch={batch} filter={filt}", flush=True)
            r = ix.knn(Q, k, path=path, batch=batch, batched_filter=filt)
            assert_same(r.ids, r.d2, ids, d2, f"L={L} {path} batch {batch} filter {filt}")
        kv = np.array([1, 3, 10, 7] * (Q.shape[0] // 4), np.uint32)
        for filt in ("items", "leaf"):
            r = ix.knn(Q, kv, path="batched", batched_filter=filt)
            for i, kq in enumerate(kv):
                assert np.array_equal(r.d2[i, :kq].view(np.uint32), d2[i, :kq].view(np.uint32)), (L, filt, i)
        print(f"lengths: L={L} epsilon 1", flush=True)
        e = ix.knn(Q, k, epsilon=1.0, approx_leaves=80)
        assert np.array_equal(e.d2.view(np.uint32), eps_d2.view(np.uint32))
