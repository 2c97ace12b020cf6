"""Every GPU path against the golden vectors (the reference CLI's answers)
on the tiny, embedding and C1 (1M x 256) fixtures: the dataset generated on
the device, indexed by the device builder, queried by each path."""
from __future__ import annotations

import numpy as np
import pytest

import golden_data as gd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


@pytest.mark.parametrize("name", gd.NAMES)
def test_gpu_paths_reproduce_reference_answers(hkg, oracle, name):
    m, a = gd.meta(name), gd.answers(name)
    print(f"golden {name}: {m['N']} x {m['L']}", flush=True)
    dX = hkg.DeviceArray(m["N"] * m["L"] * 4)
    hkg.gen_series_device(m["kind"], m["seed"], 0, m["N"], m["L"], dX)
    Q = gd.queries(oracle, m)
    kmax = a["ids"].shape[1]
    paths = ["batched", "bf16", "scan"] + (["index"] if m["N"] <= 100000 else [])
    with hkg.HerculesIndex.build_device(dX, m["N"], m["L"], leaf_size=m["leaf"]) as ix:
        for path in paths:
            r = ix.knn(Q, kmax, path=path)
            assert np.array_equal(r.ids, a["ids"]), f"{name} {path}: ids"
            assert np.array_equal(r.d2.view(np.uint32), a["d2"].view(np.uint32)), f"{name} {path}: d2"
            for key in a.files:
                if key.startswith("ref_sims_d2_k"):
                    k = int(key.rsplit("k", 1)[1])
                    assert np.array_equal(r.d2[:, :k].view(np.uint32), a[key].view(np.uint32)), (name, path, key)


def test_reference_built_c1_index(hkg, oracle, tmp_path):
    """The C1 workload (1M x 256 random walks, leaf 10000, 100 queries) on an
    index the reference itself builds (oracle/_ref/hercules, --serial 1
    --sims, every host core), read by hkg_open_dir: every path reproduces
    the golden answers -- the oracle's brute force and the reference CLI's
    own d2 at k = 1, 10, 100 -- with ids through the row map."""
    import os

    import ref_runner as rr
    from test_gpu_parity import row_map

    if not rr.available():
        pytest.skip("reference binaries (oracle/_ref) not built")
    m, a = gd.meta("c1"), gd.answers("c1")
    n, L = m["N"], m["L"]
    threads = max(2, min(16, os.cpu_count() or 2))
    print(f"c1 reference index: {n} x {L}, leaf {m['leaf']}, {threads} threads", flush=True)
    X = oracle.series(m["kind"], m["seed"], 0, n, L, threads=threads)
    rr.write_f32(str(tmp_path / "data.bin"), X)
    rr.build_index(str(tmp_path / "data.bin"), n, L, m["leaf"], str(tmp_path / "idx"), threads=threads)
    (tmp_path / "data.bin").unlink()
    idx = str(tmp_path / "idx")
    ids_map = row_map(idx, X)
    del X
    Q = gd.queries(oracle, m)
    kmax = a["ids"].shape[1]
    with hkg.HerculesIndex.open(idx) as ix:
        ix.set_ids(ids_map)
        for path in ("bf16", "batched", "index", "scan"):
            print(f"c1 reference index: path {path}", flush=True)
            r = ix.knn(Q, kmax, path=path)
            assert np.array_equal(r.ids, a["ids"]), f"c1 {path}: ids"
            assert np.array_equal(r.d2.view(np.uint32), a["d2"].view(np.uint32)), f"c1 {path}: d2"
            for k in (1, 10, 100):
                assert np.array_equal(r.d2[:, :k].view(np.uint32), a[f"ref_sims_d2_k{k}"].view(np.uint32)), (path, k)
            assert np.array_equal(r.d2[:, :1].view(np.uint32), a["ref_serial_d2_k1"].view(np.uint32)), path
