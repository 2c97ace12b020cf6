"""GPU: the bulk loader builds a valid EAPCA index (same on-disk format) --
the reference CLI answers exactly on it, and so do the GPU paths."""
from __future__ import annotations

import os

import numpy as np
import pytest

import ref_runner as rr

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def hkg():
    from hercules_amd import hkg as m
    m.lib()
    return m


@pytest.fixture(scope="module")
def data(oracle):
    X = oracle.series(0, 31, 0, 30000, 256)
    Q = np.concatenate([oracle.series(0, 32, 0, 20, 256),
                        oracle.noisy_queries(0, 31, 30000, 33, np.array([0.01, 0.1, 0.2] * 6, np.float32),
                                             256)[0]])
    return X, np.ascontiguousarray(Q)


@pytest.fixture(scope="module")
def built(hkg, data, tmp_path_factory):
    X, _ = data
    _say("build 30000 x 256 leaf 800")
    ix = hkg.HerculesIndex.build(X, leaf_size=800)
    d = tmp_path_factory.mktemp("built")
    ix.write(str(d))
    yield ix, str(d)
    ix.close()


def test_build_structure(hkg, oracle, data, built):
    X, _ = data
    ix, d = built
    info = hkg.inspect_dir(d)
    assert info["n_series"] == X.shape[0]
    assert info["n_leaves"] >= X.shape[0] // 800
    ids = np.fromfile(os.path.join(d, "ids.u64"), np.uint64)
    assert np.array_equal(np.sort(ids), np.arange(X.shape[0], dtype=np.uint64))
    raw = np.fromfile(os.path.join(d, "leaves_raw.idx"), np.float32).reshape(-1, 256)
    assert np.array_equal(raw, X[ids.astype(np.int64)])
    sims = np.fromfile(os.path.join(d, "leaves_sims.idx"), np.uint8).reshape(-1, 16)
    for r in range(0, X.shape[0], 97):
        assert np.array_equal(sims[r], oracle.sax(oracle.paa(raw[r])))


def test_build_leaf_sketches_contain_members(hkg, oracle, data, built):
    """Parse root.idx in Python (independently of the C++ parser) and check
    every leaf's intervals contain its members' FP32 segment statistics."""
    from test_index_format import parse_root_idx
    X, _ = data
    _, d = built
    tree = parse_root_idx(os.path.join(d, "root.idx"))
    raw = np.fromfile(os.path.join(d, "leaves_raw.idx"), np.float32).reshape(-1, 256)
    for leaf in tree["leaves"][::3]:
        s = 0
        for e, ind in zip(leaf["seg_end"], leaf["ind"]):
            for r in range(leaf["file_pos"], leaf["file_pos"] + leaf["size"], 7):
                # the reference's FP32 statistics and the exact ones both lie inside
                mu, sd = oracle.seg_mean_std(raw[r], s, int(e))
                assert ind[1] <= mu <= ind[0]
                assert ind[3] <= sd <= ind[2]
                seg = raw[r, s:int(e)].astype(np.float64)
                assert ind[1] - 1e-12 <= seg.mean() <= ind[0] + 1e-12
                assert ind[3] - 1e-12 <= seg.std() <= ind[2] + 1e-12
            s = int(e)


def test_reference_cli_answers_on_built_index(hkg, oracle, data, built, tmp_path):
    if not rr.available():
        pytest.skip("reference binaries not built")
    X, Q = data
    _, d = built
    _say("reference cli on built index")
    rr.write_f32(str(tmp_path / "q.bin"), Q)
    ids, d2 = oracle.knn(X, Q, 10)
    for sims, th in [(True, 8), (False, 1)]:
        rd2, _, _, _ = rr.query(d, str(tmp_path / "q.bin"), Q.shape[0], 256, 10, threads=th, sims=sims)
        assert np.array_equal(rd2.view(np.uint32), d2.view(np.uint32)), sims


@pytest.mark.parametrize("path", ["index", "batched"])
def test_gpu_paths_on_built_index(hkg, oracle, data, built, path):
    X, Q = data
    ix, _ = built
    _say(f"gpu {path} on built index")
    ids, d2 = oracle.knn(X, Q, 10)
    r = ix.knn(Q, 10, path=path)
    assert np.array_equal(r.d2.view(np.uint32), d2.view(np.uint32))
    assert np.array_equal(r.ids, ids)


def test_build_device_from_generator(hkg, oracle):
    """hkg_build_device on rows generated in HBM equals the host path."""
    _say("build_device from generator")
    n, L = 20000, 96
    dX = hkg.DeviceArray(n * L * 4)
    hkg.gen_series_device(1, 41, 0, n, L, dX)
    ix = hkg.HerculesIndex.build_device(dX, n, L, leaf_size=500)
    X = oracle.series(1, 41, 0, n, L)
    Q = np.ascontiguousarray(X[[3, 500, 19999]] + np.float32(0.05))
    ids, d2 = oracle.knn(X, Q, 5)
    r = ix.knn(Q, 5, path="batched")
    ix.close()
    assert np.array_equal(r.d2.view(np.uint32), d2.view(np.uint32))
    assert np.array_equal(r.ids, ids)
