"""Real-data query files on an MI355X, against the synthetic datasets of
their length: every exact path returns the reference CLI's answers (ids:
the oracle's brute force), and the epsilon / nprobes modes return the
reference's mode answers on its own indexes."""
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


def _check_exact(ix, Q, a, what, paths=("batched", "bf16", "scan", "index")):
    for path in paths:
        print(f"real {what}: {path}", flush=True)
        r = ix.knn(Q, 10, path=path)
        assert np.array_equal(r.d2.view(np.uint32), a["exact_d2"].view(np.uint32)), (what, path)
        assert np.array_equal(r.ids, a["exact_ids"]), (what, path)


@pytest.mark.parametrize("name", ["seismic", "deep1b"])
def test_real_queries_on_reference_indexes(hkg, oracle, name, tmp_path):
    s = gd.real_set(name)
    d = tmp_path / "idx"
    gd.reference_index(oracle, s["index"], str(d))
    rowmap = np.load(f"{gd.GOLDEN}/{s['index']}_rowmap.npy").astype(np.uint64)
    Q = gd.real_queries(name)
    a = gd.real_answers(name)
    with hkg.HerculesIndex.open(str(d)) as ix:
        for i, c in enumerate(s["modes"]):
            kw = ({"nprobes": c["nprobes"]} if c.get("nprobes") else
                  {"epsilon": c["epsilon"], "approx_leaves": c["budget"], "eapca_threshold": c["eapca"],
                   "sax_threshold": c["sax"]})
            print(f"real {name}: mode {c}", flush=True)
            r = ix.knn(Q, c["k"], **kw)
            assert np.array_equal(r.d2.view(np.uint32), a[f"m{i}"].view(np.uint32)), (name, c)
        ix.set_ids(rowmap)  # leaf rows -> dataset ordinals
        _check_exact(ix, Q, a, name)


def test_real_queries_on_device_built_index(hkg):
    s = gd.real_set("sald")
    Q = gd.real_queries("sald")
    a = gd.real_answers("sald")
    dX = hkg.DeviceArray(s["N"] * s["L"] * 4)
    hkg.gen_series_device(s["kind"], s["seed"], 0, s["N"], s["L"], dX)
    with hkg.HerculesIndex.build_device(dX, s["N"], s["L"], leaf_size=s["leaf"]) as ix:
        _check_exact(ix, Q, a, "sald")
