"""CPU: the real-data query files (tests/golden/real_queries) and the
oracle's answers for them.  The files hold 100 z-normalised queries each;
on the committed reference-built indexes the oracle's epsilon and nprobes
answers equal the reference CLI's, as they do for the synthetic queries."""
from __future__ import annotations

import os

import numpy as np
import pytest

import golden_data as gd
from root_idx import node_tables, parse_root_idx


@pytest.mark.parametrize("name", ["seismic", "deep1b", "sald"])
def test_real_query_files(name):
    Q = gd.real_queries(name)
    L = {"seismic": 256, "deep1b": 96, "sald": 128}[name]
    assert Q.shape[1] == L and Q.shape[0] % 100 == 0 and np.isfinite(Q).all()
    assert np.abs(Q.mean(1)).max() < 1e-5 and np.abs(Q.std(1) - 1).max() < 1e-5
    a = gd.real_answers(name)
    assert a["exact_d2"].shape == (Q.shape[0], 10)
    assert (np.diff(a["exact_d2"], axis=1) > 0).all()  # distinct, ascending


@pytest.mark.parametrize("name", ["seismic", "deep1b"])
def test_oracle_modes_on_real_queries(oracle, name, tmp_path):
    s = gd.real_set(name)
    raw, sims = gd.reference_index(oracle, s["index"], str(tmp_path / "idx"))
    tabs = node_tables(parse_root_idx(str(tmp_path / "idx" / "root.idx")))
    Q = gd.real_queries(name)
    a = gd.real_answers(name)
    for i, c in enumerate(s["modes"]):
        kw = ({"nprobes": c["nprobes"]} if c.get("nprobes") else
              {"epsilon": c["epsilon"], "approx_budget": c["budget"], "eapca_th": c["eapca"], "sax_th": c["sax"]})
        _, d2, _ = oracle.knn_mode(tabs, raw, sims, Q, c["k"], **kw)
        assert np.array_equal(d2.view(np.uint32), a[f"m{i}"].view(np.uint32)), (name, c)
    # epsilon 0: the exact answers
    _, d2, _ = oracle.knn_mode(tabs, raw, sims, Q, 10, epsilon=0.0, approx_budget=1)
    assert np.array_equal(d2.view(np.uint32), a["exact_d2"].view(np.uint32))
