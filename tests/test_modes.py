"""CPU: the oracle's epsilon and no-guarantee (nprobes) search modes against
the reference CLI's answers on its own indexes (tests/golden/*_modes.*),
and the properties the modes must keep: epsilon = 0 is the exact search,
every epsilon answer is within (1 + epsilon) of the exact one rank by rank
(squared distances), and the ng answers are never below the exact ones."""
from __future__ import annotations

import numpy as np
import pytest

import golden_data as gd
from root_idx import node_tables, parse_root_idx


@pytest.fixture(scope="module", params=["tiny", "emb"])
def ref_index(request, oracle, tmp_path_factory):
    name = request.param
    d = tmp_path_factory.mktemp(f"{name}_idx")
    raw, sims = gd.reference_index(oracle, name, str(d))
    tabs = node_tables(parse_root_idx(str(d / "root.idx")))
    Q = gd.queries(oracle, gd.meta(name))
    return name, tabs, raw, sims, Q


def _run(oracle, ref_index, c):
    name, tabs, raw, sims, Q = ref_index
    if c.get("nprobes"):
        return oracle.knn_mode(tabs, raw, sims, Q, c["k"], nprobes=c["nprobes"])
    return oracle.knn_mode(tabs, raw, sims, Q, c["k"], epsilon=c["epsilon"], approx_budget=c["budget"],
                           eapca_th=c["eapca"], sax_th=c["sax"])


def test_oracle_modes_equal_reference_cli(oracle, ref_index):
    name = ref_index[0]
    ans = gd.mode_answers(name)
    branches = set()
    for i, c in enumerate(gd.mode_configs(name)):
        ids, d2, info = _run(oracle, ref_index, c)
        assert np.array_equal(d2.view(np.uint32), ans[f"m{i}"].view(np.uint32)), (name, c)
        if c.get("nprobes"):
            assert (info[:, 0] <= c["nprobes"]).all()
        else:
            assert (info[:, 0] <= max(1, c["budget"])).all()
            branches |= set(np.unique(info[:, 2]).tolist())
    if name == "tiny":
        assert branches == {0, 1, 2}, branches  # every candidate branch is exercised


def test_mode_properties(oracle, ref_index):
    name, tabs, raw, sims, Q = ref_index
    exact = gd.answers(name)["d2"]
    kmax = exact.shape[1]
    k = min(10, kmax)
    # epsilon = 0 is the exact search, whatever the budget and thresholds
    for budget, eth, sth in [(1, 0.25, 0.9), (2, 0.99, 0.9), (1, 0.5, 0.999)]:
        _, d2, _ = oracle.knn_mode(tabs, raw, sims, Q, k, epsilon=0.0, approx_budget=budget, eapca_th=eth, sax_th=sth)
        assert np.array_equal(d2.view(np.uint32), exact[:, :k].view(np.uint32)), (budget, eth, sth)
    # the (1 + epsilon) guarantee on every rank, and the ids hold their distances
    for eps in [0.5, 3.0, 50.0]:
        ids, d2, _ = oracle.knn_mode(tabs, raw, sims, Q, k, epsilon=eps, approx_budget=1)
        assert (d2.astype(np.float64) <= (1.0 + eps) * exact[:, :k].astype(np.float64) * (1 + 1e-6)).all(), eps
        for qi in range(0, Q.shape[0], 7):
            for r in range(k):
                assert oracle.ed2(Q[qi], raw[int(ids[qi, r])]) == d2[qi, r]
    # ng: never better than exact; enough probes reach the exact answer on
    # queries that are copies of indexed series (rank 1)
    _, d2, info = oracle.knn_mode(tabs, raw, sims, Q, k, nprobes=3)
    assert (d2 >= exact[:, :k]).all()
    nleaf = int((tabs["nodes"][:, 0] == 1).sum())
    _, d2all, _ = oracle.knn_mode(tabs, raw, sims, Q, 1, nprobes=nleaf + 1)
    assert (d2all[:, 0] >= exact[:, 0]).all()
