"""The oracle against the golden vectors: the reference CLI's own exact
squared distances (SIMS and serial query paths, printed in hex), its PSCAN
distances, and the brute-force answers, for the tiny, embedding and C1
(1M x 256, 100 queries) fixtures."""
from __future__ import annotations

import numpy as np
import pytest

import golden_data as gd


@pytest.mark.parametrize("name", gd.NAMES)
def test_oracle_reproduces_reference_answers(oracle, name):
    m, a = gd.meta(name), gd.answers(name)
    X = oracle.series(m["kind"], m["seed"], 0, m["N"], m["L"])
    Q = gd.queries(oracle, m)
    kmax = a["ids"].shape[1]
    ids, d2 = oracle.knn(X, Q, kmax)
    assert np.array_equal(ids, a["ids"]) and np.array_equal(d2.view(np.uint32), a["d2"].view(np.uint32))
    # the reference's own query paths, distance per rank bit for bit
    for key in a.files:
        if key.startswith("ref_sims_d2_k") or key.startswith("ref_serial_d2_k"):
            k = int(key.rsplit("k", 1)[1])
            assert np.array_equal(d2[:, :k].view(np.uint32), a[key].view(np.uint32)), key
    # PSCAN prints sqrt(d2) with 6 decimals
    pd = a["ref_pscan_dist_k1"][:, 0]
    assert np.allclose(np.sqrt(d2[:, 0].astype(np.float64)), pd, rtol=0, atol=1.5e-6)
    # distances within 1e-5 relative of the FP64 distance of the same pair
    assert float(np.max(a["rel_err_fp64"])) < 1e-5
