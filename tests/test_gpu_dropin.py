"""The drop-in on an MI355X: the reference's own main (hercules.c), linked
with hercules_amd/integration/hkg_shim.c in place of its
exact_de_knn_search_psq_lf_sims_adaptive and exact_ng_knn_search, reads the
reference-built tiny index, answers the query file through libhkg, and
prints the reference's record stream -- with the reference CLI's answers,
exact, epsilon and nprobes."""
from __future__ import annotations

import os

import numpy as np
import pytest

import golden_data as gd
import ref_runner as rr

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(rr.DROPIN), reason="oracle/_ref/hercules_gpu not built (make -C oracle dropin)")
def test_reference_main_through_the_shim(oracle, tmp_path):
    d = tmp_path / "idx"
    gd.reference_index(oracle, "tiny", str(d))
    m = gd.meta("tiny")
    Q = gd.queries(oracle, m)
    qf = str(tmp_path / "q.bin")
    rr.write_f32(qf, Q)
    a = gd.answers("tiny")
    print("drop-in: exact, k = 10", flush=True)
    d2, dist, _, out = rr.query(str(d), qf, Q.shape[0], m["L"], 10, threads=4, binary=rr.DROPIN, timeout=300)
    assert np.array_equal(d2.view(np.uint32), a["ref_sims_d2_k10"].view(np.uint32))
    # the printed records: sqrtf(d2) with six decimals
    assert np.allclose(dist, np.sqrt(a["ref_sims_d2_k10"].astype(np.float64)), rtol=1e-6, atol=1e-6)
    c = gd.mode_configs("tiny")[0]
    print(f"drop-in: {c}", flush=True)
    d2, _, _, _ = rr.query(str(d), qf, Q.shape[0], m["L"], c["k"], threads=4, binary=rr.DROPIN, timeout=300,
                           eapca_threshold=c["eapca"], sax_threshold=c["sax"],
                           extra=["--epsilon", c["epsilon"], "--approx-stop-condition", c["budget"]])
    assert np.array_equal(d2.view(np.uint32), gd.mode_answers("tiny")["m0"].view(np.uint32))
    i, c = next((i, c) for i, c in enumerate(gd.mode_configs("tiny")) if c.get("nprobes") == 3)
    print(f"drop-in: {c}", flush=True)
    d2, _, _, _ = rr.query(str(d), qf, Q.shape[0], m["L"], c["k"], threads=4, binary=rr.DROPIN, timeout=300,
                           extra=["--nprobes", c["nprobes"]])
    assert np.array_equal(d2.view(np.uint32), gd.mode_answers("tiny")[f"m{i}"].view(np.uint32))


@pytest.mark.skipif(not os.path.exists(rr.DROPIN), reason="oracle/_ref/hercules_gpu not built (make -C oracle dropin)")
def test_sibling_query_functions_through_the_shim(oracle, tmp_path):
    """--num-threads 1 dispatches to exact_de_knn_search and a run without
    --sims to exact_de_knn_search_psq_lf (hercules_file_loaders.c:239-262);
    both reach the engine through the shim and print the reference's exact
    answers.  Their epsilon searches are refused, not answered."""
    d = tmp_path / "idx"
    gd.reference_index(oracle, "tiny", str(d))
    m = gd.meta("tiny")
    Q = gd.queries(oracle, m)
    qf = str(tmp_path / "q.bin")
    rr.write_f32(qf, Q)
    a = gd.answers("tiny")
    print("drop-in: --num-threads 1, k = 1 and 10", flush=True)
    d2, _, _, _ = rr.query(str(d), qf, Q.shape[0], m["L"], 1, threads=1, binary=rr.DROPIN, timeout=300)
    assert np.array_equal(d2.view(np.uint32), a["ref_serial_d2_k1"].view(np.uint32))
    d2, _, _, _ = rr.query(str(d), qf, Q.shape[0], m["L"], 10, threads=1, binary=rr.DROPIN, timeout=300)
    assert np.array_equal(d2.view(np.uint32), a["ref_sims_d2_k10"].view(np.uint32))
    print("drop-in: no --sims, 4 threads, k = 10", flush=True)
    d2, _, _, _ = rr.query(str(d), qf, Q.shape[0], m["L"], 10, threads=4, sims=False, binary=rr.DROPIN,
                           timeout=300)
    assert np.array_equal(d2.view(np.uint32), a["ref_sims_d2_k10"].view(np.uint32))
    print("drop-in: --num-threads 1 --epsilon 1 is refused", flush=True)
    with pytest.raises(RuntimeError, match="not supported on this path"):
        rr.query(str(d), qf, Q.shape[0], m["L"], 10, threads=1, binary=rr.DROPIN, timeout=300,
                 extra=["--epsilon", 1.0])
