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
