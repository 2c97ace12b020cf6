"""CPU: the on-disk index format.

A Python restatement of the root.idx reader (hercules_index.c:1793-1905,
2188-2340) checks the committed indexes the reference built
(tests/golden/*_root.idx) and the C++ parser behind hkg_inspect_dir; the
leaf files are regenerated from the generator and the committed row maps, so
the reference's iSAX words and leaf sketches are pinned against the oracle.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from root_idx import parse_root_idx

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["tiny", "emb"])
def test_golden_index_parses_and_tiles(name):
    t = parse_root_idx(os.path.join(GOLDEN, f"{name}_root.idx"))
    meta = json.load(open(os.path.join(GOLDEN, f"{name}_meta.json")))
    assert t["header"]["L"] == meta["L"] and t["header"]["max_leaf"] == meta["leaf"]
    pos = 0
    for lf in t["leaves"]:
        assert lf["file_pos"] == pos
        assert 0 < lf["size"] <= meta["leaf"]
        assert lf["seg_end"][-1] == meta["L"] and list(lf["seg_end"]) == sorted(set(lf["seg_end"]))
        pos += lf["size"]
    assert pos == meta["N"]
    # internal sizes are sums of their children
    def check(nd):
        if nd["is_leaf"]:
            return nd["size"]
        s = check(nd["left"]) + check(nd["right"])
        assert nd["size"] == s
        return s
    assert check(t["root"]) == meta["N"]


@pytest.mark.parametrize("name", ["tiny", "emb"])
def test_reference_leaf_files_pinned_by_oracle(oracle, name, tmp_path):
    """leaves_raw.idx = the generated rows in the reference's leaf order;
    leaves_sims.idx = the oracle's iSAX words of those rows (hash equal to
    the file the reference wrote); leaf sketches contain member statistics."""
    meta = json.load(open(os.path.join(GOLDEN, f"{name}_meta.json")))
    rowmap = np.load(os.path.join(GOLDEN, f"{name}_rowmap.npy"))
    X = oracle.series(meta["kind"], meta["seed"], 0, meta["N"], meta["L"])
    raw = X[rowmap.astype(np.int64)]
    sims = np.stack([oracle.sax(oracle.paa(r)) for r in raw])
    assert hashlib.sha1(sims.tobytes()).hexdigest() == meta["leaves_sims_sha1"]
    t = parse_root_idx(os.path.join(GOLDEN, f"{name}_root.idx"))
    for lf in t["leaves"][::5]:
        s = 0
        for e, ind in zip(lf["seg_end"], lf["ind"]):
            for r in range(lf["file_pos"], lf["file_pos"] + lf["size"], 3):
                mu, sd = oracle.seg_mean_std(raw[r], s, int(e))
                assert ind[1] <= mu <= ind[0]
                assert ind[3] <= sd <= ind[2]
            s = int(e)
    # the C++ parser (host only, no device) agrees
    from hercules_amd import hkg
    d = tmp_path / "idx"
    d.mkdir()
    (d / "root.idx").write_bytes(open(os.path.join(GOLDEN, f"{name}_root.idx"), "rb").read())
    raw.tofile(d / "leaves_raw.idx")
    sims.tofile(d / "leaves_sims.idx")
    info = hkg.inspect_dir(str(d))
    assert info["n_series"] == meta["N"] and info["n_leaves"] == len(t["leaves"])
    assert info["n_nodes"] == len(t["nodes"])
    # a truncated leaf file is rejected
    (d / "leaves_sims.idx").write_bytes(sims.tobytes()[:-16])
    with pytest.raises(hkg.HkgError) as e:
        hkg.inspect_dir(str(d))
    assert e.value.status == 4


def test_truncated_root_idx_rejected(tmp_path):
    from hercules_amd import hkg
    full = open(os.path.join(GOLDEN, "tiny_root.idx"), "rb").read()
    for cut in [10, 44, 60, len(full) // 2, len(full) - 1]:
        d = tmp_path / f"c{cut}"
        d.mkdir()
        (d / "root.idx").write_bytes(full[:cut])
        with pytest.raises(hkg.HkgError) as e:
            hkg.inspect_dir(str(d))
        assert e.value.status == 4, cut
