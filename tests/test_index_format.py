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
import struct

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def parse_root_idx(path: str) -> dict:
    buf = open(path, "rb").read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, buf, off)
        off += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    hdr = dict(zip(["leaf_count", "buffer_mb", "L", "init_segments", "paa_segments", "sax_bits", "max_leaf",
                    "serial", "sims"], take("QdIIiiIii")))
    leaves, nodes = [], []

    def sketches():
        S = take("h")
        seg_end = list(take(f"{S}h")) if S > 1 else [take("h")]
        ind = []
        for _ in range(S):
            n = take("i")
            assert n == 4
            ind.append(list(take("4f")))
        return seg_end, ind

    def node(depth):
        is_leaf, level = take("BI")
        me = {"is_leaf": is_leaf, "level": level}
        nodes.append(me)
        if not is_leaf:
            me["left"] = node(depth + 1)
            me["right"] = node(depth + 1)
            me["size"] = take("I")
            me["split_segment"] = take("h")
            pol = buf[off:off + 48]
            off_ = off + 48
            me["split_from"], me["split_to"] = struct.unpack_from("<hh", pol, 0)
            me["split_idx"], me["split_value"] = struct.unpack_from("<if", pol, 12)
            globals_off(off_)
            me["seg_end"], me["ind"] = sketches()
            return me
        me["size"] = take("I")
        flen = take("i")
        if flen == 0:
            me["size"] = 0
            return me
        me["name"] = buf[off:off + flen].decode(errors="replace")
        globals_off(off + flen)
        me["seg_end"], me["ind"] = sketches()
        me["file_pos"] = take("Q") if hdr["serial"] else None
        leaves.append(me)
        return me

    def globals_off(v):
        nonlocal off
        off = v

    root = node(0)
    assert off == len(buf), "trailing bytes"
    return {"header": hdr, "root": root, "leaves": leaves, "nodes": nodes}


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
