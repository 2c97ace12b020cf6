"""Test infrastructure: a Python restatement of the root.idx reader
(hercules_index.c:1793-1905, 2188-2340) and the flat node tables the
oracle's mode search takes."""
from __future__ import annotations

import struct

import numpy as np


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


NODE_FIELDS = 10  # ORC_NODE_FIELDS


def node_tables(parsed: dict) -> dict:
    """Pre-order node table {is_leaf, left, right, seg_off, nseg, split_from,
    split_to, split_idx, file_pos, size} (int64), split values, and every
    node's segments (seg_end int16, ind float32 x4)."""
    nodes = parsed["nodes"]
    index_of = {id(n): i for i, n in enumerate(nodes)}
    tab = np.zeros((len(nodes), NODE_FIELDS), np.int64)
    split_value = np.zeros(len(nodes), np.float32)
    seg_end, ind = [], []
    for i, n in enumerate(nodes):
        tab[i, 0] = n["is_leaf"]
        if not n["is_leaf"]:
            tab[i, 1] = index_of[id(n["left"])]
            tab[i, 2] = index_of[id(n["right"])]
            tab[i, 5], tab[i, 6], tab[i, 7] = n["split_from"], n["split_to"], n["split_idx"]
            split_value[i] = n["split_value"]
        else:
            tab[i, 1] = tab[i, 2] = -1
            tab[i, 8] = n.get("file_pos") or 0
        tab[i, 9] = n["size"]
        tab[i, 3] = len(seg_end)
        segs = n.get("seg_end", [])
        tab[i, 4] = len(segs)
        seg_end.extend(segs)
        ind.extend(n.get("ind", []))
    return {"nodes": np.ascontiguousarray(tab), "split_value": split_value,
            "seg_end": np.array(seg_end or [0], np.int16),
            "ind": np.array(ind or [[0, 0, 0, 0]], np.float32).reshape(-1)}
