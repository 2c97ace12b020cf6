"""Inputs of the golden fixtures, regenerated from the seeds in their
meta files (the datasets themselves are not stored)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["tiny", "emb", "c1"]


def meta(name: str) -> dict:
    with open(os.path.join(GOLDEN, f"{name}_meta.json")) as f:
        return json.load(f)


def answers(name: str):
    return np.load(os.path.join(GOLDEN, f"{name}_answers.npz"))


def queries(gen, m: dict) -> np.ndarray:
    """gen: an object with series(kind, seed, first, n, L) and
    noisy_queries(kind, data_seed, N, query_seed, sigmas, L) (the oracle)."""
    Q = gen.series(m["kind"], m["qseed"], 0, m["nq"], m["L"])
    if m["nseed"] is not None and m["sigmas"]:
        Qn, _ = gen.noisy_queries(m["kind"], m["seed"], m["N"], m["nseed"], np.array(m["sigmas"], np.float32),
                                  m["L"])
        Q = np.concatenate([Q, Qn])
    return np.ascontiguousarray(Q)


def mode_configs(name: str) -> list:
    with open(os.path.join(GOLDEN, f"{name}_modes.json")) as f:
        return json.load(f)


def mode_answers(name: str):
    return np.load(os.path.join(GOLDEN, f"{name}_modes.npz"))


def reference_index(gen, name: str, out_dir: str):
    """Write the reference-built index of a fixture (its committed root.idx,
    the generated rows in its leaf order, the oracle's iSAX words of those
    rows) into out_dir; returns (rows in leaf order, iSAX words)."""
    m = meta(name)
    rowmap = np.load(os.path.join(GOLDEN, f"{name}_rowmap.npy")).astype(np.int64)
    X = gen.series(m["kind"], m["seed"], 0, m["N"], m["L"])
    raw = np.ascontiguousarray(X[rowmap])
    sims = np.stack([gen.sax(gen.paa(r)) for r in raw])
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(GOLDEN, f"{name}_root.idx"), "rb") as f:
        root = f.read()
    with open(os.path.join(out_dir, "root.idx"), "wb") as f:
        f.write(root)
    raw.tofile(os.path.join(out_dir, "leaves_raw.idx"))
    sims.tofile(os.path.join(out_dir, "leaves_sims.idx"))
    return raw, sims


REAL_DIR = os.path.join(GOLDEN, "real_queries")


def real_set(name: str) -> dict:
    with open(os.path.join(GOLDEN, f"real_{name}.json")) as f:
        return json.load(f)


def real_answers(name: str):
    return np.load(os.path.join(GOLDEN, f"real_{name}.npz"))


def real_queries(name: str) -> np.ndarray:
    """The real query files of a set, concatenated (tests/golden/real_queries)."""
    files = {"seismic": ["queries_hard1p_seismic_len256_znorm.bin", "queries_hard10p_seismic_len256_znorm.bin",
                         "queries_size100_seismic_len256_znorm.bin"],
             "deep1b": ["queries_hard1p_deep1b_len96_znorm.bin", "queries_orig100_deep1b_len96_znorm.bin"],
             "sald": ["queries_hard1p_sald_len128_znorm.bin"]}[name]
    out = []
    for f in files:
        L = int(f.split("_len")[1].split("_")[0])
        out.append(np.fromfile(os.path.join(REAL_DIR, f), np.float32).reshape(-1, L))
    return np.ascontiguousarray(np.concatenate(out))
