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
