"""Drive the reference CLI (oracle/_ref/hercules, oracle/_ref/pscan) -- test
infrastructure only.

Only options present in the reference's option table are passed
(code/hercules/src/hercules.c:81-116): that table has no terminator, so an
unknown option makes getopt_long read past its end.  Index paths are
absolute with a trailing '/', because hercules_index_read chdir()s into the
index directory (hercules_index.c:1795).
"""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np

from oracle_bind import REF_DIR

HERCULES = os.path.join(REF_DIR, "hercules")
PSCAN = os.path.join(REF_DIR, "pscan")


def available() -> bool:
    return os.path.exists(HERCULES) and os.path.exists(PSCAN)


def write_f32(path: str, X: np.ndarray) -> None:
    np.ascontiguousarray(X, np.float32).tofile(path)


def build_index(data_path: str, n: int, L: int, leaf: int, index_dir: str, threads: int = 8,
                timeout: float = 600) -> None:
    """Parallel build (the serial build crashes) with --serial 1 --sims."""
    index_dir = os.path.abspath(index_dir) + "/"
    threads = max(2, threads)
    cmd = [HERCULES, "--dataset", os.path.abspath(data_path), "--dataset-size", str(n),
           "--timeseries-size", str(L), "--leaf-size", str(leaf), "--buffer-size", "2000",
           "--index-path", index_dir, "--serial", "1", "--sims", "--mode", "0",
           "--num-threads", str(threads), "--num-write-threads", str(max(1, threads // 2)),
           "--flush-threshold", str(max(1, threads // 2)), "--initial-db-size", str(min(n, 120000))]
    r = subprocess.run(cmd, capture_output=True, encoding="utf-8", errors="replace", timeout=timeout)
    if r.returncode != 0 or not os.path.exists(index_dir + "root.idx"):
        raise RuntimeError(f"reference build failed rc={r.returncode}: {r.stderr[-2000:]}")


_D2 = re.compile(r"^Query_exact_d2\t(\S+)\t(\d+)\t(\d+)$", re.M)
_DIST = re.compile(r"^Query_exact_distance\t(\S+)\t\S+\t(\d+)\t(\d+)$", re.M)


DROPIN = os.path.join(REF_DIR, "hercules_gpu")  # the reference's main with the engine's shim


def query(index_dir: str, queries_path: str, nq: int, L: int, k: int, threads: int = 8, sims: bool = True,
          sax_threshold: float = 0.9, eapca_threshold: float = 0.25, timeout: float = 900, extra=(),
          binary: str = HERCULES):
    """Run the reference query path; returns (d2 (nq,k) float32 from the
    exact hex print, printed sqrt distances (nq,k), wall seconds, stdout).
    sims=False with threads=1 runs the serial path exact_de_knn_search.
    extra: further options of the table (--epsilon, --nprobes,
    --approx-stop-condition)."""
    index_dir = os.path.abspath(index_dir) + "/"
    cmd = [binary, "--mode", "1", "--index-path", index_dir, "--queries", os.path.abspath(queries_path),
           "--queries-size", str(nq), "--timeseries-size", str(L), "--k", str(k), "--serial", "1",
           "--num-threads", str(threads), "--num-read-threads", "1",
           "--eapca-threshold", str(eapca_threshold), "--sax-threshold", str(sax_threshold)]
    if sims:
        cmd.append("--sims")
    cmd += [str(a) for a in extra]
    import time
    t0 = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, encoding="utf-8", errors="replace", timeout=timeout)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"reference query failed rc={r.returncode}: {r.stderr[-2000:]}")
    d2 = np.full((nq, k), np.nan, np.float32)
    dist = np.full((nq, k), np.nan, np.float64)
    for m in _D2.finditer(r.stdout):
        qi, rank = int(m.group(2)), int(m.group(3))
        if 1 <= qi <= nq and 1 <= rank <= k:
            d2[qi - 1, rank - 1] = np.float32(float.fromhex(m.group(1)))
    for m in _DIST.finditer(r.stdout):
        qi, rank = int(m.group(2)), int(m.group(3))
        if 1 <= qi <= nq and 1 <= rank <= k:
            dist[qi - 1, rank - 1] = float(m.group(1))
    return d2, dist, wall, r.stdout


_PDIST = re.compile(r"^Query_exact_distance\t(\S+)\t\S+\t(\d+)\t(\d+)$", re.M)


def pscan(data_path: str, n: int, queries_path: str, nq: int, L: int, k: int, threads: int = 8,
          timeout: float = 900):
    """Reference PSCAN (one reader + threads-1 workers); returns printed
    sqrt distances (nq, k) and wall seconds."""
    import time
    cmd = [PSCAN, "--dataset", os.path.abspath(data_path), "--dataset-size", str(n),
           "--queries", os.path.abspath(queries_path), "--queries-size", str(nq), "--timeseries-size", str(L),
           "--k", str(k), "--dbsize", str(min(n, 120000)), "--num-query-threads", str(max(2, threads + 1))]
    t0 = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, encoding="utf-8", errors="replace", timeout=timeout)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"pscan failed rc={r.returncode}: {r.stderr[-2000:]}")
    dist = np.full((nq, k), np.nan, np.float64)
    for m in _PDIST.finditer(r.stdout):
        qi, rank = int(m.group(2)), int(m.group(3))
        if 1 <= qi <= nq and 1 <= rank <= k:
            dist[qi - 1, rank - 1] = float(m.group(1))
    return dist, wall, r.stdout
