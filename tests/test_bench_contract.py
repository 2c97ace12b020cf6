"""CPU: the benchmark's host-side contract -- it refuses more than one GPU
before touching a device, and the HBM traffic it reports comes from the
counter summaries in profiles/, looked up by the exact workload."""
from __future__ import annotations

import importlib.util
import os
import re
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_lookup_by_workload():
    b = _bench()
    rows = []
    for line in open(os.path.join(REPO, "profiles", "traffic.txt")):
        if line.startswith("#") or not line.strip():
            continue
        kernel, n, L, batch, nbytes = line.split()
        rows.append((kernel, int(n), int(L), int(batch), float(nbytes)))
    assert rows
    for kernel, n, L, batch, nbytes in rows:
        assert nbytes > 0
        assert b.traffic_bytes(kernel, n, L, batch) == nbytes
    # a workload without counter runs reports no traffic rather than another's
    assert b.traffic_bytes("k_bf16_filter", 100_000_000, 256, 512) is None
    assert b.traffic_bytes("k_bf16_filter", 50_000_000, 256, 2048) is None
    assert b.traffic_bytes("no_such_kernel", 100_000_000, 256, 1024) is None


def test_profiles_named_in_traffic_file_exist():
    text = open(os.path.join(REPO, "profiles", "traffic.txt")).read()
    names = set(re.findall(r"(k_\w+\.md)", text))
    assert names
    for nm in names:
        assert os.path.exists(os.path.join(REPO, "profiles", nm)), nm
