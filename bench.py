#!/usr/bin/env python3
"""Benchmark: exact k-NN queries/s on one MI355X over an index resident in HBM.

Workload (defaults): 100M z-normalised random walks of length 256 (FP32,
102.4 GB in HBM) generated on the device, an EAPCA index built on the device
by the bulk loader (leaf 10000), and queries that are noisy copies of
dataset series with sigma ~ U[0, 0.2] and k ~ U{1..100} (seeded).  A "step"
is one batch of queries answered exactly through the BF16 path (EAPCA leaf
bounds -> best-leaf scan -> BF16 MFMA certified bound -> exact FP32
refinement -> top-k), replayed as one captured hipGraph.  Inputs are
resident in HBM when the timed region starts.  Further legs time the PSCAN
full scan, the plain FP32 paths, the approximate modes, serving under a
sweep of offered loads, and the reference's own CPU path (cpu_baseline).

Prints one JSON line (rank 0).  `python bench.py --gpus N` with N > 1 exits:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse() -> argparse.Namespace:
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=100_000_000, help="series in the index")
    p.add_argument("--length", type=int, default=256)
    p.add_argument("--batch", type=int, default=0,
                   help="queries per step (0: the batch size hkg_autotune picks for --path on this index, timed "
                        "before the timed steps)")
    p.add_argument("--leaf", type=int, default=10000)
    p.add_argument("--path", default="bf16", choices=["batched", "index", "scan", "bf16"])
    p.add_argument("--sigma-max", type=float, default=0.2)
    p.add_argument("--kmax", type=int, default=100)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-series", type=int, default=10_000_000,
                   help="CPU baseline: series of the reference-built index the reference's SIMS path answers on")
    p.add_argument("--cpu-queries", type=int, default=32, help="CPU baseline: queries (k = --cpu-k)")
    p.add_argument("--cpu-k", type=int, default=50, help="CPU baseline: k (the reference answers one k per run)")
    p.add_argument("--cpu-pscan-queries", type=int, default=4, help="CPU baseline: queries of the reference PSCAN")
    p.add_argument("--scan-passes", type=int, default=8, help="scan leg: single-query passes timed (0: skip)")
    p.add_argument("--fp32-batch", type=int, default=2048, help="FP32 batched leg: queries per step (0: skip)")
    p.add_argument("--fp32-steps", type=int, default=3, help="FP32 batched leg: timed steps")
    p.add_argument("--index-queries", type=int, default=64, help="plain per-query FP32 leg: queries (0: skip)")
    p.add_argument("--parity-queries", type=int, default=8)
    p.add_argument("--row-probes", type=int, default=32, help="dataset rows looked up exactly after the build")
    p.add_argument("--bf16-bound-pass", type=int, default=0, choices=[0, 1, 2])
    p.add_argument("--bf16-query-tiles", type=int, default=0, choices=[0, 2, 3, 4],
                   help="BF16 filter: 32-query MFMA tiles per wave (0: the autotuner's choice, else 2)")
    p.add_argument("--serve-qps", type=float, default=0.0,
                   help="offered load of the serving leg (0: a sweep at --serve-loads of the measured throughput)")
    p.add_argument("--serve-loads", default="0.5,0.8,0.95",
                   help="serving sweep: offered loads as fractions of the measured throughput")
    p.add_argument("--serve-inflight", type=int, default=2, help="serving: batches on the device at once")
    p.add_argument("--serve-requests", type=int, default=16384, help="requests of the serving leg (0: skip)")
    p.add_argument("--mode-epsilon", type=float, default=1.0,
                   help="epsilon of the approximate-mode leg (0: skip that leg)")
    p.add_argument("--mode-nprobes", type=int, default=10, help="nprobes of the ng-mode leg (0: skip that leg)")
    p.add_argument("--mode-queries", type=int, default=256, help="queries of each mode leg (from the first step)")
    p.add_argument("--mode-cpu-queries", type=int, default=32,
                   help="modes beside the reference CLI on a 1M-series index: queries (0: skip)")
    p.add_argument("--real-queries", type=int, default=16,
                   help="real-data leg: Seismic query files on the index; this many checked against the full "
                        "scan (0: skip the leg)")
    p.add_argument("--loader-series", type=int, default=0,
                   help="loader leg: write an index of this many series to TMPDIR and time hkg_open_dir (0: skip)")
    p.add_argument("--tune-queries", type=int, default=4096,
                   help="with --batch 0: sample queries (seed 6, same noise and k) the autotuner times the "
                        "candidate batch sizes of --path on")
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="one device-to-device copy of this size first (reference dispatch for counter runs)")
    return p.parse_args()


def host_info() -> dict:
    """The host the CPU baseline ran on: threads, CPU model, RAM."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    ram = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                ram = round(int(line.split()[1]) / 1024 ** 2, 1)
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "ram_gib": ram}


def cpu_baseline(hkg, args) -> dict:
    """The reference's own multithreaded SIMS query path
    (exact_de_knn_search_psq_lf_sims_adaptive, hercules_query_engine.c:706-1083,
    compiled from its sources into oracle/_ref with the fixes the oracle
    applies) on every host core, over an index the reference itself builds
    from the first cpu_series rows of the same dataset (leaf as the GPU
    run), at k = cpu_k; measured at that N, not scaled.  The reference PSCAN
    on the same rows is a second figure.  The GPU answers the same queries
    on the same reference-built index, and the squared distances are
    compared bit for bit."""
    import shutil
    sys.path.insert(0, os.path.join(REPO, "tests"))
    try:
        import ref_runner as rr
    except Exception as e:  # pragma: no cover
        return {"error": f"reference runner unavailable: {e}"}
    if not rr.available():
        return {"error": "reference binaries not built (oracle/_ref)"}
    host = host_info()
    threads = max(2, host["nproc"] or 2)
    n, L, nq, k = min(args.cpu_series, args.n), args.length, args.cpu_queries, args.cpu_k
    d = tempfile.mkdtemp(prefix="hkg_cpu_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        t0 = time.perf_counter()
        data = os.path.join(d, "data.bin")
        chunk = 1_000_000
        dX = hkg.DeviceArray(min(chunk, n) * L * 4)
        with open(data, "wb") as f:
            for first in range(0, n, chunk):
                m = min(chunk, n - first)
                hkg.gen_series_device(0, 1, first, m, L, dX)
                f.write(dX.download(np.empty((m, L), np.float32)).tobytes())
        dX.close()
        gen_s = time.perf_counter() - t0
        rq = np.random.default_rng(15)
        sig = rq.uniform(0.0, args.sigma_max, nq).astype(np.float32)
        Q, _ = hkg.gen_noisy_queries(0, 1, n, 15, sig, L)
        qf = os.path.join(d, "q.bin")
        rr.write_f32(qf, Q)
        idx = os.path.join(d, "idx")
        log(f"cpu baseline: {n} x {L} written in {gen_s:.1f} s; reference build on {threads} threads")
        tb = time.perf_counter()
        rr.build_index(data, n, L, args.leaf, idx, threads=threads, timeout=1800)
        build_s = time.perf_counter() - tb
        with open(os.path.join(idx, "leaves_raw.idx"), "rb") as f:  # page cache warm
            while f.read(64 << 20):
                pass
        log(f"cpu baseline: reference index built in {build_s:.1f} s; {nq} queries, k = {k}")
        rd2, _, wall, out = rr.query(idx, qf, nq, L, k, threads=threads, sims=True, timeout=1800)
        # the same run with one query: the CLI's start-up and index load, which
        # the per-query rate leaves out
        _, _, wall1, _ = rr.query(idx, qf, 1, L, k, threads=threads, sims=True, timeout=1800)
        per = max(wall - wall1, 1e-9) / max(nq - 1, 1)
        # the same queries through the engine on the same index
        with hkg.HerculesIndex.open(idx) as ix:
            same = True
            gpu = {}
            for path in ("bf16", "batched"):
                ix.knn(Q[:2], k, path=path)
                hkg.lib().hkg_device_sync()
                tg = time.perf_counter()
                r = ix.knn(Q, k, path=path, batch=nq)
                gpu[path] = nq / (time.perf_counter() - tg)
                same &= bool(np.array_equal(r.d2.view(np.uint32), rd2.view(np.uint32)))
        npq = min(args.cpu_pscan_queries, nq)
        pscan_qps = None
        if npq:
            rr.write_f32(os.path.join(d, "qp.bin"), Q[:npq])
            _, pwall, _ = rr.pscan(data, n, os.path.join(d, "qp.bin"), npq, L, k, threads=threads, timeout=1800)
            pscan_qps = npq / pwall
        return {"value": 1.0 / per, "unit": "queries/s", "cores": threads, "kind": "reference",
                "sample": f"the reference's SIMS index path (exact_de_knn_search_psq_lf_sims_adaptive, "
                          f"--num-threads {threads} --sims --eapca-threshold 0.25 --sax-threshold 0.9) on an "
                          f"index it built from the first {n} rows of the same dataset (leaf {args.leaf}); {nq} "
                          f"noisy queries of those rows (sigma ~ U[0, {args.sigma_max}], seed 15), k = {k}; "
                          f"measured at {n} series (not scaled); queries/s = (nq - 1) / (wall(nq) - wall(1))",
                "n_series": n, "k": k, "queries": nq, "host": host,
                "reference_wall_s": round(wall, 2), "reference_wall_1_query_s": round(wall1, 2),
                "reference_build_s": round(build_s, 1),
                "pscan_queries_per_s": pscan_qps, "pscan_queries": npq,
                "gpu_same_index_queries_per_s": gpu,
                "identical_d2_to_reference": same}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def mode_cpu_leg(hkg, args, Q: np.ndarray) -> dict:
    """The exact search and the epsilon and nprobes modes beside the
    reference CLI on the same index: a 1M-series index (the first 1M rows of
    the dataset, the reference's CPU-sized case) built on the device and
    written in the reference's format; the reference (compiled from its
    sources into oracle/_ref) and the device answer the same queries with
    the same settings, and the answers are compared bit for bit."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    try:
        import ref_runner as rr
    except Exception as e:  # pragma: no cover
        return {"error": f"reference runner unavailable: {e}"}
    if not rr.available():
        return {"error": "reference binaries not built (oracle/_ref)"}
    import shutil
    threads = max(2, min(16, os.cpu_count() or 2))
    n, L, nq, k = min(1_000_000, args.n), args.length, args.mode_cpu_queries, 10
    out = {"index": f"first {n} rows, leaf {args.leaf}, built on the device", "queries": nq, "k": k,
           "reference_threads": threads}
    d = tempfile.mkdtemp(prefix="hkg_mode_cpu_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        dX = hkg.DeviceArray(n * L * 4)
        hkg.gen_series_device(0, 1, 0, n, L, dX)
        with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=args.leaf) as built:
            built.write(os.path.join(d, "idx"))
        rr.write_f32(os.path.join(d, "q.bin"), Q[:nq])
        with hkg.HerculesIndex.open(os.path.join(d, "idx")) as ix:
            legs = ([("epsilon", {"epsilon": args.mode_epsilon},
                      ["--epsilon", args.mode_epsilon, "--approx-stop-condition", 80])] if args.mode_epsilon > 0 else []) + \
                   ([("ng", {"nprobes": args.mode_nprobes}, ["--nprobes", args.mode_nprobes])]
                    if args.mode_nprobes > 0 else [])
            # the exact search too: the reference's SIMS path against the
            # batched index path
            legs = [("exact", {"path": "batched"}, [])] + legs
            for name, kw, flags in legs:
                ix.knn(Q[:2], k, **kw)  # warm the code path
                hkg.lib().hkg_device_sync()
                t = time.perf_counter()
                r = ix.knn(Q[:nq], k, batch=nq, **kw)
                gpu_s = time.perf_counter() - t
                rd2, _, wall, _ = rr.query(os.path.join(d, "idx"), os.path.join(d, "q.bin"), nq, L, k,
                                           threads=threads, sims=True, extra=flags)
                # the same run with one query: its wall time is the CLI's start-up
                # and index load, which the per-query rate leaves out
                _, _, wall1, _ = rr.query(os.path.join(d, "idx"), os.path.join(d, "q.bin"), 1, L, k,
                                          threads=threads, sims=True, extra=flags)
                per = max(wall - wall1, 1e-9) / (nq - 1) if nq > 1 else wall
                out[name] = {"settings": kw, "gpu_queries_per_s": nq / gpu_s, "reference_queries_per_s": 1.0 / per,
                             "reference_wall_s": wall, "reference_wall_1_query_s": wall1,
                             "identical_to_reference": bool(np.array_equal(r.d2.view(np.uint32), rd2.view(np.uint32)))}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def traffic_bytes(kernel: str, n: int, L: int, batch: int):
    """HBM bytes per launch of a kernel at this workload (series, length,
    queries per launch), from the counter runs summarised in profiles/
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy of known
    size); None when that workload was not measured."""
    path = os.path.join(REPO, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if line.startswith("#") or len(parts) < 5:
            continue
        if parts[0] == kernel and (int(parts[1]), int(parts[2]), int(parts[3])) == (n, L, batch):
            return float(parts[4])
    return None


def serving_leg(hkg, ix, args, Qh: np.ndarray, kq: np.ndarray, qps: float) -> dict:
    """Poisson stream of single-query requests through the continuous-
    batching server; latency = submit -> result ready, as the server reports
    it.  A sample of the served answers is compared with direct queries.
    Before the stream the new server is warmed, untimed: two full batches at
    once (every worker runs one) and one request of each k class (their
    buffer sets), as a server takes some traffic before it is put in
    service."""
    import threading

    nreq = args.serve_requests
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, nreq))
    rows = rng.integers(0, Qh.shape[0], nreq)
    kmax = int(kq.max())
    srv = hkg.Server(ix, max_batch=args.batch, path=args.path, max_inflight=args.serve_inflight, max_k=kmax,
                     bf16_bound_pass=args.bf16_bound_pass, scratch_bytes=32 << 30)
    t_warm = time.perf_counter()
    warm = [(i % Qh.shape[0], kmax) for i in range(args.serve_inflight * args.batch)]
    warm_cls = sorted({min(1 << i, kmax) for i in range(kmax.bit_length() + 1)})
    tw =[srv.submit(Qh[q], k) for q, k in warm]
    for t_, (q, k) in zip(tw, warm):
        srv.wait(t_, k)
    for c in warm_cls:
        srv.wait(srv.submit(Qh[0], c), c)
    warm_s = time.perf_counter() - t_warm
    tickets = [None] * nreq
    ready = threading.Semaphore(0)

    def producer():
        t0 = time.perf_counter()
        for i in range(nreq):
            while True:
                wait = arrivals[i] - (time.perf_counter() - t0)
                if wait <= 0:
                    break
                if wait > 2e-3:
                    time.sleep(wait - 1e-3)
            tickets[i] = srv.submit(Qh[rows[i]], int(kq[rows[i]]))
            ready.release()

    res = [None] * nreq
    lat = np.empty(nreq)
    th = threading.Thread(target=producer)
    t_start = time.perf_counter()
    th.start()
    for i in range(nreq):
        ready.acquire()
        ids, d2, lat_us = srv.wait(tickets[i], int(kq[rows[i]]))
        res[i] = (ids, d2)
        lat[i] = lat_us
    wall = time.perf_counter() - t_start
    th.join()
    srv.stop()
    sample = np.unique(rng.integers(0, nreq, 64))
    same = True
    for i in sample:
        r = ix.knn(Qh[rows[i]][None, :], int(kq[rows[i]]), path=args.path)
        same &= bool(np.array_equal(r.ids[0], res[i][0]) and
                     np.array_equal(r.d2[0].view(np.uint32), res[i][1].view(np.uint32)))
    return {"offered_qps": round(qps, 1), "achieved_qps": round(nreq / wall, 1), "requests": nreq,
            "p50_ms": round(float(np.percentile(lat, 50)) / 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) / 1e3, 3),
            "max_batch": args.batch, "inflight": args.serve_inflight, "path": args.path, "scratch_gib": 32,
            "warm_up_s": round(warm_s, 3), "warm_up_requests": len(warm) + len(warm_cls),
            "answers_identical_to_direct_queries": same, "sampled": int(len(sample))}


def scan_leg(hkg, ix, Qh, kq, dQ, dk, kmax: int, dids, dd2, args, N: int, L: int) -> dict:
    """PSCAN (hkg path scan) over the whole index: passes of one query and of
    four queries (the workload's queries and k), each timed on its stream by
    the stage stamps, in the scan grid hkg_autotune picks on 8 other queries
    of the workload; algorithmic bytes 4 N L per pass."""
    tune = ix.autotune(Qh[-8:], kq[-8:], path="scan")
    out = {"tuning": tune}
    nbytes = 4.0 * N * L
    for qn, passes in ((1, args.scan_passes), (4, max(2, args.scan_passes // 2))):
        qv, kv = _views(hkg, dQ, dk, 0, qn, L)
        ix.knn_device(qv, qn, kv, kmax, dids, dd2, path="scan", batch=qn, profile=True)  # warm
        ms = []
        for p in range(passes):
            qv, kv = _views(hkg, dQ, dk, p, qn, L)
            st = ix.knn_device(qv, qn, kv, kmax, dids, dd2, path="scan", batch=qn, profile=True)
            ms.append(st["stage_ms"]["scan"] / max(st["stage_launches"]["scan"], 1))
        avg = float(np.mean(ms))
        tb = traffic_bytes("k_scan", N, L, qn)
        out[f"queries_per_pass_{qn}"] = {
            "passes": passes, "avg_pass_ms": avg, "min_pass_ms": float(np.min(ms)), "max_pass_ms": float(np.max(ms)),
            "algorithmic_bytes_per_pass": nbytes, "algorithmic_gbs": nbytes / (avg * 1e6),
            "algorithmic_frac_of_hbm_peak": nbytes / (avg * 1e6) / HBM_PEAK_GBS,
            "traffic_bytes_per_pass": tb, "traffic_gbs": tb / (avg * 1e6) if tb else None,
            "queries_per_s": qn / (avg * 1e-3)}
    return out


def fp32_legs(hkg, ix, dQ, dk, kmax: int, dids, dd2, args, L: int, bf16_qps: float) -> dict:
    """The plain FP32 GPU paths at the headline workload (the same queries):
    the batched path (persistent (query, leaf) work queue, iSAX filter +
    exact refinement, captured graph) and the per-query index path."""
    out = {}
    B = args.fp32_batch
    if B:
        def step(i):
            qv, kv = _views(hkg, dQ, dk, i, B, L)
            return ix.knn_device(qv, B, kv, kmax, dids, dd2, path="batched", batch=B, profile=True)
        step(0)  # warm: buffer set and graph
        hkg.lib().hkg_device_sync()
        t = time.perf_counter()
        sts = [step(i) for i in range(args.fp32_steps)]
        wall = time.perf_counter() - t
        ref_ms = sum(st["stage_ms"]["refine"] for st in sts) / len(sts)
        ref_b = sum(st["stage_bytes"]["refine"] for st in sts) / len(sts)
        tb = traffic_bytes("k_isax_filter", args.n, L, B)
        out["batched"] = {
            "batch": B, "steps": args.fp32_steps, "queries_per_s": args.fp32_steps * B / wall,
            "ms_per_step": wall * 1e3 / args.fp32_steps,
            "stages_ms": {kk: round(sum(st["stage_ms"][kk] for st in sts) / len(sts), 3)
                          for kk in sts[0]["stage_ms"] if sts[0]["stage_ms"][kk] > 0},
            "kernel": "k_isax_filter", "kernel_ms": ref_ms,
            # 16 B per iSAX word streamed (once per leaf per 64 queries) + at
            # most 4 L B per refined (query, series) pair (an upper bound: the
            # refinement abandons early)
            "kernel_algorithmic_bytes_upper_bound": ref_b,
            "kernel_algorithmic_gbs_upper_bound": ref_b / (ref_ms * 1e6) if ref_ms else None,
            # HBM bytes the counters measured per launch (profiles/traffic.txt)
            "kernel_traffic_bytes": tb,
            "kernel_traffic_gbs": tb / (ref_ms * 1e6) if tb and ref_ms else None,
            "kernel_traffic_frac_of_hbm_peak": tb / (ref_ms * 1e6) / HBM_PEAK_GBS if tb and ref_ms else None,
            "series_lb_checked_per_query": sum(st["series_lb_checked"] for st in sts) / (len(sts) * B),
            "series_refined_per_query": sum(st["series_refined"] for st in sts) / (len(sts) * B),
            "leaves_after_eapca_per_query": sum(st["leaves_after_eapca"] for st in sts) / (len(sts) * B)}
        out["bf16_over_fp32_batched"] = bf16_qps / out["batched"]["queries_per_s"]
    M = args.index_queries
    if M:
        qv, kv = _views(hkg, dQ, dk, 0, M, L)
        ix.knn_device(qv, 2, kv, kmax, dids, dd2, path="index")  # warm
        hkg.lib().hkg_device_sync()
        t = time.perf_counter()
        st = ix.knn_device(qv, M, kv, kmax, dids, dd2, path="index")
        wall = time.perf_counter() - t
        out["index"] = {"queries": M, "queries_per_s": M / wall, "ms_per_query": wall * 1e3 / M,
                        "series_refined_per_query": st["series_refined"] / M}
        out["bf16_over_fp32_index"] = bf16_qps / out["index"]["queries_per_s"]
    return out


def loader_leg(hkg, args, L: int, Qh: np.ndarray) -> dict:
    """An index of loader_series series built on the device, written in the
    reference's format (root.idx, leaves_raw.idx, leaves_sims.idx) and read
    back by hkg_open_dir; the files are fresh, so they come from the page
    cache: the figure is the loader's pinned-staging upload, not a disk's.
    Both indexes answer the same queries identically."""
    import shutil
    n = args.loader_series
    dX = hkg.DeviceArray(n * L * 4)
    hkg.gen_series_device(0, 7, 0, n, L, dX)
    d = tempfile.mkdtemp(prefix="hkg_loader_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        built = hkg.HerculesIndex.build_device(dX, n, L, leaf_size=args.leaf)
        del dX
        tw = time.perf_counter()
        built.write(d)
        write_s = time.perf_counter() - tw
        Qs = Qh[:16]
        a = built.knn(Qs, 10, path="batched")
        built.close()
        t0 = time.perf_counter()
        ix2 = hkg.HerculesIndex.open(d)
        open_s = time.perf_counter() - t0
        b = ix2.knn(Qs, 10, path="batched")
        ix2.close()
        nbytes = n * L * 4 + n * 16
        return {"series": n, "bytes": nbytes, "open_s": open_s, "open_gbs": nbytes / open_s / 1e9,
                "write_s": write_s, "answers_identical": bool(np.array_equal(a.d2.view(np.uint32), b.d2.view(np.uint32)))}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _views(hkg, dQ, dk, step: int, B: int, L: int):
    """Non-owning views of one step's queries and k values."""
    return (hkg.DeviceArray.view(dQ.ptr + step * B * L * 4, B * L * 4),
            hkg.DeviceArray.view(dk.ptr + step * B * 4, B * 4))


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the index would shard by leaf, one contiguous range of whole leaves per GPU)", file=sys.stderr)
        return 2
    from hercules_amd import hkg

    L, N, B = args.length, args.n, args.batch
    if args.calibrate_copy_mib:
        nb = args.calibrate_copy_mib << 20
        ca, cb = hkg.DeviceArray(nb), hkg.DeviceArray(nb)
        hkg._check(hkg.lib().hkg_memcpy(cb.ptr, ca.ptr, nb, 3))
        del ca, cb
        log(f"calibration copy: {args.calibrate_copy_mib} MiB device to device")
    t0 = time.perf_counter()
    dX = hkg.DeviceArray(N * L * 4)
    hkg.gen_series_device(0, 1, 0, N, L, dX)
    log(f"generated {N} x {L} in HBM: {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    ix = hkg.HerculesIndex.build_device(dX, N, L, leaf_size=args.leaf)
    build_s = time.perf_counter() - t1
    info = ix.info()
    log(f"built index: {info['n_leaves']} leaves, {build_s:.1f} s, {info['device_bytes'] / 1e9:.1f} GB")

    # launch shape: the autotuner's choice for the path on this index, from
    # sample queries of their own (not the timed ones)
    tuning = {}
    if B == 0:
        rt = np.random.default_rng(6)
        m = args.tune_queries
        sigt = rt.uniform(0.0, args.sigma_max, m).astype(np.float32)
        kt = rt.integers(1, args.kmax + 1, m).astype(np.uint32)
        Qt, _ = hkg.gen_noisy_queries(0, 1, N, 6, sigt, L)
        t2 = time.perf_counter()
        tuning = ix.autotune(Qt, kt, path=args.path)
        tuning["queries"] = m
        tuning["seconds"] = round(time.perf_counter() - t2, 2)
        B = tuning["batch"]
        del Qt
        log(f"autotuned {args.path} on {m} sample queries: {tuning}")

    steps, warm = args.steps, args.warmup
    total_q = (steps + warm) * B
    rng = np.random.default_rng(5)
    sig = rng.uniform(0.0, args.sigma_max, total_q).astype(np.float32)
    kq = rng.integers(1, args.kmax + 1, total_q).astype(np.uint32)
    Qh, src = hkg.gen_noisy_queries(0, 1, N, 5, sig, L)
    dQ = hkg.DeviceArray(Qh.nbytes)
    dQ.upload(Qh)
    dk = hkg.DeviceArray(kq.nbytes)
    dk.upload(kq)
    kmax = int(kq.max())
    dids = hkg.DeviceArray(B * kmax * 8)
    dd2 = hkg.DeviceArray(B * kmax * 4)
    qt = args.bf16_query_tiles if args.path == "bf16" else 0
    # per-step views into the resident query array
    def run(step: int, profile: bool) -> dict:
        qv, kv = _views(hkg, dQ, dk, step, B, L)
        return ix.knn_device(qv, B, kv, kmax, dids, dd2, path=args.path, batch=B, profile=profile,
                             bf16_bound_pass=args.bf16_bound_pass, bf16_query_tiles=qt)

    for s in range(warm):
        st = run(s, True)
        log(f"warmup step {s}: {st['device_ms']:.1f} ms")
    hkg.lib().hkg_device_sync()
    stage_ms = {}
    stage_bytes = {}
    stage_launch = {}
    step_ms = []
    refined = checked = pairs = 0
    t = time.perf_counter()
    for s in range(warm, warm + steps):
        ts = time.perf_counter()
        st = run(s, True)
        step_ms.append((time.perf_counter() - ts) * 1e3)
        for kk, v in st["stage_ms"].items():
            stage_ms[kk] = stage_ms.get(kk, 0.0) + v
            stage_bytes[kk] = stage_bytes.get(kk, 0) + st["stage_bytes"][kk]
            stage_launch[kk] = stage_launch.get(kk, 0) + st["stage_launches"][kk]
        refined += st["series_refined"]
        checked += st["series_lb_checked"]
        pairs += st.get("bf16_pairs", 0)
        log(f"step {s}: {step_ms[-1]:.1f} ms, refined {st['series_refined']}, "
            f"lb-checked {st['series_lb_checked']}, leaves {st['leaves_after_eapca']}")
    hkg.lib().hkg_device_sync()
    elapsed = time.perf_counter() - t
    qps = steps * B / elapsed
    # the same steps again without the stage stamps (a cross-check that the
    # stamp kernels in the graph cost nothing measurable)
    nchk = min(steps, 5)
    hkg.lib().hkg_device_sync()
    t = time.perf_counter()
    for s in range(warm, warm + nchk):
        run(s, False)
    hkg.lib().hkg_device_sync()
    qps_nostamps = nchk * B / (time.perf_counter() - t)

    # dominant kernel and its roofline
    dom = max(stage_ms, key=lambda k: stage_ms[k])
    launches = max(stage_launch[dom], 1)
    avg_ms = stage_ms[dom] / launches
    bytes_per_launch = stage_bytes[dom] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    mfma = args.path == "bf16" and dom == "refine"
    if mfma:
        # the BF16 filter is a contraction: 2 L N B FLOPs per batch pass
        flops_per_launch = 2.0 * L * N * B
        achieved_tf = flops_per_launch / (avg_ms * 1e-3) / 1e12
        # the MFMAs actually issued: waves skip the row tiles none of their
        # queries needs, so the executed pairs fall below N B
        exec_flops = 2.0 * L * pairs / launches
        exec_tf = exec_flops / (avg_ms * 1e-3) / 1e12
    kernel_names = {"refine": "k_filter_refine", "approx": "k_scan_best", "eapca": "k_eapca", "scan": "k_scan",
                    "topk": "k_select", "collect": "k_collect", "prep": "k_prep", "select_leaf": "k_best_leaf"}
    if args.path == "bf16":
        kernel_names.update({"refine": "k_bf16_filter", "bound": "k_bf16_filter (bound pass)",
                             "topk": "k_bf16_refine"})

    # parity at full size: index path == full scan on a sample of the queries
    npar = min(args.parity_queries, B)
    par = {}
    if npar:
        Qs = Qh[:npar]
        ks = kq[:npar]
        a = ix.knn(Qs, ks, path=args.path)
        b = ix.knn(Qs, ks, path="scan")
        same = bool(np.array_equal(a.ids, b.ids) and np.array_equal(a.d2.view(np.uint32), b.d2.view(np.uint32)))
        # error of every returned d2 against the FP64 distance to the row with
        # that id (the row regenerated on the device from (seed, id))
        worst, nres = 0.0, 0
        for i in range(npar):
            q64 = Qs[i].astype(np.float64)
            for r in range(int(ks[i])):
                x64 = hkg.gen_series(0, 1, int(a.ids[i, r]), 1, L)[0].astype(np.float64)
                e64 = float(np.sum((q64 - x64) ** 2))
                worst = max(worst, abs(float(a.d2[i, r]) - e64) / max(e64, 1e-30))
                nres += 1
        par = {"index_vs_full_scan_identical": same, "queries_checked": npar,
               "fp64_max_rel_error": worst, "results_checked_fp64": nres}
        log(f"parity index vs scan on {npar} queries: {same}; worst FP64 relative error of {nres} results: "
            f"{worst:.3g}")
    if args.row_probes:
        # every probed dataset row is in the index: found at distance 0 under its own id
        rows = np.unique(np.concatenate([np.linspace(0, N - 1, args.row_probes).astype(np.int64),
                                         np.random.default_rng(3).integers(0, N, args.row_probes)]))
        P = np.concatenate([hkg.gen_series(0, 1, int(r), 1, L) for r in rows])
        rr = ix.knn(P, 1, path=args.path)
        found = int(np.sum((rr.d2[:, 0] == 0.0) & (rr.ids[:, 0] == rows.astype(np.uint64))))
        par["rows_found_exactly"] = f"{found}/{len(rows)}"
        log(f"dataset rows found at distance 0: {found}/{len(rows)}")

    scan = scan_leg(hkg, ix, Qh, kq, dQ, dk, kmax, dids, dd2, args, N, L) if args.scan_passes else {}
    if scan:
        log(f"scan: {scan}")
    args.fp32_batch = min(args.fp32_batch, B)
    args.fp32_steps = max(1, min(args.fp32_steps, (steps + warm) - 1))
    fp32 = fp32_legs(hkg, ix, dQ, dk, kmax, dids, dd2, args, L, qps) if (args.fp32_batch or args.index_queries) else {}
    if fp32:
        log(f"fp32 paths: {fp32}")

    # the reference's epsilon and no-guarantee modes on one batch of the same
    # queries: throughput, and how their answers compare with the exact ones
    modes = {}
    legs = ([("epsilon", {"epsilon": args.mode_epsilon})] if args.mode_epsilon > 0 else []) + \
           ([("ng", {"nprobes": args.mode_nprobes})] if args.mode_nprobes > 0 else [])
    M = min(args.mode_queries, B)
    if legs and M:
        exact = ix.knn(Qh[warm * B:warm * B + M], kq[warm * B:warm * B + M], path=args.path)
        kv = kq[warm * B:warm * B + M]
        w = exact.d2.shape[1]
        ok = np.arange(w)[None, :] < kv[:, None]
        for name, kw in legs:
            qv, kvv = _views(hkg, dQ, dk, warm, B, L)
            hkg.lib().hkg_device_sync()
            tm = time.perf_counter()
            st = ix.knn_device(qv, M, kvv, kmax, dids, dd2, batch=M, profile=True, **kw)
            hkg.lib().hkg_device_sync()
            ms = (time.perf_counter() - tm) * 1e3
            d2m = dd2.download(np.empty((M, kmax), np.float32))[:, :w]
            ex = exact.d2.astype(np.float64)
            ratio = np.where(ok, d2m.astype(np.float64) / np.maximum(ex, 1e-30), 1.0)
            leg = {"settings": kw, "queries": M, "queries_per_s": M / (ms * 1e-3), "ms_per_batch": ms,
                   "queries_with_exact_answer": float(np.mean(np.all(~ok | (d2m == exact.d2), axis=1))),
                   "max_d2_ratio_to_exact": float(ratio.max()),
                   "leaves_scanned_first_per_query": st["mode_leaves_first"] / M,
                   "candidate_leaves_per_query": st["mode_candidates"] / M,
                   "sequential_branch_queries": st["mode_sequential"],
                   "stages_ms": {kk: round(v, 3) for kk, v in st["stage_ms"].items() if v > 0},
                   # k_mode_search: the rows of every leaf it scans, 4 L bytes each
                   "mode_search_algorithmic_gbs": (st["stage_bytes"]["approx"] / (st["stage_ms"]["approx"] * 1e6)
                                                   if st["stage_ms"]["approx"] > 0 else None)}
            if name == "epsilon":
                leg["within_1_plus_epsilon"] = bool(ratio.max() <= (1.0 + kw["epsilon"]) * (1 + 1e-6))
            modes[name] = leg
            log(f"mode {name}: {leg}")

    # real-data queries (the reference's Seismic query files, len 256) on the
    # same index: throughput of one batch, and the answers of a few against
    # the full scan
    if legs and args.mode_cpu_queries and not args.no_cpu_baseline:
        modes["beside_reference_1M"] = mode_cpu_leg(hkg, args, Qh)
        log(f"modes beside the reference: {modes['beside_reference_1M']}")

    real = {}
    if args.real_queries and L == 256:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import golden_data as gd
        Qr = gd.real_queries("seismic")
        kr = np.full(Qr.shape[0], 10, np.uint32)
        ix.knn(Qr[:8], 10, path=args.path)  # warm the shape
        hkg.lib().hkg_device_sync()
        tr = time.perf_counter()
        rr_ = ix.knn(Qr, kr, path=args.path, batch=Qr.shape[0])
        hkg.lib().hkg_device_sync()
        ms = (time.perf_counter() - tr) * 1e3
        sc = ix.knn(Qr[:args.real_queries], 10, path="scan")
        same = bool(np.array_equal(rr_.ids[:args.real_queries], sc.ids) and
                    np.array_equal(rr_.d2[:args.real_queries].view(np.uint32), sc.d2.view(np.uint32)))
        real = {"queries": "Seismic hard1p + hard10p + size100 (300, len 256), k = 10",
                "queries_per_s_host_to_host": Qr.shape[0] / (ms * 1e-3), "ms": ms,
                "device_ms": rr_.stats["device_ms"],
                "series_refined_per_query": rr_.stats["series_refined"] / Qr.shape[0],
                "identical_to_full_scan": same, "checked": int(args.real_queries)}
        log(f"real queries: {real}")

    loader = {}
    if args.loader_series:
        loader = loader_leg(hkg, args, L, Qh)
        log(f"loader: {loader}")

    serving = None
    if args.serve_requests:
        loads = [args.serve_qps] if args.serve_qps > 0 else [f * qps for f in map(float, args.serve_loads.split(","))]
        serving = []
        for inflight in sorted({1, args.serve_inflight}):
            # batches of up to the step's (tuned) size, so that the server can
            # carry the throughput the offered loads are fractions of
            args_i = argparse.Namespace(**{**vars(args), "serve_inflight": inflight, "batch": B})
            for sq in loads:
                leg = serving_leg(hkg, ix, args_i, Qh, kq, sq)
                leg["offered_fraction_of_throughput"] = round(sq / qps, 3)
                serving.append(leg)
                log(f"serving: {leg}")

    cpu = None if args.no_cpu_baseline else cpu_baseline(hkg, args)
    if cpu:
        log(f"cpu baseline: {cpu}")
    tb = traffic_bytes(kernel_names.get(dom, dom), N, L, B)
    # measured HBM rate of the dominant kernel: counter bytes per launch / launch time
    traffic_gbs = tb / (avg_ms * 1e6) if tb and avg_ms else None
    out = {
        "metric": "exact k-NN queries/s at 100M\u00d7len256; achieved HBM GB/s; p50/p99 latency",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": elapsed * 1e3 / steps,
        "step_ms_p50": float(np.percentile(step_ms, 50)),
        "step_ms_p99": float(np.percentile(step_ms, 99)),
        "step_ms_max": float(np.max(step_ms)),
        "value_without_stage_stamps": qps_nostamps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: z-normalised random walks generated on the device (seed 1); queries are dataset "
                "series plus N(0, sigma^2) noise, sigma ~ U[0, 0.2], renormalised (seed 5)",
        "config": {"workload": f"{N // 1_000_000}M series x len {L}, leaf {args.leaf}, {B} queries per step, "
                               f"k ~ U{{1..{args.kmax}}}, path {args.path}",
                   "n_series": N, "length": L, "batch": B, "leaf_size": args.leaf, "n_leaves": info["n_leaves"],
                   "bf16_query_tiles": (qt or tuning.get("bf16_query_tiles") or 2) if args.path == "bf16" else None,
                   "build_s": round(build_s, 2)},
        "roofline": ({"bound": "mfma", "kernel": kernel_names.get(dom, dom), "achieved": achieved_tf,
                      "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved_tf / MFMA_BF16_PEAK_TFLOPS,
                      "traffic": tb, "traffic_gbs": traffic_gbs, "avg_launch_ms": avg_ms,
                      "flops_per_launch": flops_per_launch, "executed_flops_per_launch": exec_flops,
                      "executed_tflops": exec_tf, "executed_frac": exec_tf / MFMA_BF16_PEAK_TFLOPS,
                      "tile_bytes_per_launch": bytes_per_launch,
                      "tile_bytes_gbs": achieved}
                     if mfma else
                     {"bound": "hbm", "kernel": kernel_names.get(dom, dom), "achieved": achieved,
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                      "traffic": tb, "traffic_gbs": traffic_gbs, "avg_launch_ms": avg_ms,
                      "bytes_per_launch": bytes_per_launch}),
        "stages_ms": {k: round(v / steps, 3) for k, v in stage_ms.items() if v > 0},
        "series_refined_per_query": refined / (steps * B),
        "series_lb_checked_per_query": checked / (steps * B),
        "parity": par,
        "scan": scan,
        "fp32_paths": fp32,
        "modes": modes,
        "real_queries": real,
        "loader": loader,
        "tuning": tuning,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ix.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
