"""Write the golden fixtures under tests/golden/ (run in the build container,
where oracle/_ref -- the reference compiled from its sources -- exists).

Each fixture holds inputs' seeds and the answers: the reference CLI's exact
FP32 squared distances per rank (printed in hex by oracle/ref_interpose.c)
and the oracle's brute-force ids and distances.  Datasets are not stored:
they are regenerated from the seeds by the generator specification.

    python tests/golden/make_golden.py [tiny|c1|emb|all|modes|real]
"""
from __future__ import annotations

import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ref_runner as rr  # noqa: E402
from oracle_bind import Oracle  # noqa: E402

CONFIGS = {
    # name: kind, N, L, data_seed, leaf, queries (independent walks seed, n) + noisy (seed, sigmas)
    "tiny": dict(kind=0, N=20000, L=256, seed=11, leaf=500, qseed=12, nq=40, nseed=13,
                 sigmas=[0.0, 0.01, 0.1, 0.2] * 5, ks=[1, 10, 100]),
    "c1": dict(kind=0, N=1_000_000, L=256, seed=1, leaf=10000, qseed=2, nq=100, nseed=None, sigmas=[],
               ks=[1, 10, 100]),
    "emb": dict(kind=1, N=20000, L=96, seed=21, leaf=400, qseed=22, nq=20, nseed=23,
                sigmas=[0.05, 0.1, 0.2, 0.4] * 5, ks=[1, 10]),
}


def queries(o: Oracle, c: dict) -> np.ndarray:
    Q = o.series(c["kind"], c["qseed"], 0, c["nq"], c["L"])
    if c["nseed"] is not None and c["sigmas"]:
        Qn, _ = o.noisy_queries(c["kind"], c["seed"], c["N"], c["nseed"], np.array(c["sigmas"], np.float32),
                                c["L"])
        Q = np.concatenate([Q, Qn])
    return np.ascontiguousarray(Q)


def make(name: str) -> None:
    c = CONFIGS[name]
    o = Oracle()
    work = tempfile.mkdtemp(prefix=f"golden_{name}_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        t0 = time.time()
        X = o.series(c["kind"], c["seed"], 0, c["N"], c["L"])
        Q = queries(o, c)
        nq = Q.shape[0]
        rr.write_f32(os.path.join(work, "data.bin"), X)
        rr.write_f32(os.path.join(work, "q.bin"), Q)
        idx = os.path.join(work, "idx")
        rr.build_index(os.path.join(work, "data.bin"), c["N"], c["L"], c["leaf"], idx, threads=8)
        kmax = max(c["ks"])
        ids, d2 = o.knn(X, Q, kmax, threads=8)
        out = {"ids": ids, "d2": d2, "queries_hash": np.frombuffer(Q.tobytes(), np.uint8)[:0]}
        timing = {}
        for k in c["ks"]:
            rd2, rdist, wall, _ = rr.query(idx, os.path.join(work, "q.bin"), nq, c["L"], k, threads=8, sims=True)
            out[f"ref_sims_d2_k{k}"] = rd2
            timing[f"ref_sims_k{k}_wall_s"] = wall
            if not np.array_equal(rd2.view(np.uint32), d2[:, :k].view(np.uint32)):
                bad = np.nonzero((rd2 != d2[:, :k]).any(1))[0]
                print(f"{name} k={k}: reference SIMS differs from brute force on queries {bad}")
        # the serial path on the smallest k
        k0 = min(c["ks"])
        rd2, _, wall, _ = rr.query(idx, os.path.join(work, "q.bin"), nq, c["L"], k0, threads=1, sims=False)
        out[f"ref_serial_d2_k{k0}"] = rd2
        timing[f"ref_serial_k{k0}_wall_s"] = wall
        pdist, pwall, _ = rr.pscan(os.path.join(work, "data.bin"), c["N"], os.path.join(work, "q.bin"), nq,
                                   c["L"], k0, threads=8)
        out[f"ref_pscan_dist_k{k0}"] = pdist
        timing[f"ref_pscan_k{k0}_wall_s"] = pwall
        # FP64 error of every returned d2
        err = np.zeros_like(d2, dtype=np.float64)
        for qi in range(nq):
            sel = ids[qi] != np.uint64(2**64 - 1)
            x = X[ids[qi][sel].astype(np.int64)].astype(np.float64)
            d64 = np.sum((Q[qi].astype(np.float64) - x) ** 2, 1)
            err[qi, sel] = np.abs(d2[qi, sel] - d64) / np.maximum(d64, 1e-30)
        out["rel_err_fp64"] = err.astype(np.float32)
        del out["queries_hash"]
        np.savez_compressed(os.path.join(HERE, f"{name}_answers.npz"), **out)
        meta = dict(c, n_queries=nq, timing_8threads_build_container=timing,
                    max_rel_err_fp64=float(err.max()), generated_s=round(time.time() - t0, 1))
        if name in ("tiny", "emb"):
            shutil.copy(os.path.join(idx, "root.idx"), os.path.join(HERE, f"{name}_root.idx"))
            L = c["L"]
            raw = np.fromfile(os.path.join(idx, "leaves_raw.idx"), np.float32).reshape(-1, L)
            lookup = {X[i].tobytes(): i for i in range(X.shape[0])}
            rowmap = np.array([lookup[r.tobytes()] for r in raw], np.uint32)
            np.save(os.path.join(HERE, f"{name}_rowmap.npy"), rowmap)
            sims = np.fromfile(os.path.join(idx, "leaves_sims.idx"), np.uint8)
            meta["leaves_sims_sha1"] = __import__("hashlib").sha1(sims.tobytes()).hexdigest()
        with open(os.path.join(HERE, f"{name}_meta.json"), "w") as fh:
            json.dump(meta, fh, indent=1)
        print(name, json.dumps(meta["timing_8threads_build_container"]), "max rel err", meta["max_rel_err_fp64"])
    finally:
        shutil.rmtree(work, ignore_errors=True)


# approximate modes on the committed reference-built indexes:
# (k, epsilon, approx budget, eapca threshold, sax threshold) and (k, nprobes)
MODE_CONFIGS = {
    "tiny": [dict(k=10, epsilon=5.0, budget=1, eapca=0.25, sax=0.9),
             dict(k=10, epsilon=20.0, budget=1, eapca=0.5, sax=0.999),
             dict(k=100, epsilon=0.5, budget=3, eapca=0.25, sax=0.9),
             dict(k=1, epsilon=2.0, budget=1, eapca=0.9, sax=0.99),
             dict(k=100, epsilon=3.0, budget=1, eapca=0.99, sax=0.99),
             dict(k=10, epsilon=1.0, budget=80, eapca=0.25, sax=0.9),
             dict(k=1, nprobes=1), dict(k=10, nprobes=3), dict(k=100, nprobes=7), dict(k=10, nprobes=25)],
    "emb": [dict(k=10, epsilon=4.0, budget=1, eapca=0.25, sax=0.9),
            dict(k=10, epsilon=10.0, budget=2, eapca=0.8, sax=0.99),
            dict(k=1, epsilon=1.0, budget=1, eapca=0.25, sax=0.999),
            dict(k=1, nprobes=2), dict(k=10, nprobes=5)],
}


def mode_flags(c: dict) -> list:
    if c.get("nprobes"):
        return ["--nprobes", c["nprobes"]]
    return ["--epsilon", c["epsilon"], "--approx-stop-condition", c["budget"]]


def make_modes(name: str) -> None:
    """The reference CLI's answers in its epsilon and nprobes modes on the
    committed reference-built index (same rows, same tree)."""
    import golden_data as gd
    o = Oracle()
    m = gd.meta(name)
    work = tempfile.mkdtemp(prefix=f"golden_modes_{name}_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        gd.reference_index(o, name, os.path.join(work, "idx"))
        Q = gd.queries(o, m)
        rr.write_f32(os.path.join(work, "q.bin"), Q)
        out = {}
        for i, c in enumerate(MODE_CONFIGS[name]):
            rd2, _, _, _ = rr.query(os.path.join(work, "idx"), os.path.join(work, "q.bin"), Q.shape[0], m["L"],
                                    c["k"], threads=8, sims=True, eapca_threshold=c.get("eapca", 0.25),
                                    sax_threshold=c.get("sax", 0.9), extra=mode_flags(c))
            out[f"m{i}"] = rd2
        np.savez_compressed(os.path.join(HERE, f"{name}_modes.npz"), **out)
        with open(os.path.join(HERE, f"{name}_modes.json"), "w") as fh:
            json.dump(MODE_CONFIGS[name], fh, indent=1)
        print(name, "modes", len(out))
    finally:
        shutil.rmtree(work, ignore_errors=True)


# real-data query files (tests/golden/real_queries) on the synthetic
# datasets of their length: the committed tiny (len 256) and embedding
# (len 96) indexes, and a 20K-series random-walk set of len 128
REAL_SETS = {
    "seismic": dict(index="tiny", files=["queries_hard1p_seismic_len256_znorm.bin",
                                         "queries_hard10p_seismic_len256_znorm.bin",
                                         "queries_size100_seismic_len256_znorm.bin"]),
    "deep1b": dict(index="emb", files=["queries_hard1p_deep1b_len96_znorm.bin",
                                       "queries_orig100_deep1b_len96_znorm.bin"]),
    "sald": dict(index=None, kind=0, N=20000, L=128, seed=31, leaf=500,
                 files=["queries_hard1p_sald_len128_znorm.bin"]),
}
REAL_MODES = [dict(k=10, epsilon=2.0, budget=1, eapca=0.25, sax=0.9), dict(k=10, nprobes=5)]


def real_queries(name: str) -> np.ndarray:
    import golden_data as gd
    return gd.real_queries(name)


def make_real(name: str) -> None:
    """The reference CLI's exact answers (k = 10), and on the committed
    indexes its epsilon / nprobes answers, for a real query set."""
    import golden_data as gd
    c = REAL_SETS[name]
    o = Oracle()
    Q = real_queries(name)
    work = tempfile.mkdtemp(prefix=f"golden_real_{name}_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        idx = os.path.join(work, "idx")
        if c["index"]:
            m = gd.meta(c["index"])
            raw, _ = gd.reference_index(o, c["index"], idx)
            X = o.series(m["kind"], m["seed"], 0, m["N"], m["L"])
            L = m["L"]
        else:
            L = c["L"]
            X = o.series(c["kind"], c["seed"], 0, c["N"], L)
            rr.write_f32(os.path.join(work, "data.bin"), X)
            rr.build_index(os.path.join(work, "data.bin"), c["N"], L, c["leaf"], idx, threads=8)
        rr.write_f32(os.path.join(work, "q.bin"), Q)
        out = {}
        rd2, _, _, _ = rr.query(idx, os.path.join(work, "q.bin"), Q.shape[0], L, 10, threads=8, sims=True)
        ids, d2 = o.knn(X, Q, 10, threads=8)
        if not np.array_equal(rd2.view(np.uint32), d2.view(np.uint32)):
            raise RuntimeError(f"{name}: reference SIMS differs from brute force")
        out["exact_d2"], out["exact_ids"] = d2, ids
        if c["index"]:
            for i, mc in enumerate(REAL_MODES):
                md2, _, _, _ = rr.query(idx, os.path.join(work, "q.bin"), Q.shape[0], L, mc["k"], threads=8,
                                        sims=True, eapca_threshold=mc.get("eapca", 0.25),
                                        sax_threshold=mc.get("sax", 0.9), extra=mode_flags(mc))
                out[f"m{i}"] = md2
        np.savez_compressed(os.path.join(HERE, f"real_{name}.npz"), **out)
        with open(os.path.join(HERE, f"real_{name}.json"), "w") as fh:
            json.dump(dict(c, modes=REAL_MODES if c["index"] else []), fh, indent=1)
        print("real", name, Q.shape)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which == "real":
        for n in REAL_SETS:
            make_real(n)
    elif which == "modes":
        for n in MODE_CONFIGS:
            make_modes(n)
    else:
        for n in (CONFIGS if which == "all" else [which]):
            make(n)
