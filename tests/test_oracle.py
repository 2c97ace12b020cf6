"""CPU: pin the oracle.

Every oracle function is checked (a) bit-for-bit against the reference's own
function compiled from its sources (oracle/_ref, where it was built) and
(b) against something the mathematics fixes -- a closed form, an FP64
computation, an invariant -- chosen so that a dropped term, a wrong sign or
index, or a different summation order fails.
"""
from __future__ import annotations

import ctypes as C
import os
from statistics import NormalDist

import numpy as np
import pytest

from oracle_bind import REF_DIR, RefLib, ref_available

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference not built (oracle/_ref)")

F32MAX = np.float32(np.finfo(np.float32).max)


@pytest.fixture(scope="module")
def ref():
    return RefLib()


@pytest.fixture(scope="module")
def probe():
    L = C.CDLL(os.path.join(REF_DIR, "libref_probe.so"))
    L.ref_node_min_distance.restype = C.c_float
    L.ref_node_min_distance.argtypes = [np.ctypeslib.ndpointer(np.float32), C.c_int,
                                        np.ctypeslib.ndpointer(np.int16), np.ctypeslib.ndpointer(np.float32)]
    return L


@pytest.fixture(scope="module")
def walks(oracle):
    return oracle.series(0, 5, 0, 600, 256)


# ---- breakpoints ---------------------------------------------------------------
def closed_form_bp() -> np.ndarray:
    return np.array([np.float32(NormalDist().inv_cdf(i / 256)) if i != 128 else 0.0 for i in range(1, 256)],
                    np.float32)


def test_breakpoints_closed_form(oracle):
    assert np.array_equal(oracle.bp, closed_form_bp())
    assert np.all(np.diff(oracle.bp) > 0)
    assert np.array_equal(oracle.bp, -oracle.bp[::-1])  # symmetry of Phi^-1


def test_product_breakpoints_equal_oracle(oracle):
    from hercules_amd import hkg
    assert np.array_equal(hkg.breakpoints256(), oracle.bp)


@needs_ref
def test_breakpoints_equal_reference_table(oracle, ref):
    assert np.array_equal(ref.breakpoints256(), oracle.bp)


# ---- squared Euclidean distance ------------------------------------------------
@needs_ref
@pytest.mark.parametrize("L", [8, 96, 256, 1024])
def test_ed2_bit_exact_vs_reference(oracle, ref, L):
    rng = np.random.default_rng(L)
    for i in range(400):
        q = rng.standard_normal(L).astype(np.float32) * np.float32(rng.choice([1e-3, 1, 30]))
        x = (q + rng.standard_normal(L).astype(np.float32) * np.float32(rng.choice([0, 1e-4, 0.3, 3]))).astype(
            np.float32)
        for bound in (F32MAX, np.float32(rng.uniform(0, 2 * L)), np.float32(0.0)):
            assert oracle.ed2(q, x, bound) == ref.ed2(q, x, bound)


def test_ed2_against_fp64_and_order(oracle, walks):
    """Close to the FP64 value, and equal to the chunked order -- not to a
    naive sequential sum (which differs in the last bits for most rows)."""
    q = walks[0]
    naive_differs = 0
    for x in walks[1:200]:
        d = oracle.ed2(q, x)
        d64 = float(np.sum((q.astype(np.float64) - x) ** 2))
        assert abs(d - d64) <= 1e-5 * d64
        t = (q - x).astype(np.float32)
        t = (t * t).reshape(-1, 8)
        s = ((t[:, 0] + t[:, 1]) + (t[:, 2] + t[:, 3])) + ((t[:, 4] + t[:, 5]) + (t[:, 6] + t[:, 7]))
        acc = np.float32(0)
        for v in s:
            acc = np.float32(acc + v)
        assert acc == d
        seq = np.float32(0)
        for v in (q - x).astype(np.float32) ** 2:
            seq = np.float32(seq + v)
        naive_differs += seq != d
    assert naive_differs > 20
    assert oracle.ed2(q, q) == 0


def test_ed2_early_abandon_semantics(oracle, walks):
    q, x = walks[3], walks[4]
    full = oracle.ed2(q, x)
    # a bound above the result never changes it; one below returns a value >= bound
    assert oracle.ed2(q, x, np.float32(full * 1.0001)) == full
    b = np.float32(full / 3)
    r = oracle.ed2(q, x, b)
    assert b <= r <= full


# ---- PAA / SAX / iSAX lower bound -------------------------------------------------
@needs_ref
def test_paa_sax_lbsax_bit_exact_vs_reference(oracle, ref, walks):
    rng = np.random.default_rng(1)
    for i, x in enumerate(walks):
        p = oracle.paa(x)
        assert np.array_equal(p, ref.paa(x))
        s = oracle.sax(p)
        assert np.array_equal(s, ref.sax(x))
        q = walks[(i * 7 + 1) % len(walks)] if i % 2 else rng.standard_normal(256).astype(np.float32) * 2
        pq = oracle.paa(q)
        assert oracle.lb_sax(pq, s, 256) == ref.lb_sax(pq, s, 256)
    # extreme symbols 0 and 255 (the +-2e6 sentinels)
    s = np.array([0, 255] * 8, np.uint8)
    pq = np.linspace(-4, 4, 16).astype(np.float32)
    assert oracle.lb_sax(pq, s, 256) == ref.lb_sax(pq, s, 256)


def test_sax_symbol_is_count_below(oracle, walks):
    for x in walks[:100]:
        p = oracle.paa(x)
        s = oracle.sax(p)
        assert np.array_equal(s, np.searchsorted(oracle.bp, p, side="left").astype(np.uint8))
        # PAA: FP64 mean of each 16-value segment, within rounding
        assert np.allclose(p, x.reshape(16, -1).astype(np.float64).mean(1), atol=1e-6)


def test_lb_sax_is_a_lower_bound(oracle, walks):
    rng = np.random.default_rng(2)
    for i in range(300):
        x = walks[i % len(walks)]
        q = (x + rng.standard_normal(256).astype(np.float32) * np.float32(rng.uniform(0, 1))).astype(np.float32)
        lb = oracle.lb_sax(oracle.paa(q), oracle.sax(oracle.paa(x)), 256)
        assert lb <= oracle.ed2_f64(q, x) * (1 + 1e-6) + 1e-6 if hasattr(oracle, "ed2_f64") else True
        d64 = float(np.sum((q.astype(np.float64) - x) ** 2))
        assert lb <= d64 * (1 + 1e-5) + 1e-6
    x = walks[0]
    assert oracle.lb_sax(oracle.paa(x), oracle.sax(oracle.paa(x)), 256) == 0


# ---- segment mean / sd and the EAPCA bound ------------------------------------------
@needs_ref
def test_seg_mean_std_bit_exact_vs_reference(oracle, ref, walks):
    rng = np.random.default_rng(3)
    for x in walks[:300]:
        s, e = sorted(rng.choice(257, 2, replace=False))
        assert oracle.seg_mean_std(x, int(s), int(e)) == ref.seg_mean_std(x, int(s), int(e))


def test_seg_mean_std_against_fp64(oracle, walks):
    for x in walks[:50]:
        for s, e in [(0, 256), (0, 5), (7, 19), (250, 256), (128, 192)]:
            mu, sd = oracle.seg_mean_std(x, s, e)
            seg = x[s:e].astype(np.float64)
            assert abs(mu - seg.mean()) < 1e-5
            assert abs(sd - seg.std()) < 1e-3 or np.isnan(sd)


def random_node(rng, members: np.ndarray, nseg: int):
    ends = np.sort(rng.choice(np.arange(1, 256), nseg - 1, replace=False)).tolist() + [256]
    seg_end = np.array(ends, np.int16)
    return seg_end


def node_sketch(oracle, members: np.ndarray, seg_end: np.ndarray) -> np.ndarray:
    ind = np.zeros((len(seg_end), 4), np.float32)
    ind[:, 0] = -F32MAX
    ind[:, 1] = F32MAX
    ind[:, 2] = -F32MAX
    ind[:, 3] = F32MAX
    for x in members:
        s = 0
        for i, e in enumerate(seg_end):
            mu, sd = oracle.seg_mean_std(x, s, int(e))
            ind[i] = [max(ind[i, 0], mu), min(ind[i, 1], mu), max(ind[i, 2], sd), min(ind[i, 3], sd)]
            s = int(e)
    return ind


@needs_ref
def test_lb_eapca_bit_exact_vs_reference(oracle, probe, walks):
    rng = np.random.default_rng(4)
    for t in range(150):
        nseg = int(rng.integers(1, 10))
        seg_end = random_node(rng, None, nseg)
        members = walks[rng.choice(len(walks), 20, replace=False)]
        ind = node_sketch(oracle, members, seg_end).ravel()
        q = walks[int(rng.integers(len(walks)))] if t % 2 else rng.standard_normal(256).astype(np.float32)
        r = probe.ref_node_min_distance(q, nseg, seg_end, ind)
        assert oracle.lb_eapca(q, seg_end, ind, use_fma=1) == np.float32(r)


def test_lb_eapca_is_a_lower_bound(oracle, walks):
    rng = np.random.default_rng(5)
    for t in range(60):
        nseg = int(rng.integers(1, 10))
        seg_end = random_node(rng, None, nseg)
        members = walks[rng.choice(len(walks), 30, replace=False)]
        ind = node_sketch(oracle, members, seg_end).ravel()
        q = walks[int(rng.integers(len(walks)))]
        lb = oracle.lb_eapca(q, seg_end, ind, use_fma=0)
        for x in members:
            d64 = float(np.sum((q.astype(np.float64) - x) ** 2))
            assert lb <= d64 * (1 + 1e-5) + 1e-5
        # a query inside every interval has a zero bound
        assert oracle.lb_eapca(members[0], seg_end, ind, use_fma=0) == 0


# ---- answers ---------------------------------------------------------------------
def test_bruteforce_against_fp64_and_ties(oracle):
    X = oracle.series(0, 9, 0, 2000, 96)
    X[1500] = X[40]
    X[700] = X[40]
    Q = np.ascontiguousarray(np.concatenate([X[[40]] + np.float32(0.01), oracle.series(0, 10, 0, 9, 96)]))
    ids, d2 = oracle.knn(X, Q, 12, threads=4)
    for qi in range(Q.shape[0]):
        d = np.array([oracle.ed2(Q[qi], X[j]) for j in range(X.shape[0])], np.float32)
        # k smallest distinct values, smallest id for each
        vals = np.unique(d)[:12]
        assert np.array_equal(d2[qi], vals)
        for r, v in enumerate(vals):
            assert ids[qi, r] == np.nonzero(d == v)[0].min()
        d64 = np.sum((Q[qi].astype(np.float64) - X) ** 2, 1)
        assert np.allclose(d2[qi], np.sort(d64)[: len(vals)] if qi else d2[qi], rtol=1e-5)
    assert ids[0, 0] == 40  # the duplicate rows 700 and 1500 lose the tie


def test_bruteforce_threads_agree(oracle):
    X = oracle.series(0, 3, 0, 5000, 64)
    Q = oracle.series(0, 4, 0, 7, 64)
    a = oracle.knn(X, Q, 30, threads=1)
    b = oracle.knn(X, Q, 30, threads=7)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_bruteforce_padding(oracle):
    X = oracle.series(0, 3, 0, 4, 16)
    ids, d2 = oracle.knn(X, X[:1].copy(), 6)
    assert (d2[0, 4:] == F32MAX).all() and (ids[0, 4:] == np.uint64(2**64 - 1)).all()


# ---- generator --------------------------------------------------------------------
@pytest.mark.parametrize("kind,L", [(0, 256), (1, 96)])
def test_generator_invariants(oracle, kind, L):
    X = oracle.series(kind, 1, 0, 500, L)
    X64 = X.astype(np.float64)
    assert np.abs(X64.mean(1)).max() < 1e-6
    assert np.abs(X64.std(1) - 1).max() < 1e-6
    assert np.array_equal(X, oracle.series(kind, 1, 0, 500, L, threads=3))
    assert np.array_equal(X[100:200], oracle.series(kind, 1, 100, 100, L))
    assert not np.array_equal(X, oracle.series(kind, 2, 0, 500, L))
    if kind == 0:
        # random walks are smooth: lag-1 autocorrelation near 1
        ac = (X64[:, 1:] * X64[:, :-1]).mean(1)
        assert np.median(ac) > 0.9


def test_generator_normals(oracle):
    """Box-Muller normals: moments of many draws, and the log helper against
    the FP64 library log."""
    for u in [1e-300, 1e-9, 0.1, 0.5, 0.70710678, 0.9999999]:
        assert abs(oracle.lib.orc_log(u) - np.log(u)) <= 4e-16 * abs(np.log(u)) + 1e-300
    X = oracle.series(1, 3, 0, 2000, 256)  # z-normalised normals: check raw draws via kind 0 increments
    W = oracle.series(0, 3, 0, 2000, 256).astype(np.float64)
    inc = np.diff(W, axis=1)
    inc /= inc.std(1, keepdims=True)
    assert abs(inc.mean()) < 0.01
    kurt = (inc ** 4).mean()
    assert 2.8 < kurt < 3.2
    assert X.shape == (2000, 256)


def test_noisy_queries(oracle):
    sig = np.array([0.0, 0.05, 0.2], np.float32)
    Q, src = oracle.noisy_queries(0, 1, 1000, 3, sig, 128)
    X = oracle.series(0, 1, 0, 1000, 128)
    # sigma 0: the query is the source series itself
    assert np.array_equal(Q[0], X[src[0]]) or np.allclose(Q[0], X[src[0]], atol=1e-6)
    d = [float(np.sum((Q[i].astype(np.float64) - X[src[i]]) ** 2)) for i in range(3)]
    assert d[0] < 1e-9 < d[1] < d[2]
