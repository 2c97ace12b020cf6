"""hkg_scan_knn_device_noalloc on an MI355X: the full scan with its scratch in
the library's own device array gives the answers of the allocating entry
point (hkg_knn_device, path scan) bit for bit, on two calls in a row, the
second with fewer queries and a smaller kmax; and the oracle's brute force
agrees.  Requests beyond the array are refused before any launch."""
from __future__ import annotations

import numpy as np
import pytest

from hercules_amd import hkg

pytestmark = pytest.mark.gpu


def _device(a: np.ndarray) -> hkg.DeviceArray:
    d = hkg.DeviceArray(max(a.nbytes, 4))
    d.upload(a)
    return d


def _run_both(ix, Q: np.ndarray, k: np.ndarray, kmax: int):
    nq = Q.shape[0]
    dQ, dk = _device(Q), _device(k)
    outs = []
    for noalloc in (True, False):
        dids, dd2 = hkg.DeviceArray(nq * kmax * 8), hkg.DeviceArray(nq * kmax * 4)
        if noalloc:
            ix.scan_knn_device_noalloc(dQ, nq, dk, kmax, dids, dd2)
        else:
            ix.knn_device(dQ, nq, dk, kmax, dids, dd2, path="scan")
        outs.append((dids.download(np.empty((nq, kmax), np.uint64)), dd2.download(np.empty((nq, kmax), np.float32))))
    return outs


def test_noalloc_scan_equals_allocating_scan(oracle):
    n, L = 30000, 256
    X = oracle.series(0, 1, 0, n, L)
    dX = hkg.DeviceArray(X.nbytes)
    dX.upload(X)
    rng = np.random.default_rng(17)
    with hkg.HerculesIndex.build_device(dX, n, L, leaf_size=1000) as ix:
        for call, (nq, kmax, seed) in enumerate([(40, 50, 3), (7, 10, 4)]):
            print(f"no-alloc call {call}: {nq} queries, kmax {kmax}", flush=True)
            sig = rng.uniform(0.0, 0.2, nq).astype(np.float32)
            Q, _ = oracle.noisy_queries(0, 1, n, seed, sig, L)
            k = rng.integers(1, kmax + 1, nq).astype(np.uint32)
            k[0] = kmax
            (ia, da), (ib, db) = _run_both(ix, Q, k, kmax)
            assert np.array_equal(ia, ib)
            assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
            oid, od2 = oracle.knn(X, Q, kmax)
            for q in range(nq):
                kq = int(k[q])
                assert np.array_equal(ia[q, :kq], oid[q, :kq])
                assert np.array_equal(da[q, :kq].view(np.uint32), od2[q, :kq].view(np.uint32))
                assert np.all(da[q, kq:] == hkg.FLT_MAX) and np.all(ia[q, kq:] == hkg.NO_ID)


def test_noalloc_refuses_beyond_its_scratch(oracle):
    n, L = 2000, 256
    X = oracle.series(0, 1, 0, n, L)
    with hkg.HerculesIndex.open_raw(X) as ix:
        nq, kmax = 512, 256
        dQ = hkg.DeviceArray(nq * L * 4)
        dk = _device(np.full(nq, kmax, np.uint32))
        dids, dd2 = hkg.DeviceArray(nq * kmax * 8), hkg.DeviceArray(nq * kmax * 4)
        with pytest.raises(hkg.HkgError) as e:
            ix.scan_knn_device_noalloc(dQ, nq, dk, kmax, dids, dd2)
        assert e.value.status == hkg.HKG_EINVAL and "scratch" in str(e.value)
        # a per-query k outside [1, kmax] is caught by the device check
        Q = X[:3].copy()
        dQ3, dk3 = _device(Q), _device(np.array([1, 0, 2], np.uint32))
        with pytest.raises(hkg.HkgError) as e:
            ix.scan_knn_device_noalloc(dQ3, 3, dk3, 2, dids, dd2)
        assert e.value.status == hkg.HKG_EINVAL
