"""Continuous-batching server over a resident-index session (bm25mi_server_* in the library;
no reference counterpart).

The server keeps up to `slots` batches in flight (a Session created with slots=N): requests
arriving while others run on the GPU are admitted into the next free slot, up to max_batch
queries and the slot's candidate-pool memory, and their work overlaps on the device. Each
request's result equals a plain Search of the same query alone.
"""

import ctypes as C

import numpy as np

from . import _lib
from .corpus import pack


def percentiles(latency_us):
    """p50 / p95 / p99 (nearest rank) and mean of latencies in milliseconds."""
    v = np.sort(np.asarray(latency_us, np.float64)) / 1e3
    rank = lambda p: float(v[max(0, int(np.ceil(p * len(v))) - 1)])  # noqa: E731
    return {'p50_ms': rank(0.50), 'p95_ms': rank(0.95), 'p99_ms': rank(0.99), 'mean_ms': float(v.mean())}


class Server:
    def __init__(self, session, max_batch=None):
        self._lib = _lib.load()
        self.session = session
        handle = C.c_void_p()
        max_batch = session.config.max_queries if max_batch is None else max_batch
        _lib.check(self._lib.bm25mi_server_start(session.handle, max_batch, C.byref(handle)))
        self._h = handle

    def submit(self, terms, k, threshold=None, learned=False):
        """Queues one query; with a threshold (or learned: predicted on the device) it runs
        pruned, and reruns plain if it must."""
        terms = np.ascontiguousarray(terms, np.int32)
        ticket = C.c_uint64()
        tp = _lib.ptr(terms) if terms.size else None
        if learned:
            if threshold is not None:
                raise ValueError('a learned request predicts its own threshold')
            _lib.check(self._lib.bm25mi_server_submit_learned(self._h, tp, len(terms), k, C.byref(ticket)))
        elif threshold is None:
            _lib.check(self._lib.bm25mi_server_submit(self._h, tp, len(terms), k, C.byref(ticket)))
        else:
            _lib.check(self._lib.bm25mi_server_submit_pruned(self._h, tp, len(terms), k, float(threshold),
                                                             C.byref(ticket)))
        return ticket.value

    def wait(self, ticket, k):
        """(scores f32[k], ids i32[k], latency in microseconds)."""
        scores = np.empty(k, np.float32)
        ids = np.empty(k, np.int32)
        lat = C.c_double()
        _lib.check(self._lib.bm25mi_server_wait(self._h, C.c_uint64(ticket), _lib.ptr(scores), _lib.ptr(ids),
                                                C.byref(lat)))
        return scores, ids, lat.value

    def run_poisson(self, queries, ks, qps, seed=0, keep_results=True, thresholds=None, learned=False):
        """Open-loop load: query i is submitted at the i-th arrival of a seeded Poisson process
        of rate qps. Returns latencies (us), sustained QPS, percentiles and, when asked, the
        concatenated result rows."""
        offsets, terms = pack(queries) if not isinstance(queries, tuple) else queries
        offsets = np.ascontiguousarray(offsets, np.int64)
        terms = np.ascontiguousarray(terms, np.int32)
        ks = np.ascontiguousarray(ks, np.int32)
        thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
        n = len(ks)
        if ks.ndim != 1 or offsets.shape != (n + 1,) or (thr is not None and thr.shape != (n,)):
            raise ValueError('offsets, ks and thresholds must describe the same queries')
        if n and (offsets[0] != 0 or offsets[-1] > len(terms) or (np.diff(offsets) < 0).any()):
            raise ValueError('query offsets must start at 0, ascend and stay within the terms')
        lat = np.empty(n, np.float64)
        rows = int(ks.sum())
        scores = np.empty(rows, np.float32) if keep_results else None
        ids = np.empty(rows, np.int32) if keep_results else None
        elapsed = C.c_double()
        _lib.check(self._lib.bm25mi_server_run_poisson(self._h, _lib.ptr(offsets), _lib.ptr(terms), n, _lib.ptr(ks),
                                                       _lib.ptr(thr), 1 if learned else 0, float(qps),
                                                       C.c_uint64(seed), _lib.ptr(lat),
                                                       _lib.ptr(scores), _lib.ptr(ids), C.byref(elapsed)))
        out = {'latency_us': lat, 'elapsed_s': elapsed.value, 'offered_qps': float(qps),
               'sustained_qps': n / elapsed.value, **percentiles(lat)}
        if keep_results:
            out['scores'], out['ids'] = scores, ids
        return out

    def counters(self):
        b, q, m, f = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64()
        _lib.check(self._lib.bm25mi_server_counters(self._h, C.byref(b), C.byref(q), C.byref(m), C.byref(f)))
        return {'batches': b.value, 'queries': q.value, 'max_inflight': m.value, 'fallbacks': f.value}

    def stop(self):
        if self._h is not None and self._h.value:
            _lib.check(self._lib.bm25mi_server_stop(self._h))
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
