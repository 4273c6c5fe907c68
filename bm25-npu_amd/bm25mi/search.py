"""Search surfaces over the C ABI.

* Session            - the resident-index session (bm25mi_create / search / destroy).
* SearchSession      - Python Search(query term IDs, K) -> (scores, ids) with the
                       request/response checks of the reference's raw-ID wrapper
                       (standalone/run_bm25s_index_search.py:15-76).
* AscendIndexBm25    - Init / Search / Close / GetLastError with the facade's
                       status codes, output atomicity and state machine
                       (FR/ascendhost/include/index/AscendIndexBm25.h:17-57,
                       FR/ascendhost/src/bm25/AscendIndexBm25.cpp:33-84).
"""

import ctypes as C
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .corpus import pack

LIMIT = 2**31 - 1


def _is_int(value, low, high):
    return type(value) is int and low <= value <= high


class Session:
    """A resident index on one GPU. All searches go through libbm25mi."""

    def __init__(self, index, max_queries=8, max_words=24, max_k=10, block_docs=0, k1=None, device=0,
                 use_graph=False, grid_blocks=0, task_cost=0, slots=1, pool_entries=0, live_max_k=0, keep_cap=0):
        self._lib = _lib.load()
        cfg = _lib.Config()
        self._lib.bm25mi_config_default(C.byref(cfg))
        cfg.device_id = device
        cfg.max_queries = max_queries
        cfg.max_words = max_words
        cfg.max_k = max_k
        cfg.block_docs = block_docs
        cfg.k1 = index.k1 if k1 is None else k1
        cfg.use_graph = 1 if use_graph else 0
        cfg.grid_blocks = grid_blocks
        cfg.task_cost = task_cost
        cfg.slots = slots
        cfg.pool_entries = pool_entries
        cfg.live_max_k = live_max_k
        cfg.keep_cap = keep_cap
        self.config = cfg
        self._kcache = {}
        self.num_docs, self.vocab = index.num_docs, index.vocab
        handle = C.c_void_p()
        _lib.check(self._lib.bm25mi_create(_lib.ptr(index.offsets), index.vocab, _lib.ptr(index.docs),
                                           _lib.ptr(index.tf), index.postings, _lib.ptr(index.norm),
                                           index.num_docs, _lib.ptr(index.idf), C.byref(cfg), C.byref(handle)))
        self._h = handle

    @property
    def handle(self):
        if self._h is None or not self._h.value:
            raise _lib.Bm25Error(_lib.ILLEGAL_OPERATION, 'Session closed')
        return self._h

    def close(self):
        if self._h is not None and self._h.value:
            _lib.check(self._lib.bm25mi_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _pack(queries):
        if isinstance(queries, tuple):
            offsets, terms = queries
            return np.ascontiguousarray(offsets, np.int64), np.ascontiguousarray(terms, np.int32)
        return pack(queries)

    @classmethod
    def _prepare(cls, queries, k):
        offsets, terms = cls._pack(queries)
        nq = len(offsets) - 1
        ks = np.ascontiguousarray(np.full(nq, k, np.int32) if np.isscalar(k) else np.asarray(k, np.int32))
        if len(ks) != nq:
            raise ValueError('one K per query required')
        return offsets, terms, nq, ks

    def search(self, queries, k, thresholds=None, timed=False, learned=False, out=None):
        """Exact top-K per query. queries: list of term-ID lists or (offsets, terms).

        k: int or per-query array. thresholds: optional per-query predicted K-th
        scores (threshold path; results are still exact). learned: thresholds predicted on
        the GPU by the session's resident model (set_threshold_model) inside the Search.
        out: optional int32 buffer of 2 * (result rows) + Q entries that receives the scores
        (as FP32 bits), IDs and fallback flags, e.g. pre-filled with a poison pattern so that
        a caller sees every entry written. Returns (scores, ids, fallback) with rows
        concatenated per query (shape (Q, K) when k is an int), views of `out` when given.
        """
        offsets, terms = self._pack(queries)
        nq = len(offsets) - 1
        if np.isscalar(k):
            # K arrays of scalar-K calls are kept with their addresses (the library only reads them).
            hit = self._kcache.get((nq, k))
            if hit is None:
                arr = np.full(nq, k, np.int32)
                hit = (arr, _lib.ptr(arr))
                if len(self._kcache) < 64:
                    self._kcache[(nq, k)] = hit
            ks, pk = hit
            rows = nq * int(k)
        else:
            ks = np.ascontiguousarray(np.asarray(k, np.int32))
            if len(ks) != nq:
                raise ValueError('one K per query required')
            pk = _lib.ptr(ks)
            rows = int(ks.sum())
        # Scores, IDs and fallback flags share one buffer: one address lookup per call.
        if out is None:
            out = np.zeros(2 * rows + nq, np.int32)
        elif not (isinstance(out, np.ndarray) and out.dtype == np.int32 and out.shape == (2 * rows + nq,)
                  and out.flags.c_contiguous and out.flags.writeable):
            raise ValueError(f'out must be a writeable contiguous int32 array of {2 * rows + nq} entries')
        else:
            out[2 * rows:] = 0  # fallback flags: the plain entry point does not write them
        base = _lib.ptr(out)
        scores, ids, fallback = out[:rows].view(np.float32), out[rows:2 * rows], out[2 * rows:]
        pScores, pIds, pFallback = base, base + 4 * rows, base + 8 * rows
        # The converted thresholds stay bound to a local until the call returns: the library
        # reads them (twice: checks, then packing) while ctypes has released the GIL.
        thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float32)
        if thr is not None and thr.shape != (nq,):
            raise ValueError('one threshold per query required')
        pThr = _lib.ptr(thr)
        h = self.handle
        po, pt = _lib.ptr(offsets), _lib.ptr(terms)
        if learned:
            if pThr is not None:
                raise ValueError('a learned Search predicts its own thresholds')
            rc = self._lib.bm25mi_search_learned(h, po, pt, nq, pk, 1 if timed else 0, pScores, pIds, pFallback)
        elif timed:
            rc = self._lib.bm25mi_search_timed(h, po, pt, nq, pk, pThr, pScores, pIds, pFallback)
        elif pThr is not None:
            rc = self._lib.bm25mi_search_pruned(h, po, pt, nq, pk, pThr, pScores, pIds, pFallback)
        else:
            rc = self._lib.bm25mi_search_batch(h, po, pt, nq, pk, pScores, pIds)
        _lib.check(rc, h)
        if np.isscalar(k):
            scores, ids = scores.reshape(nq, int(k)), ids.reshape(nq, int(k))
        return scores, ids, fallback

    def features(self, queries, k):
        offsets, terms, nq, ks = self._prepare(queries, k)
        out = np.empty((nq, _lib.FEATURES), np.float32)
        _lib.check(self._lib.bm25mi_query_features(self.handle, _lib.ptr(offsets), _lib.ptr(terms), nq,
                                                   _lib.ptr(ks), _lib.ptr(out)), self.handle)
        return out

    def bounds(self, queries, k):
        """(lower, upper) float32[Q]: the K-th largest term impact (a lower bound of each
        query's K-th score) and the FP32 sum of term maxima (an upper bound of any score)."""
        offsets, terms, nq, ks = self._prepare(queries, k)
        lower = np.empty(nq, np.float32)
        upper = np.empty(nq, np.float32)
        _lib.check(self._lib.bm25mi_query_bounds(self.handle, _lib.ptr(offsets), _lib.ptr(terms), nq, _lib.ptr(ks),
                                                 _lib.ptr(lower), _lib.ptr(upper)), self.handle)
        return lower, upper

    def set_threshold_model(self, model, mean, scale, y_mean, y_scale, safety):
        """Makes a K-th-score model resident on the session's GPU for learned Searches: MLP
        parameters (layout of the bm25mi_mlp_* kernels), feature mean and scale, and the output
        transform threshold = safety * upper * exp(z * y_scale + y_mean)."""
        blob = np.ascontiguousarray(np.concatenate([np.asarray(model, np.float32), np.asarray(mean, np.float32),
                                                    np.asarray(scale, np.float32)]), np.float32)
        _lib.check(self._lib.bm25mi_set_threshold_model(self.handle, _lib.ptr(blob), len(blob), float(y_mean),
                                                        float(y_scale), float(safety)), self.handle)

    def learned_thresholds(self, queries, k):
        """float32[Q]: the thresholds a learned Search of these queries would use, predicted on
        the GPU by the resident model (0 = the query takes the plain path)."""
        offsets, terms, nq, ks = self._prepare(queries, k)
        out = np.empty(nq, np.float32)
        _lib.check(self._lib.bm25mi_learned_thresholds(self.handle, _lib.ptr(offsets), _lib.ptr(terms), nq,
                                                       _lib.ptr(ks), _lib.ptr(out)), self.handle)
        return out

    def set_launch(self, grid_waves=0, task_cost=0):
        """Re-plans the scoring launch (waves, schedule task cost); 0 keeps a value."""
        _lib.check(self._lib.bm25mi_set_launch(self.handle, grid_waves, task_cost), self.handle)

    def stats(self):
        st = _lib.Stats()
        _lib.check(self._lib.bm25mi_last_stats(self.handle, C.byref(st)), self.handle)
        return st.as_dict()


def check_request(queries, max_queries, max_words):
    """Request checks of the reference's raw-ID wrapper (standalone/run_bm25s_index_search.py:38-46):
    a list of 1..max_queries rows, each a list; at most max_words words in total; every word an
    int (bool rejected) in [0, INT32_MAX-1]. Raises ValueError."""
    if not (type(queries) is list and 1 <= len(queries) <= max_queries):
        raise ValueError('Query batch is invalid or exceeds declared capacity')
    total = 0
    for row in queries:
        if not (type(row) is list and len(row) <= max_words - total):
            raise ValueError('Query words exceed declared capacity')
        for word in row:
            if not _is_int(word, 0, LIMIT - 1):
                raise ValueError('Word IDs must be non-negative INT32 integers (bool rejected)')
        total += len(row)


def check_response(scores, ids, queries, k, num_docs):
    """Response checks of the same wrapper (:56-75) under strict ranking: shape (Q, K), float32
    scores and int32 IDs, finite non-negative scores, IDs in [0, num_docs) and unique per row,
    scores descending, IDs ascending among equal scores. Raises RuntimeError."""
    if not (scores.shape == ids.shape == (len(queries), k) and scores.dtype == np.float32
            and ids.dtype == np.int32 and np.isfinite(scores).all() and (scores >= 0).all()
            and (ids >= 0).all() and (ids < num_docs).all()):
        raise RuntimeError('Search output shape/dtype/score/id invalid')
    for s, i in zip(scores, ids):
        if len(np.unique(i)) != k or not (s[:-1] >= s[1:]).all() \
                or ((s[:-1] == s[1:]) & (i[:-1] >= i[1:])).any():
            raise RuntimeError('Search returned duplicate IDs or broke the strict order')


class SearchSession:
    """Search(query term IDs) -> (scores f32[Q,K], ids i32[Q,K]), K fixed per session.

    The Python raw-ID Search surface of standalone/run_bm25s_index_search.py:15-76: the
    configuration checks of its constructor (:21-29, with K up to the library's 1024 instead of
    32), the request checks of check_request and the response checks of check_response, around a
    resident-index Session. Term IDs at or above the vocabulary are skipped (no empty-token
    column is needed: an empty query returns IDs 0..K-1 with score 0).
    """

    def __init__(self, index, k, max_queries=8, max_words=24, **kwargs):
        if not (_is_int(k, 1, min(_lib.MAX_K, index.num_docs)) and _is_int(max_queries, 1, 10000)
                and _is_int(max_words, 0, LIMIT)):
            raise ValueError('Invalid Search configuration')
        self.k, self.maxq, self.maxwords = k, max_queries, max_words
        self.num_docs = index.num_docs
        self.session = Session(index, max_queries=max_queries, max_words=max(1, max_words), max_k=k, **kwargs)

    def search(self, queries):
        check_request(queries, self.maxq, self.maxwords)
        scores, ids, _ = self.session.search(queries, self.k)
        check_response(scores, ids, queries, self.k, self.num_docs)
        return scores, ids

    def close(self):
        self.session.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# APP_ERR classes of the facade (Bm25ErrorAdapter.h:11-45).
APP_ERR_OK = _lib.OK
APP_ERR_INVALID_PARAM = _lib.INVALID_PARAM
APP_ERR_ILLEGAL_OPERATION = _lib.ILLEGAL_OPERATION
APP_ERR_INNER_ERROR = _lib.INNER


@dataclass
class Bm25Config:
    """Bm25Config (AscendIndexBm25.h:17-25); the SDK context/resources/vendor fields
    have no counterpart on this stack."""
    deviceId: int = 0
    queryCount: int = 8
    maxQueryWords: int = 24
    topK: int = 10
    k1: float = 1.2
    blockDocs: int = 0


@dataclass
class Bm25SearchResult:
    scores: np.ndarray = field(default_factory=lambda: np.empty(0, np.float32))
    labels: np.ndarray = field(default_factory=lambda: np.empty(0, np.int64))


@dataclass
class ErrorInfo:
    aclCode: int = 0
    message: str = ''


class AscendIndexBm25:
    """Init/Search/Close/GetLastError with the reference facade's semantics:
    a second Init without Close is ILLEGAL_OPERATION; Search before Init is
    ILLEGAL_OPERATION; the result is replaced only on success; input errors keep
    the session usable; Close is idempotent; messages are cut to 1023 chars."""

    def __init__(self):
        self._mutex = threading.Lock()
        self._session = None
        self._k = 0
        self._queries = 0
        self._error = ErrorInfo()

    def _invoke(self, op):
        self._error = ErrorInfo()
        try:
            op()
            return APP_ERR_OK
        except _lib.Bm25Error as e:
            self._error = ErrorInfo(e.device_code, e.message[:1023])
            return e.code
        except ValueError as e:
            self._error = ErrorInfo(0, str(e)[:1023])
            return APP_ERR_INVALID_PARAM
        except Exception as e:  # noqa: BLE001 - the facade maps everything to a code
            self._error = ErrorInfo(0, str(e)[:1023])
            return APP_ERR_INNER_ERROR

    def Init(self, index, config):
        with self._mutex:
            def op():
                if self._session is not None:
                    raise _lib.Bm25Error(APP_ERR_ILLEGAL_OPERATION, 'BM25 is already initialized; Close before Init')
                checked = index.copy()
                checked.validate()
                self._session = Session(checked, max_queries=config.queryCount, max_words=config.maxQueryWords,
                                        max_k=config.topK, block_docs=config.blockDocs, k1=config.k1,
                                        device=config.deviceId)
                self._k = config.topK
                self._queries = config.queryCount
            return self._invoke(op)

    def Search(self, queries, result):
        with self._mutex:
            def op():
                if self._session is None:
                    raise _lib.Bm25Error(APP_ERR_ILLEGAL_OPERATION, 'BM25 is not initialized')
                if not (isinstance(queries, list) and all(isinstance(r, list) for r in queries)):
                    raise ValueError('Queries must be a list of lists of word IDs')
                if len(queries) != self._queries:
                    raise ValueError('Fixed session requires its declared query count')
                for r in queries:
                    for w in r:
                        if not _is_int(w, 0, LIMIT - 1):
                            raise ValueError('Invalid public query word ID')
                scores, ids, _ = self._session.search(queries, self._k)
                result.scores = scores.reshape(-1).copy()
                result.labels = ids.reshape(-1).astype(np.int64)
            return self._invoke(op)

    def Close(self):
        with self._mutex:
            def op():
                if self._session is not None:
                    self._session.close()
                    self._session = None
            return self._invoke(op)

    def GetLastError(self):
        with self._mutex:
            return ErrorInfo(self._error.aclCode, self._error.message)
