"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the reference's
own CPU Search built into oracle/_ref/. TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.
"""

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / 'liboracle.so'
REF_LIB = HERE / '_ref' / 'libref_cpu.so'
REF_DIR = HERE / '_ref'
P, i32, i64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double

_lib = None
_ref = None


def build(reference=None):
    """Compiles liboracle.so, and oracle/_ref when the reference checkout is present."""
    subprocess.run(['make', '-s', '-C', str(HERE), 'oracle'], check=True)
    ref = Path(reference or os.environ.get('BM25_REFERENCE', '/root/reference'))
    if (ref / 'feature_retrieval').is_dir():
        subprocess.run(['make', '-s', '-C', str(HERE), 'ref', f'REF={ref}'], check=True)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.oracle_search.restype = C.c_int
        _lib.oracle_search.argtypes = [P, i64, P, P, P, i64, P, f32, P, P, i32, P, i32, i32, i32, P, P, P]
        _lib.oracle_bm25s_search.restype = C.c_int
        _lib.oracle_bm25s_search.argtypes = [P, i64, P, P, i64, P, P, i32, i32, i32, P, P]
        _lib.oracle_bm25_weights.restype = None
        _lib.oracle_bm25_weights.argtypes = [P, i64, P, i64, f64, f64, P, P]
        _lib.oracle_contribution.restype = f32
        _lib.oracle_contribution.argtypes = [f32, f32, f32, f32]
    return _lib


def ref_available():
    return REF_LIB.exists()


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(str(REF_LIB))
        _ref.ref_cpu_search.restype = C.c_int
        _ref.ref_cpu_search.argtypes = [P, i64, P, P, i64, P, i64, P, P, P, i32, i32, i32, i32, P, P, P]
        _ref.ref_cpu_create.restype = P
        _ref.ref_cpu_create.argtypes = [P, i64, P, P, i64, P, i64, P, i32, i32, i32, i32, i32, i64]
        _ref.ref_cpu_run.restype = C.c_int
        _ref.ref_cpu_run.argtypes = [P, P, P, i32, P, P]
        _ref.ref_cpu_destroy.restype = None
        _ref.ref_cpu_destroy.argtypes = [P]
    return _ref


def kp1(k1):
    """(k1 + 1) rounded to FP32, the factor the reference writes as 2.2F for k1 = 1.2."""
    return float(np.float32(k1 + 1.0))


def search(index, offsets, terms, k, shards=8, threads=0, chunk=64, full=False, k1=None):
    """Exact top-K per query; k is an int or per-query array. Returns (scores, ids[, full rows])."""
    offsets = np.ascontiguousarray(offsets, np.int64)
    terms = np.ascontiguousarray(terms, np.int32)
    nq = len(offsets) - 1
    ks = np.ascontiguousarray(np.full(nq, k, np.int32) if np.isscalar(k) else np.asarray(k, np.int32))
    rows = int(ks.sum())
    scores = np.empty(rows, np.float32)
    ids = np.empty(rows, np.int32)
    fulls = np.empty(nq * index.num_docs, np.float32) if full else None
    rc = lib().oracle_search(_p(index.offsets), index.vocab, _p(index.docs), _p(index.tf), _p(index.norm),
                             index.num_docs, _p(index.idf), kp1(index.k1 if k1 is None else k1), _p(offsets),
                             _p(terms if terms.size else np.zeros(1, np.int32)), nq, _p(ks), shards, threads,
                             chunk, _p(scores), _p(ids), _p(fulls))
    if rc != 0:
        raise ValueError('oracle_search rejected its input')
    if np.isscalar(k):
        scores, ids = scores.reshape(nq, int(k)), ids.reshape(nq, int(k))
    if full:
        return scores, ids, fulls.reshape(nq, index.num_docs)
    return scores, ids


def bm25s_search(index, impacts, offsets, terms, k, threads=0):
    """The bm25s-numba retrieve algorithm restated (dense row + heap top-k)."""
    offsets = np.ascontiguousarray(offsets, np.int64)
    terms = np.ascontiguousarray(terms, np.int32)
    nq = len(offsets) - 1
    scores = np.empty((nq, k), np.float32)
    ids = np.empty((nq, k), np.int32)
    rc = lib().oracle_bm25s_search(_p(index.offsets), index.vocab, _p(index.docs), _p(impacts), index.num_docs,
                                   _p(offsets), _p(terms if terms.size else np.zeros(1, np.int32)), nq, k, threads,
                                   _p(scores), _p(ids))
    if rc != 0:
        raise ValueError('oracle_bm25s_search rejected its input')
    return scores, ids


def weights(offsets, doc_len, k1, b):
    offsets = np.ascontiguousarray(offsets, np.int64)
    doc_len = np.ascontiguousarray(doc_len, np.int32)
    idf = np.empty(len(offsets) - 1, np.float32)
    norm = np.empty(len(doc_len), np.float32)
    lib().oracle_bm25_weights(_p(offsets), len(idf), _p(doc_len), len(doc_len), k1, b, _p(idf), _p(norm))
    return idf, norm


def ref_search(index, offsets, terms, k, threads=4, shards=None, full=False):
    """The reference's own bm25cpu::CpuSearchSession (k1 fixed at 1.2, K <= 32)."""
    offsets = np.ascontiguousarray(offsets, np.int64)
    terms = np.ascontiguousarray(terms, np.int32)
    nq = len(offsets) - 1
    shards = min(index.num_docs, threads) if shards is None else shards
    scores = np.empty((nq, k), np.float32)
    ids = np.empty((nq, k), np.int32)
    fulls = np.empty(nq * index.num_docs, np.float32) if full else None
    rc = ref_lib().ref_cpu_search(_p(index.offsets), index.vocab, _p(index.docs), _p(index.tf), index.postings,
                                  _p(index.norm), index.num_docs, _p(index.idf), _p(offsets),
                                  _p(terms if terms.size else np.zeros(1, np.int32)), nq, k, threads, shards,
                                  _p(scores), _p(ids), _p(fulls))
    if rc != 0:
        raise ValueError('reference CPU Search rejected its input')
    if full:
        return scores, ids, fulls.reshape(nq, index.num_docs)
    return scores, ids


class RefSession:
    """The reference's own CPU Search session built once (variant 'plain' = CpuSearchSession,
    'directory' = DirectoryCpuSearchSession), for timing. Its arithmetic uses the reference's
    constant k1 = 1.2 (2.2F) whatever the index was built with."""

    def __init__(self, index, k, threads, shards=None, variant='plain', max_queries=8, max_words=24):
        shards = threads if shards is None else shards
        self._h = ref_lib().ref_cpu_create(_p(index.offsets), index.vocab, _p(index.docs), _p(index.tf),
                                           index.postings, _p(index.norm), index.num_docs, _p(index.idf),
                                           0 if variant == 'plain' else 1, threads, shards, k, max_queries,
                                           max_words)
        if not self._h:
            raise ValueError('reference CPU session rejected its configuration')
        self.k = k

    def search(self, offsets, terms):
        offsets = np.ascontiguousarray(offsets, np.int64)
        terms = np.ascontiguousarray(terms, np.int32)
        nq = len(offsets) - 1
        scores = np.empty((nq, self.k), np.float32)
        ids = np.empty((nq, self.k), np.int32)
        if ref_lib().ref_cpu_run(self._h, _p(offsets), _p(terms if terms.size else np.zeros(1, np.int32)), nq,
                                 _p(scores), _p(ids)) != 0:
            raise ValueError('reference CPU Search rejected its input')
        return scores, ids

    def close(self):
        if self._h:
            ref_lib().ref_cpu_destroy(self._h)
            self._h = None
