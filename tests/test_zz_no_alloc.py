"""bm25mi_search_batch_noalloc (the Search whose device scratch lives in the library's static
array) against bm25mi_search_batch on the same session, bit for bit, on two calls in a row, the
second with a smaller batch; a batch whose candidate pools exceed the array is rejected and the
session stays usable. Calls go through the C ABI."""

import ctypes as C

import numpy as np
import pytest

import bm25mi
from bm25mi import _lib

pytestmark = pytest.mark.gpu


def call(session, entry, queries, ks):
    offsets, terms = bm25mi.pack(queries)
    ks = np.ascontiguousarray(ks, np.int32)
    rows = int(ks.sum())
    scores = np.full(rows, -1.0, np.float32)
    ids = np.full(rows, -1, np.int32)
    fn = getattr(_lib.load(), entry)
    rc = fn(session.handle, _lib.ptr(offsets), _lib.ptr(terms), len(ks), _lib.ptr(ks), _lib.ptr(scores),
            _lib.ptr(ids))
    return rc, scores, ids


@pytest.fixture(scope='module')
def million():
    corpus = bm25mi.generate(bm25mi.CorpusConfig(1_000_000, 1_000_000, 1.07, 56.0, seed=3))
    return corpus.offsets.copy(), corpus.index()


def test_noalloc_search_equals_batch_search(million, need_gpu):
    offsets, index = million
    rng = np.random.default_rng(12)
    big = bm25mi.draw_queries(offsets, 12, rng.integers(1, 13, size=12), seed=31)
    big_k = rng.choice(np.array([1, 10, 37, 100], np.int32), size=12).astype(np.int32)
    small = bm25mi.draw_queries(offsets, 3, 3, seed=32) + [[]]
    small_k = np.array([10, 5, 1, 10], np.int32)
    with bm25mi.Session(index, max_queries=32, max_words=32 * 16, max_k=1024) as s:
        for rep in range(2):
            for queries, ks in ((big, big_k), (small, small_k)):
                rc, want_s, want_i = call(s, 'bm25mi_search_batch', queries, ks)
                assert rc == _lib.OK
                rc, got_s, got_i = call(s, 'bm25mi_search_batch_noalloc', queries, ks)
                assert rc == _lib.OK, _lib.last_error(s.handle)
                assert np.array_equal(got_i, want_i), f'rep {rep} Q{len(ks)}: doc IDs differ'
                assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32)), \
                    f'rep {rep} Q{len(ks)}: score bits differ'
        # 32 queries of K = 1024 need 32 x 1M pool entries (256 MB): more than the array holds.
        huge = bm25mi.draw_queries(offsets, 32, 2, seed=33)
        rc, got_s, got_i = call(s, 'bm25mi_search_batch_noalloc', huge, np.full(32, 1024, np.int32))
        assert rc == _lib.INVALID_PARAM and (got_i == -1).all()
        assert 'no-allocation scratch' in _lib.last_error(s.handle)[0]
        rc, got_s, got_i = call(s, 'bm25mi_search_batch_noalloc', small, small_k)
        rc2, want_s, want_i = call(s, 'bm25mi_search_batch', small, small_k)
        assert rc == rc2 == _lib.OK and np.array_equal(got_i, want_i)
        assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32))
