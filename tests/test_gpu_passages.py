"""GPU parity of uneven batches at full size: the 8,841,823-passage synthetic corpus
(vocabulary 2,600,000, Zipf 1.05, mean length 56, k1 1.5), batches of 1, 8, 64 and 256 queries
with geometric 1-32 terms (mean 6) and K drawn per query from {10, 100, 1000}, each batch one
persistent launch. Every batch is compared with the CPU oracle bit for bit (doc IDs and FP32
score bits), at the default 2048-doc tiles and at 4096- and 8192-doc tiles, and with candidate
capacities (keep_cap) of 1024 and 8192 per task."""

import numpy as np
import pytest

import bm25mi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def passages():
    corpus = bm25mi.generate('passages8p8m')
    index = corpus.index()
    offsets = corpus.offsets.copy()
    del corpus
    print(f'passages8p8m: {index.num_docs} docs, {index.postings} postings', flush=True)
    return index, offsets


def batch(offsets, nq, j):
    lengths = bm25mi.corpus.geometric_lengths(nq, 6, 1, 32, seed=1000 * nq + j)
    ks = np.random.default_rng(77 + 1000 * nq + j).choice(np.array([10, 100, 1000], np.int32),
                                                           size=nq).astype(np.int32)
    return bm25mi.draw_queries(offsets, nq, lengths, seed=2000 * nq + j), ks


@pytest.mark.parametrize('block', [2048, 4096, 8192])
def test_uneven_batches_passages8p8m(passages, block, need_gpu):
    index, offsets = passages
    sizes = (1, 8, 64, 256) if block == 2048 else (64,)
    with bm25mi.Session(index, max_queries=256, max_words=256 * 32, max_k=1000, block_docs=block) as s:
        for nq in sizes:
            for j in (0, 1):
                queries, ks = batch(offsets, nq, j)
                want_s, want_i = orc.search(index, *bm25mi.pack(queries), ks, shards=16, threads=16, chunk=32)
                got_s, got_i, _ = s.search(queries, ks)
                assert np.array_equal(got_i, want_i), f'Q{nq} batch {j} block {block}: doc IDs differ'
                assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32)), \
                    f'Q{nq} batch {j} block {block}: score bits differ'


@pytest.mark.parametrize('cap', [1024, 8192])
def test_keep_caps_passages8p8m(passages, cap, need_gpu):
    """A task's candidate capacity (the reference's capacity variants) changes no result."""
    index, offsets = passages
    with bm25mi.Session(index, max_queries=64, max_words=64 * 32, max_k=1000, keep_cap=cap) as s:
        queries, ks = batch(offsets, 64, 0)
        want_s, want_i = orc.search(index, *bm25mi.pack(queries), ks, shards=16, threads=16, chunk=32)
        got_s, got_i, _ = s.search(queries, ks)
        assert np.array_equal(got_i, want_i), f'keep_cap {cap}: doc IDs differ'
        assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32)), f'keep_cap {cap}: score bits differ'
