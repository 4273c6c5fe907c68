"""GPU parity at the full size of the 10M-doc workloads, bit for bit against the CPU oracle
(doc IDs and FP32 score bits), through the C ABI:

* the 10,000,000-doc synthetic corpus (vocabulary 1,000,000, Zipf 1.07, mean length 56,
  k1 1.5): Q8 x 3-term K10 Searches, the bench's two batches plus four more seeds;
* learned thresholds on the same corpus: a model trained on plain-path outputs, evaluated on
  held-out Q8 and Q64 batches of 2-12-term queries at K10 and K100;
* the serving loop with that index resident: Poisson arrivals at 10K and 40K queries/s,
  1-32 terms, K mixed 10/100/1000 (70/25/5%), plain and learned requests, every request's
  rows checked.

The reference checks its 10M-doc runs the same way: final IDs byte-identical to the expected
IDs and CPU scores identical (standalone/final_support.py:245-276).
"""

import numpy as np
import pytest

import bm25mi
from bm25mi import learned
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def assert_same(got_scores, got_ids, want_scores, want_ids, label=''):
    got_scores, want_scores = np.asarray(got_scores, np.float32), np.asarray(want_scores, np.float32)
    assert np.array_equal(np.asarray(got_ids), np.asarray(want_ids)), f'{label}: doc IDs differ'
    assert np.array_equal(got_scores.view(np.uint32), want_scores.view(np.uint32)), f'{label}: score bits differ'


def oracle_rows(index, queries, ks, chunk=64):
    """Oracle rows of many queries (per-query K), concatenated in query order."""
    out_s, out_i = [], []
    for b in range(0, len(queries), chunk):
        s, i = orc.search(index, *bm25mi.pack(queries[b:b + chunk]), np.asarray(ks[b:b + chunk], np.int32),
                          shards=16, threads=16, chunk=16)
        out_s.append(s.ravel())
        out_i.append(i.ravel())
    return np.concatenate(out_s), np.concatenate(out_i)


@pytest.fixture(scope='module')
def corpus10m():
    corpus = bm25mi.generate('synthetic10m')
    index = corpus.index()
    offsets = corpus.offsets.copy()
    del corpus
    print(f'synthetic10m: {index.num_docs} docs, {index.postings} postings', flush=True)
    return index, offsets


def test_synthetic10m_q8x3_k10(corpus10m, need_gpu):
    index, offsets = corpus10m
    with bm25mi.Session(index, max_queries=8, max_words=24, max_k=10) as s:
        # Seeds 1 and 2 are the bench's A/B batches.
        for seed in (1, 2, 3, 4, 5, 6):
            queries = bm25mi.draw_queries(offsets, 8, 3, seed=seed)
            want_s, want_i = orc.search(index, *bm25mi.pack(queries), 10, shards=16, threads=16, chunk=8)
            for timed in (False, True):
                got_s, got_i, _ = s.search(queries, 10, timed=timed)
                assert_same(got_s, got_i, want_s, want_i, f'seed {seed} timed={timed}')


@pytest.fixture(scope='module')
def trained(corpus10m):
    """A threshold model trained on exact plain-path K-th scores of seeded 2-12-term queries."""
    index, offsets = corpus10m
    s = bm25mi.Session(index, max_queries=64, max_words=64 * 12, max_k=100)
    qs, ks = learned.make_queries(offsets, 4096, seed=100)
    model = learned.ThresholdModel(seed=0).fit(learned.collect(s, qs, ks, batch=64), epochs=40, batch=256,
                                               lr=2e-3, seed=0)
    vq, vk = learned.make_queries(offsets, 1024, seed=200)
    model.calibrate(learned.collect(s, vq, vk, batch=64), fallback_rate=0.01)
    model.upload(s)
    yield s, model
    s.close()


@pytest.mark.parametrize('nq', [8, 64])
@pytest.mark.parametrize('k', [10, 100])
def test_learned_thresholds_synthetic10m(corpus10m, trained, nq, k):
    """Held-out batches of 2-12-term queries: the learned Search (thresholds predicted on the
    GPU inside the Search), the threshold path with the same model's host-side predictions, and
    the plain path all equal the oracle."""
    index, offsets = corpus10m
    s, model = trained
    for seed in (0, 1):
        queries = learned.make_queries(offsets, nq, seed=900 + 10 * nq + k + seed, k_choices=(k,))[0]
        want_s, want_i = orc.search(index, *bm25mi.pack(queries), k, shards=16, threads=16, chunk=16)
        got_s, got_i, fb = s.search(queries, k, learned=True)
        assert_same(got_s, got_i, want_s, want_i, f'learned (device) Q{nq} K{k} seed {seed}')
        lower, upper = s.bounds(queries, k)
        thr = model.thresholds(s.features(queries, k), lower, upper)
        got_s, got_i, fb = s.search(queries, k, thresholds=thr)
        assert_same(got_s, got_i, want_s, want_i, f'learned (host) Q{nq} K{k} seed {seed}')
        plain_s, plain_i, _ = s.search(queries, k)
        assert_same(plain_s, plain_i, want_s, want_i, f'plain Q{nq} K{k} seed {seed}')


def test_learned_headline_batches_synthetic10m(corpus10m, trained):
    """The headline shape (Q8 x 3 terms, K10) through the learned Search: results equal the
    oracle and far fewer postings are streamed than on the plain path."""
    index, offsets = corpus10m
    s, _ = trained
    for seed in (1, 2, 3):
        queries = bm25mi.draw_queries(offsets, 8, 3, seed=seed)
        want_s, want_i = orc.search(index, *bm25mi.pack(queries), 10, shards=16, threads=16, chunk=8)
        got_s, got_i, _ = s.search(queries, 10, learned=True, timed=True)
        assert_same(got_s, got_i, want_s, want_i, f'learned headline seed {seed}')
        st = s.stats()
        assert st['postings_scanned'] < 0.5 * st['postings_total']


@pytest.mark.parametrize('qps', [10000.0, 40000.0])
def test_serving_synthetic10m_every_request(corpus10m, trained, qps, need_gpu):
    index, offsets = corpus10m
    rng = np.random.default_rng(int(qps))
    n = 600
    lengths = rng.integers(1, 33, size=n)
    ks = rng.choice(np.array([10, 100, 1000], np.int32), size=n, p=[0.7, 0.25, 0.05]).astype(np.int32)
    queries = bm25mi.draw_queries(offsets, n, lengths, seed=int(qps) + 3)
    want_s, want_i = oracle_rows(index, queries, ks)
    ends = np.cumsum(ks) - 1
    kth = want_s[ends]
    # Thresholds around the true K-th score: every fifth above it (those rerun on the plain
    # path), the rest below it.
    thr = (kth * np.where(np.arange(n) % 5 == 0, 1.02, 0.9)).astype(np.float32)
    _, model = trained
    with bm25mi.Session(index, max_queries=256, max_words=256 * 32, max_k=1000, slots=4,
                        pool_entries=1 << 26) as s:
        model.upload(s)
        with bm25mi.Server(s, max_batch=256) as srv:
            for mode, t, lrn in (('plain', None, False), ('thresholds', thr, False), ('learned', None, True)):
                r = srv.run_poisson(queries, ks, qps=qps, seed=int(qps) + 7, thresholds=t, learned=lrn)
                assert_same(r['scores'], r['ids'], want_s, want_i, f'serving {mode} at {qps} QPS')
            c = srv.counters()
            assert c['queries'] >= 3 * n and c['batches'] < c['queries']
