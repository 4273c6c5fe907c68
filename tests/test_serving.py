"""Serving loop with continuous batching: every result equals the same query searched
alone (bit for bit), batches form and overlap under load, admission by candidate-pool
memory splits large-K batches, and the session's own entry points are fenced while a
server runs."""

import threading

import numpy as np
import pytest

import bm25mi
from bm25mi import _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def corpus():
    c = bm25mi.generate(bm25mi.CorpusConfig(200_000, 100_000, 1.07, 56.0, seed=21))
    return c, c.index()


def workload(corpus, n, seed):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(1, 33, size=n)
    ks = rng.choice(np.array([10, 100, 1000], np.int32), size=n, p=[0.7, 0.25, 0.05]).astype(np.int32)
    return bm25mi.draw_queries(corpus[0].offsets, n, lengths, seed=seed + 1), ks


def alone(session, queries, ks):
    rows = [session.search([q], int(k)) for q, k in zip(queries, ks)]
    return np.concatenate([r[0].ravel() for r in rows]), np.concatenate([r[1].ravel() for r in rows])


def test_results_equal_alone_under_poisson_load(corpus, need_gpu):
    _, index = corpus
    queries, ks = workload(corpus, 400, seed=1)
    with bm25mi.Session(index, max_queries=64, max_words=64 * 32, max_k=1000, slots=4) as s:
        want_s, want_i = alone(s, queries, ks)
        # The plain path itself is pinned against the oracle on a sample.
        for i in (0, 7, 99):
            os_, oi = orc.search(index, *bm25mi.pack([queries[i]]), int(ks[i]), threads=8)
            got_s, got_i, _ = s.search([queries[i]], int(ks[i]))
            assert np.array_equal(got_i, oi) and np.array_equal(got_s.view(np.uint32), os_.view(np.uint32))
        with bm25mi.Server(s, max_batch=64) as srv:
            for qps, seed in ((20000.0, 3), (1e7, 4)):
                r = srv.run_poisson(queries, ks, qps=qps, seed=seed)
                assert np.array_equal(r['ids'], want_i), qps
                assert np.array_equal(r['scores'].view(np.uint32), want_s.view(np.uint32)), qps
                assert r['p50_ms'] <= r['p95_ms'] <= r['p99_ms']
            c = srv.counters()
            assert c['queries'] == 800
            assert c['batches'] < c['queries']  # requests were batched
            assert c['max_inflight'] >= 2       # batches overlapped on the device


def test_threads_submit_and_wait_concurrently(corpus, need_gpu):
    _, index = corpus
    queries, ks = workload(corpus, 120, seed=7)
    with bm25mi.Session(index, max_queries=32, max_words=32 * 32, max_k=1000, slots=3) as s:
        want_s, want_i = alone(s, queries, ks)
        rows = np.concatenate([[0], np.cumsum(ks)])
        got_s = np.empty_like(want_s)
        got_i = np.empty_like(want_i)
        with bm25mi.Server(s) as srv:
            def worker(part):
                for i in part:
                    t = srv.submit(queries[i], int(ks[i]))
                    sc, ids, lat = srv.wait(t, int(ks[i]))
                    assert lat > 0
                    got_s[rows[i]:rows[i + 1]], got_i[rows[i]:rows[i + 1]] = sc, ids
            threads = [threading.Thread(target=worker, args=(range(j, 120, 6),)) for j in range(6)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        assert np.array_equal(got_i, want_i) and np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32))


def test_pool_admission_splits_batches_and_fences_session(corpus, need_gpu):
    _, index = corpus
    queries, _ = workload(corpus, 24, seed=11)
    ks = np.full(24, 1000, np.int32)
    with bm25mi.Session(index, max_queries=16, max_words=16 * 32, max_k=1000, slots=2) as s:
        per_query = min(s.num_docs, -(-s.num_docs // 2048) * 2000)
        # A slot holds 4 queries of K = 1000: larger batches must be split by admission.
        small = bm25mi.Session(index, max_queries=16, max_words=16 * 32, max_k=1000, slots=2,
                               pool_entries=4 * per_query)
        want_s, want_i = alone(s, queries, ks)
        with bm25mi.Server(small, max_batch=16) as srv:
            r = srv.run_poisson(queries, ks, qps=1e7, seed=5)
            assert np.array_equal(r['ids'], want_i)
            assert np.array_equal(r['scores'].view(np.uint32), want_s.view(np.uint32))
            assert srv.counters()['batches'] >= 6
            with pytest.raises(bm25mi.Bm25Error) as err:
                small.search([[1]], 10)
            assert err.value.code == _lib.ILLEGAL_OPERATION
            with pytest.raises(bm25mi.Bm25Error) as err:
                srv.submit([-1], 10)
            assert err.value.code == _lib.INVALID_PARAM
            assert _lib.load().bm25mi_destroy(small.handle) == _lib.ILLEGAL_OPERATION
        got_s, got_i, _ = small.search([queries[0]], 1000)  # usable again after the server stops
        assert np.array_equal(got_i.ravel(), want_i[:1000])
        small.close()


def test_pruned_requests_fall_back_and_match_plain(corpus, need_gpu):
    _, index = corpus
    queries, ks = workload(corpus, 96, seed=31)
    with bm25mi.Session(index, max_queries=32, max_words=32 * 32, max_k=1000, slots=3) as s:
        want_s, want_i = alone(s, queries, ks)
        ends = np.cumsum(ks) - 1
        kth = want_s[ends]
        # Every third threshold sits above the true K-th score (forcing a plain rerun), the
        # others below it.
        factor = np.where(np.arange(len(ks)) % 3 == 0, 1.05, 0.6).astype(np.float32)
        thr = (kth * factor).astype(np.float32)
        with bm25mi.Server(s, max_batch=32) as srv:
            r = srv.run_poisson(queries, ks, qps=1e6, seed=9, thresholds=thr)
            assert np.array_equal(r['ids'], want_i)
            assert np.array_equal(r['scores'].view(np.uint32), want_s.view(np.uint32))
            c = srv.counters()
            # Requests scored on the live path (threshold at least 0.3 of the FP32 sum of the
            # query's term maxima, K at most 64; the library scores the others on the plain
            # path) rerun.
            _, upper = s.bounds(queries, ks)
            live = (thr >= np.float32(0.3) * upper) & (np.asarray(ks) <= 64)
            assert c['fallbacks'] >= int(((factor > 1) & (kth > 0) & live).sum()) > 0


def test_learned_requests_match_plain(corpus, need_gpu):
    """Learned requests (thresholds predicted on the GPU in each admitted batch) mixed with plain
    ones under Poisson load: every result equals the same query searched alone."""
    from bm25mi import learned
    c, index = corpus
    queries, ks = workload(corpus, 160, seed=41)
    with bm25mi.Session(index, max_queries=32, max_words=32 * 32, max_k=1000, slots=3) as s:
        tq, tk = learned.make_queries(c.offsets, 1024, seed=5, k_choices=(10, 100))
        model = learned.ThresholdModel(seed=0).fit(learned.collect(s, tq, tk, batch=32), epochs=20, batch=128)
        model.calibrate(learned.collect(s, tq[:256], tk[:256], batch=32), fallback_rate=0.05)
        model.upload(s)
        want_s, want_i = alone(s, queries, ks)
        with bm25mi.Server(s, max_batch=32) as srv:
            r = srv.run_poisson(queries, ks, qps=1e5, seed=3, learned=True)
            assert np.array_equal(r['ids'], want_i)
            assert np.array_equal(r['scores'].view(np.uint32), want_s.view(np.uint32))
            tickets = [srv.submit(q, int(k), learned=bool(i % 2)) for i, (q, k) in enumerate(zip(queries[:40], ks[:40]))]
            rows = np.concatenate([[0], np.cumsum(ks[:40])])
            for i, t in enumerate(tickets):
                sc, ids, _ = srv.wait(t, int(ks[i]))
                assert np.array_equal(ids, want_i[rows[i]:rows[i + 1]])
                assert np.array_equal(sc.view(np.uint32), want_s[rows[i]:rows[i + 1]].view(np.uint32))
