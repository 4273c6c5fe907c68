"""GPU parity: libbm25mi Search vs the CPU oracle, bit for bit (doc IDs and FP32 score
bits), through the C ABI. Cases follow the reference's own checks: its tiny
known-answer corpora (FR/ut/cases/bm25/cpu_index_search_tests.cpp:95-134), its
selector edge cases (FR/test/ascendhost/bm25/selector_check.cpp:63-118: zero, K-1,
K, K+1 positive docs, ties), empty/unknown/duplicate/2147483646 terms, the fixed
query sets fixtures/A.txt and B.txt, and the golden vectors of the reference binary.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import bm25mi
from oracle import oracle as orc
from tests import corpora

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / 'golden'


def assert_same(got_scores, got_ids, want_scores, want_ids, label=''):
    got_scores, want_scores = np.asarray(got_scores, np.float32), np.asarray(want_scores, np.float32)
    assert np.array_equal(np.asarray(got_ids), np.asarray(want_ids)), f'{label}: doc IDs differ'
    assert np.array_equal(got_scores.view(np.uint32), want_scores.view(np.uint32)), f'{label}: score bits differ'


def takes_live_path(session, queries, k, thr):
    """Threshold queries the library scores on the live path: a threshold of at least 0.3 of
    the FP32 sum of the query's term maxima and K at most 64 (the others are scored on the
    plain path)."""
    _, upper = session.bounds(queries, k)
    ks = np.broadcast_to(np.asarray(k), np.shape(thr))
    return (thr > 0) & (thr >= np.float32(0.3) * upper) & (ks <= 64)


@pytest.fixture(scope='module')
def synth():
    corpus = bm25mi.generate('synthetic8k')
    return corpus, corpus.index()


@pytest.fixture(scope='module')
def synth_session(synth, need_gpu):
    _, index = synth
    s = bm25mi.Session(index, max_queries=64, max_words=64 * 64, max_k=1024, block_docs=8192)
    yield s
    s.close()


def test_synthetic8k_q8x3_k10(synth, synth_session):
    corpus, index = synth
    for seed in range(5):
        queries = bm25mi.draw_queries(corpus.offsets, 8, 3, seed=seed)
        offsets, terms = bm25mi.pack(queries)
        want_s, want_i = orc.search(index, offsets, terms, 10)
        got_s, got_i, _ = synth_session.search(queries, 10)
        assert_same(got_s, got_i, want_s, want_i, f'seed {seed}')


def test_poisoned_output_buffer_is_fully_written(synth, synth_session):
    """A caller-supplied output buffer pre-filled with a NaN / invalid-ID pattern (the
    reference's output poisoning, FR/test/ascendhost/bm25/search_diagnostic.cpp:151-161) holds
    exactly the oracle's rows afterwards; a buffer of the wrong size is rejected."""
    corpus, index = synth
    queries = bm25mi.draw_queries(corpus.offsets, 8, 3, seed=11)
    want_s, want_i = orc.search(index, *bm25mi.pack(queries), 10)
    for learned_path in (False, True):
        buf = np.full(2 * 80 + 8, 0x7FC0DEAD, np.int32)
        thr = None
        if learned_path:
            lower, _ = synth_session.bounds(queries, 10)
            thr = lower  # a provable lower bound of each K-th score: the threshold path, same rows
        got_s, got_i, fb = synth_session.search(queries, 10, thresholds=thr, out=buf)
        assert_same(got_s, got_i, want_s, want_i, f'poisoned buffer, thresholds={learned_path}')
        assert np.shares_memory(got_s, buf) and not fb.any()
    with pytest.raises(ValueError):
        synth_session.search(queries, 10, out=np.zeros(10, np.int32))


@pytest.mark.parametrize('k', [1, 2, 10, 32, 100, 1000, 1024])
def test_synthetic8k_all_k(synth, synth_session, k):
    corpus, index = synth
    queries = bm25mi.draw_queries(corpus.offsets, 16, 4, seed=100 + k)
    offsets, terms = bm25mi.pack(queries)
    want_s, want_i = orc.search(index, offsets, terms, k)
    got_s, got_i, _ = synth_session.search(queries, k)
    assert_same(got_s, got_i, want_s, want_i, f'k={k}')


def test_uneven_batch(synth, synth_session):
    corpus, index = synth
    rng = np.random.default_rng(9)
    for trial in range(4):
        nq = int(rng.integers(1, 65))
        lengths = rng.integers(0, 33, size=nq)
        queries = bm25mi.draw_queries(corpus.offsets, nq, lengths, seed=trial)
        ks = rng.choice([1, 10, 100, 1000], size=nq).astype(np.int32)
        offsets, terms = bm25mi.pack(queries)
        want_s, want_i = orc.search(index, offsets, terms, ks)
        got_s, got_i, _ = synth_session.search(queries, ks)
        assert_same(got_s, got_i, want_s, want_i, f'trial {trial}')


def test_edge_queries(synth, synth_session):
    _, index = synth
    queries = [[], [2147483646], [index.vocab], [0, 0], [4999, 4999, 4999], [3, 2147483646, 3],
               list(range(64)), [4998], [0]]
    offsets, terms = bm25mi.pack(queries)
    for k in (1, 10, 1000):
        want_s, want_i = orc.search(index, offsets, terms, k)
        got_s, got_i, _ = synth_session.search(queries, k)
        assert_same(got_s, got_i, want_s, want_i, f'edge k={k}')
    # Empty query: IDs 0..K-1 with score 0 (large_search.cpp:184-187).
    s, i, _ = synth_session.search([[]], 10)
    assert i[0].tolist() == list(range(10)) and not s.any()


def test_rejections_keep_session_usable(synth, synth_session):
    bad = [([[-1]], 10), ([[1] * 65], 10), ([[1]] * 65, 10), ([[1]], 0), ([[1]], 1025)]
    for queries, k in bad:
        with pytest.raises(bm25mi.Bm25Error) as err:
            synth_session.search(queries, k)
        assert err.value.code == bm25mi._lib.INVALID_PARAM
    test_synthetic8k_q8x3_k10(synth, synth_session)


@pytest.mark.parametrize('n', [1, 6, 129, 2048])
def test_reference_kat_tiny_corpora(n, need_gpu):
    docs = corpora.tiny_documents(n)
    index = corpora.index_from_documents(docs, 7)
    queries_sets = ([[0, 0], [], [6], [7], [4, 5, 0], [2147483646, 2], [0, 1, 2, 3]], [[4], [0], [], [7], [2, 2, 1]],
                    [[], [], []])
    ks = sorted({1, min(n, 3), min(n, 10), min(n, 32)})
    for block in (2048, 8192):
        with bm25mi.Session(index, max_queries=7, max_words=128, max_k=max(ks), block_docs=block) as s:
            for queries in queries_sets:
                expected = corpora.independent_scores(docs, queries, 7)
                for k in ks:
                    got_s, got_i, _ = s.search(queries, k)
                    for q in range(len(queries)):
                        top = corpora.strict_top(expected[q], k)
                        assert_same(got_s[q], got_i[q], expected[q, top], top, f'n={n} k={k} q={q}')


def test_selector_edge_counts(need_gpu):
    """Positive-doc counts around K and massive ties (selector_check.cpp:63-118)."""
    n = 8192 * 3 + 77
    rng = np.random.default_rng(5)
    for positives in (0, 1, 9, 10, 11, 63, 64, 65, 1023, 1024, 1025, 4097, n):
        docs_t0 = np.sort(rng.choice(n, size=positives, replace=False)).astype(np.int32)
        # Term 0: the chosen docs with tf 1 and equal norms -> all tied; term 1: ties at two levels.
        docs_t1 = np.arange(0, n, 7, dtype=np.int32)
        offsets = np.array([0, len(docs_t0), len(docs_t0) + len(docs_t1)], np.int64)
        docs = np.concatenate([docs_t0, docs_t1])
        tf = np.concatenate([np.ones(len(docs_t0)), 1 + (docs_t1 % 2)]).astype(np.float32)
        norm = np.ones(n, np.float32)
        idf = np.array([2.0, 0.5], np.float32)
        index = bm25mi.Index(offsets, docs, tf, norm, idf, k1=1.2)
        queries = [[0], [0, 1], [1, 0], [1]]
        offs, terms = bm25mi.pack(queries)
        with bm25mi.Session(index, max_queries=4, max_words=8, max_k=1024, block_docs=4096) as s:
            for k in (1, 10, 32, 1000):
                want_s, want_i = orc.search(index, offs, terms, k)
                got_s, got_i, _ = s.search(queries, k)
                assert_same(got_s, got_i, want_s, want_i, f'positives={positives} k={k}')


def test_golden_vectors_of_reference_binary(need_gpu):
    record = json.loads((GOLDEN / 'reference_cpu_golden.json').read_text())
    sessions = {}
    try:
        for case in record['cases']:
            if case['corpus'] not in sessions:
                index = corpora.golden_index(case['corpus'])
                sessions[case['corpus']] = bm25mi.Session(index, max_queries=16, max_words=256, max_k=32)
            got_s, got_i, _ = sessions[case['corpus']].search(case['queries'], case['k'])
            assert got_i.tolist() == case['ids'], case['name']
            assert got_s.view(np.uint32).tolist() == case['score_bits'], case['name']
    finally:
        for s in sessions.values():
            s.close()


@pytest.mark.parametrize('block', [2048, 4096, 8192])
@pytest.mark.parametrize('graph', [True, False])
def test_million_docs_block_sizes(block, graph, need_gpu, million):
    corpus, index, queries = million
    offsets, terms = bm25mi.pack(queries)
    want_s, want_i = orc.search(index, offsets, terms, 10, threads=8)
    with bm25mi.Session(index, max_queries=16, max_words=512, max_k=100, block_docs=block,
                        use_graph=graph) as s:
        got_s, got_i, _ = s.search(queries, 10)
        assert_same(got_s, got_i, want_s, want_i, f'block={block} graph={graph}')
        ks = np.array([1, 10, 100, 37] * 4, np.int32)
        want_s, want_i = orc.search(index, offsets, terms, ks, threads=8)
        got_s, got_i, _ = s.search(queries, ks)
        assert_same(got_s, got_i, want_s, want_i, f'uneven block={block}')


@pytest.mark.parametrize('block', [2048, 4096, 8192])
def test_million_docs_pruned(block, need_gpu, million):
    """Threshold path through the planning kernel and the live list: uniform random terms
    (mostly short lists) mixed into df-weighted queries, plain queries in the same batch, graph
    replay and direct launches, thresholds at and below the K-th score (no rerun) and above it
    (every such query reruns on the plain path); results equal the oracle bit for bit."""
    corpus, index, queries = million
    rng = np.random.default_rng(block)
    extra = [rng.integers(0, index.vocab, size=3).tolist() + q[:2] for q in queries[:8]]
    qs = queries + extra
    offsets, terms = bm25mi.pack(qs)
    ks = np.array([10, 100, 1, 37] * 6, np.int32)
    want_s, want_i = orc.search(index, offsets, terms, ks, threads=8)
    kth = want_s[np.cumsum(ks) - 1].astype(np.float32)
    # use_graph: untimed calls replay the captured graphs, timed calls launch directly.
    with bm25mi.Session(index, max_queries=32, max_words=1024, max_k=100, block_docs=block,
                        use_graph=True) as s:
        for factor in (0.5, 0.9, 1.0):
            thr = (kth * np.float32(factor)).astype(np.float32)
            thr[::3] = 0.0
            for timed in (False, True):
                got_s, got_i, fb = s.search((offsets, terms), ks, thresholds=thr, timed=timed)
                assert_same(got_s, got_i, want_s, want_i, f'pruned block={block} f={factor} timed={timed}')
                assert not fb.any()
            if factor == 1.0:
                st = s.stats()
                assert st['blocks_skipped'] > 0 and st['postings_scanned'] < st['postings_total']
        thr = (kth * np.float32(1.01)).astype(np.float32)
        got_s, got_i, fb = s.search((offsets, terms), ks, thresholds=thr)
        assert_same(got_s, got_i, want_s, want_i, f'rerun block={block}')
        live = takes_live_path(s, (offsets, terms), ks, thr)
        assert live.any() and (fb == (live & (kth > 0))).all()


@pytest.fixture(scope='module')
def million():
    config = bm25mi.CorpusConfig(1_000_000, 1_000_000, 1.07, 56.0, seed=3)
    corpus = bm25mi.generate(config)
    index = corpus.index()
    queries = bm25mi.draw_queries(corpus.offsets, 8, 3, seed=4) + bm25mi.draw_queries(corpus.offsets, 8, 12, seed=5)
    return corpus, index, queries


@pytest.mark.parametrize('graph', [True, False])
def test_all_threshold_batch_launch_modes(graph, need_gpu, million):
    """A batch whose queries all take the live path (planning and live kernels only): replayed
    from its captured graph, then launched directly, several times in a row; results equal the
    oracle bit for bit and no query reruns."""
    corpus, index, queries = million
    offsets, terms = bm25mi.pack(queries)
    want_s, want_i = orc.search(index, offsets, terms, 10, threads=8)
    kth = want_s[:, -1].astype(np.float32)
    with bm25mi.Session(index, max_queries=16, max_words=512, max_k=10, use_graph=graph) as s:
        live = takes_live_path(s, (offsets, terms), 10, kth)
        keep = np.flatnonzero(live & (kth > 0))
        assert len(keep) >= 4
        sub = [queries[i] for i in keep]
        for rep in range(3):
            got_s, got_i, fb = s.search(sub, 10, thresholds=kth[keep])
            assert_same(got_s, got_i, want_s[keep], want_i[keep], f'graph={graph} rep={rep}')
            assert not fb.any()
        st = s.stats()
        assert st['blocks_skipped'] > 0 and st['postings_scanned'] < st['postings_total']


def test_pruned_path_equals_plain(synth, synth_session):
    corpus, index = synth
    queries = bm25mi.draw_queries(corpus.offsets, 16, 3, seed=77)
    offsets, terms = bm25mi.pack(queries)
    for k in (10, 100, 1000):
        want_s, want_i = orc.search(index, offsets, terms, k)
        kth = want_s[:, -1].astype(np.float32)
        for factor in (0.5, 1.0, 1.01):
            thr = (kth * np.float32(factor)).astype(np.float32)
            got_s, got_i, fb = synth_session.search(queries, k, thresholds=thr)
            assert_same(got_s, got_i, want_s, want_i, f'pruned k={k} f={factor}')
            if factor > 1.0:
                # Above the K-th score fewer than K docs survive: every positive-K query that
                # takes the live path reruns.
                live = takes_live_path(synth_session, queries, k, thr)
                assert (fb == (live & (kth > 0))).all()
            elif factor == 1.0:
                # At the K-th score only the top K (and ties) survive: no rerun.
                assert not fb.any()
            # Far below it, more docs may survive than the query's pool holds (one 8192-doc
            # block here: BlockKeepCap(K) entries); such a query reruns, with the same results.
        st = synth_session.stats()
        assert st['postings_total'] > 0


def test_timed_path_reports_kernel_time(synth, synth_session):
    corpus, index = synth
    queries = bm25mi.draw_queries(corpus.offsets, 8, 3, seed=1)
    offsets, terms = bm25mi.pack(queries)
    want_s, want_i = orc.search(index, offsets, terms, 10)
    got_s, got_i, _ = synth_session.search(queries, 10, timed=True)
    assert_same(got_s, got_i, want_s, want_i, 'timed')
    st = synth_session.stats()
    assert st['score_kernel_ms'] > 0 and st['device_ms'] >= st['score_kernel_ms']
    assert st['postings_in_range'] == st['postings_total'] <= st['postings_scanned']


def test_search_session_golden_vectors_and_rejects(need_gpu):
    """The raw-ID Search surface (standalone/run_bm25s_index_search.py:15-76) on the golden
    vectors the reference's CPU Search wrote, and its request rejections, which leave the
    session usable."""
    record = json.loads((GOLDEN / 'reference_cpu_golden.json').read_text())
    by_corpus = {}
    for case in record['cases']:
        by_corpus.setdefault(case['corpus'], []).append(case)
    for name, cases in by_corpus.items():
        index = corpora.golden_index(name)
        for k in sorted({c['k'] for c in cases}):
            with bm25mi.SearchSession(index, k, max_queries=16, max_words=256) as ss:
                for case in (c for c in cases if c['k'] == k):
                    scores, ids = ss.search([list(map(int, q)) for q in case['queries']])
                    assert ids.tolist() == case['ids'], case['name']
                    assert scores.view(np.uint32).tolist() == case['score_bits'], case['name']
                for bad in ([], [[1]] * 17, [[True]], [[-1]], [[2**31 - 1]], [(1,)], [[1] * 257]):
                    with pytest.raises(ValueError):
                        ss.search(bad)
                case = next(c for c in cases if c['k'] == k)
                scores, ids = ss.search([list(map(int, q)) for q in case['queries']])
                assert ids.tolist() == case['ids']
    with pytest.raises(ValueError):
        bm25mi.SearchSession(corpora.golden_index('tiny129'), True)
    with pytest.raises(ValueError):
        bm25mi.SearchSession(corpora.golden_index('tiny129'), 130)


def test_facade_device_failure_state_machine(need_gpu):
    """sdk_facade_tests.cpp:38-62: an Init failure keeps its device code and a retry succeeds; a
    device failure during Search returns INNER_ERROR with its code kept and the output
    untouched, after which Search is an ILLEGAL_OPERATION; a failed Close keeps the state and
    can be retried; Close is idempotent. The failures are injected from the host before any
    device work of the failing call."""
    lib = bm25mi._lib.load()
    index = bm25mi.Index([0, 1], [0], [1], [1], [2])
    cfg = bm25mi.Bm25Config(queryCount=1, maxQueryWords=1, topK=1)
    model = bm25mi.AscendIndexBm25()
    ok, inner, illegal = (bm25mi.search.APP_ERR_OK, bm25mi.search.APP_ERR_INNER_ERROR,
                          bm25mi.search.APP_ERR_ILLEGAL_OPERATION)
    try:
        assert lib.bm25mi_inject_failure(b'create', 42, b'injected Init failure') == 0
        assert model.Init(index, cfg) == inner and model.GetLastError().aclCode == 42
        assert model.Init(index, cfg) == ok
        assert model.Init(index, cfg) == illegal
        out = bm25mi.Bm25SearchResult(np.array([9.0], np.float32), np.array([9], np.int64))
        assert model.Search([[-1]], out) == bm25mi.search.APP_ERR_INVALID_PARAM
        assert out.scores[0] == 9 and out.labels[0] == 9
        assert model.Search([[0]], out) == ok and out.labels.tolist() == [0]
        good = out.scores.copy()
        assert lib.bm25mi_inject_failure(b'search', 161001, b'injected device failure') == 0
        assert model.Search([[0]], out) == inner
        err = model.GetLastError()
        assert err.aclCode == 161001 and err.message == 'injected device failure'
        assert np.array_equal(out.scores, good) and out.labels.tolist() == [0]
        assert model.Search([[0]], out) == illegal
        assert lib.bm25mi_inject_failure(b'close', 7, b'injected Close failure') == 0
        assert model.Close() == inner and model.GetLastError().aclCode == 7
        assert model.Close() == ok and model.Close() == ok
        assert model.Search([[0]], out) == illegal
    finally:
        lib.bm25mi_inject_failure(None, 0, None)


def test_result_checks_reject_corrupt_rows(synth, need_gpu):
    """The result checks of Session::Receive (large_search.cpp:64-85) reject rows out of strict
    order and doc IDs out of range as device errors, which fail the session."""
    corpus, index = synth
    lib = bm25mi._lib.load()
    queries = bm25mi.draw_queries(corpus.offsets, 2, 3, seed=3)
    for k, message in ((10, 'Strict global ranking failure'), (1, 'Invalid real-document result')):
        with bm25mi.Session(index, max_queries=2, max_words=8, max_k=10) as s:
            s.search(queries, k)
            assert lib.bm25mi_inject_failure(b'result', 0, None) == 0
            with pytest.raises(bm25mi.Bm25Error) as err:
                s.search(queries, k)
            assert err.value.code == bm25mi._lib.INNER and message in err.value.message
            with pytest.raises(bm25mi.Bm25Error) as err:
                s.search(queries, k)
            assert err.value.code == bm25mi._lib.ILLEGAL_OPERATION


@pytest.mark.parametrize('block', [2048, 8192])
def test_final_kernel_beyond_sort_capacity(block, need_gpu):
    """Tie mass that leaves more than 16,384 candidates at or above a query's final bound: the
    final kernel selects the K-th key by bisection over the whole pool instead of sorting it in
    LDS. 200,000 docs share one score (term 0) and every third one a second, higher one (term 1
    added); merge_check.h:23-114 checks all-1.0f ties across lists the same way."""
    n = 200_000
    docs0 = np.arange(n, dtype=np.int32)
    docs1 = np.arange(0, n, 3, dtype=np.int32)
    offsets = np.array([0, n, n + len(docs1)], np.int64)
    index = bm25mi.Index(offsets, np.concatenate([docs0, docs1]), np.ones(n + len(docs1), np.float32),
                         np.ones(n, np.float32), np.array([1.0, 0.25], np.float32), k1=1.2)
    queries = [[0], [0, 1], [1, 0], [0, 0]]
    offs, terms = bm25mi.pack(queries)
    with bm25mi.Session(index, max_queries=4, max_words=8, max_k=1024, block_docs=block) as s:
        for k in (1000, 1024):
            want_s, want_i = orc.search(index, offs, terms, k)
            got_s, got_i, _ = s.search(queries, k)
            assert_same(got_s, got_i, want_s, want_i, f'ties block={block} k={k}')
            # Query [0] alone: every block keeps at least K of its tied docs, so its pool holds
            # more candidates than the final kernel's LDS sort (16,384).
            got_s, got_i, _ = s.search([[0]], k)
            assert_same(got_s, got_i, want_s[:1], want_i[:1], f'ties [0] block={block} k={k}')
            assert s.stats()['candidates'] > 16384
