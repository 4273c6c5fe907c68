"""Pins the CPU oracle (oracle/bm25_oracle.c) before anything is checked against it.

* the reference's own known-answer corpora and independent per-document FP32
  oracle (FR/ut/cases/bm25/cpu_index_search_tests.cpp:22-134), restated here;
* the reference's compiled CPU Search and unit tests (oracle/_ref, built from the
  reference sources in this container; skipped where absent);
* golden vectors that the reference binary produced (tests/golden/);
* brute-force full sorts, and planted mistakes that these checks must catch.
"""

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import bm25mi
from oracle import oracle as orc
from tests import corpora

GOLDEN = Path(__file__).resolve().parent / 'golden'
QUERY_A = [[0, 0], [], [6], [7], [4, 5, 0], [2147483646, 2], [0, 1, 2, 3]]
QUERY_B = [[4], [0], [], [7], [2, 2, 1]]
QUERY_EMPTY = [[], [], []]


@pytest.mark.parametrize('n', [1, 6, 129, 2048])
def test_reference_kat_tiny_corpora(n):
    docs = corpora.tiny_documents(n)
    index = corpora.index_from_documents(docs, 7)
    for queries in (QUERY_A, QUERY_B, QUERY_A, QUERY_EMPTY, QUERY_B):
        expected = corpora.independent_scores(docs, queries, 7)
        offsets, terms = bm25mi.pack(queries)
        for k in sorted({1, min(n, 3), min(n, 10), min(n, 32)}):
            for threads in (1, 4, 8):
                scores, ids, full = orc.search(index, offsets, terms, k, shards=min(n, threads), threads=threads,
                                               full=True)
                assert np.array_equal(full.view(np.uint32), expected.view(np.uint32))
                for q in range(len(queries)):
                    top = corpora.strict_top(expected[q], k)
                    assert np.array_equal(ids[q], top)
                    assert np.array_equal(scores[q].view(np.uint32), expected[q, top].view(np.uint32))


def test_bruteforce_synthetic8k_uneven_k():
    corpus = bm25mi.generate('synthetic8k')
    index = corpus.index()
    queries = bm25mi.draw_queries(corpus.offsets, 12, 3, seed=5)
    queries += [[], [4999, 4999], [10**6, 3], [0, 1, 2, 3, 4, 5, 6, 7]]
    ks = [1, 10, 100, 1000, 8192, 3, 10, 57, 1024, 2, 10, 500, 10, 10, 10, 999]
    offsets, terms = bm25mi.pack(queries)
    scores, ids, full = orc.search(index, offsets, terms, ks, shards=5, threads=4, full=True)
    row = 0
    for q, k in enumerate(ks):
        top = corpora.strict_top(full[q], k)
        assert np.array_equal(ids[row:row + k], top)
        assert np.array_equal(scores[row:row + k], full[q, top])
        row += k
    # Dense rows equal an independent numpy recomputation in query-term order.
    impacts = index.impacts()
    for q, query in enumerate(queries):
        dense = np.zeros(index.num_docs, np.float32)
        for t in query:
            if t < index.vocab:
                a, b = index.offsets[t], index.offsets[t + 1]
                dense[index.docs[a:b]] = dense[index.docs[a:b]] + impacts[a:b]
        assert np.array_equal(dense.view(np.uint32), full[q].view(np.uint32))


def test_weights_match_reference_formulas():
    corpus = bm25mi.generate('synthetic8k')
    for k1, b in ((1.2, 0.75), (1.5, 0.75), (0.9, 0.4)):
        idf, norm = orc.weights(corpus.offsets, corpus.doc_len, k1, b)
        n = corpus.num_docs
        df = corpus.df().astype(np.float64)
        want_idf = np.log1p((n - df + 0.5) / (df + 0.5)).astype(np.float32)
        avg = corpus.doc_len.astype(np.float64).sum() / n
        want_norm = (k1 * (1 - b + b * corpus.doc_len.astype(np.float64) / avg)).astype(np.float32)
        assert np.array_equal(idf.view(np.uint32), want_idf.view(np.uint32))
        assert np.array_equal(norm.view(np.uint32), want_norm.view(np.uint32))


def test_contribution_closed_form():
    lib = orc.lib()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        idf, tf, norm = np.float32(rng.uniform(0, 15)), np.float32(rng.integers(1, 50)), np.float32(
            rng.uniform(0.2, 4))
        kp1 = np.float32(2.2)
        want = np.float32(np.float32(idf * np.float32(tf * kp1)) / np.float32(tf + norm))
        got = np.float32(lib.oracle_contribution(float(idf), float(tf), float(norm), float(kp1)))
        assert got.view(np.uint32) == want.view(np.uint32)
    # k1 + 1 = 2.2 at the reference's k1.
    assert np.float32(orc.kp1(1.2)) == np.float32(2.2)


def test_planted_mistakes_are_caught():
    """Dense rows are pinned bit for bit: a reversed term order, a wrong k1 and a
    dropped repeated term each change them."""
    corpus = bm25mi.generate('synthetic8k')
    index = corpus.index()
    impacts = index.impacts()
    queries = bm25mi.draw_queries(corpus.offsets, 8, 4, seed=21)
    queries[0] = queries[0] + [queries[0][0]]

    def dense(qs):
        out = np.zeros((len(qs), index.num_docs), np.float32)
        for q, query in enumerate(qs):
            for t in query:
                a, b = index.offsets[t], index.offsets[t + 1]
                out[q, index.docs[a:b]] = out[q, index.docs[a:b]] + impacts[a:b]
        return out

    expected = dense(queries)
    offsets, terms = bm25mi.pack(queries)
    full = orc.search(index, offsets, terms, 3, full=True)[2]
    assert np.array_equal(full.view(np.uint32), expected.view(np.uint32))
    reversed_q = [list(reversed(q)) for q in queries]
    assert not np.array_equal(orc.search(index, *bm25mi.pack(reversed_q), 3, full=True)[2].view(np.uint32),
                              expected.view(np.uint32))
    assert not np.array_equal(orc.search(index, offsets, terms, 3, full=True, k1=1.2)[2], expected)
    dropped = [list(dict.fromkeys(q)) for q in queries]
    assert not np.array_equal(orc.search(index, *bm25mi.pack(dropped), 3, full=True)[2], expected)


def test_golden_vectors_from_reference_binary():
    """tests/golden/reference_cpu_golden.json was written by the reference's own CPU
    Search (oracle/_ref) by tests/golden/make_golden.py."""
    record = json.loads((GOLDEN / 'reference_cpu_golden.json').read_text())
    for case in record['cases']:
        index = corpora.golden_index(case['corpus'])
        offsets, terms = bm25mi.pack(case['queries'])
        scores, ids = orc.search(index, offsets, terms, case['k'], shards=case['shards'], threads=4)
        assert ids.tolist() == case['ids'], case['name']
        assert scores.view(np.uint32).tolist() == case['score_bits'], case['name']


needs_ref = pytest.mark.skipif(not orc.ref_available(), reason='oracle/_ref (reference build) not present')


@needs_ref
@pytest.mark.parametrize('name', ['cpu_index_search_tests', 'index_data_tests', 'error_contract_tests'])
def test_reference_unit_tests_pass(name):
    out = subprocess.run([str(orc.REF_DIR / name)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr


@needs_ref
def test_oracle_equals_reference_binary():
    corpus = bm25mi.generate('synthetic8k')
    index = corpus.index(k1=1.2)
    queries = bm25mi.draw_queries(corpus.offsets, 16, 3, seed=11)
    queries += [[], [5, 5, 5], [2147483646], [0, 4999, 17, 17]]
    offsets, terms = bm25mi.pack(queries)
    for k in (1, 10, 32):
        for threads in (1, 8):
            rs, ri, rf = orc.ref_search(index, offsets, terms, k, threads=threads, full=True)
            s, i, f = orc.search(index, offsets, terms, k, shards=min(threads, 8), threads=threads, full=True)
            assert np.array_equal(rf.view(np.uint32), f.view(np.uint32))
            assert np.array_equal(ri, i)
            assert np.array_equal(rs.view(np.uint32), s.view(np.uint32))


def test_bm25s_restatement_equals_oracle():
    """The bm25s-algorithm CPU baseline (oracle_bm25s_search: dense FP32 row per query,
    scatter-add of precomputed impacts in query-token order, bounded heap top-k) is pinned the
    way the reference checks bm25s (standalone/bm25s_benchmark.py:145-154): its selected scores
    equal the full-score rows at the selected IDs, its score multiset equals the exact top-K,
    and its IDs equal the strict contract's. Impacts are precomputed exactly as the oracle
    computes each contribution, so here the scores agree bit for bit."""
    corpus = bm25mi.generate('synthetic8k')
    index = corpus.index()
    impacts = index.impacts()
    queries = bm25mi.draw_queries(corpus.offsets, 10, 3, seed=21) + [[], [4999, 4999], [10**6, 3], [7, 7, 7, 1]]
    offsets, terms = bm25mi.pack(queries)
    for k in (1, 10, 100, 1000):
        want_s, want_i, full = orc.search(index, offsets, terms, k, full=True)
        got_s, got_i = orc.bm25s_search(index, impacts, offsets, terms, k, threads=3)
        assert np.array_equal(np.take_along_axis(full, got_i.astype(np.int64), axis=1).view(np.uint32),
                              got_s.view(np.uint32))
        assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32))
        assert np.array_equal(got_i, want_i)
    # A planted mistake in the impacts (one posting's impact scaled) is caught.
    bad = impacts.copy()
    hot = int(np.argmax(impacts))
    bad[hot] *= np.float32(0.5)
    got_s, got_i = orc.bm25s_search(index, bad, *bm25mi.pack([[int(np.searchsorted(index.offsets, hot, 'right') - 1)]]),
                                    10)
    want_s, want_i = orc.search(index, *bm25mi.pack([[int(np.searchsorted(index.offsets, hot, 'right') - 1)]]), 10)
    assert not (np.array_equal(got_s, want_s) and np.array_equal(got_i, want_i))
