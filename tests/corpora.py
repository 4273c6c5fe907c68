"""Small deterministic corpora shared by the tests."""

import math
import random
from collections import Counter

import numpy as np

import bm25mi


def tiny_documents(n):
    """Documents(n) of FR/ut/cases/bm25/cpu_index_search_tests.cpp:23-36 (vocabulary 7)."""
    docs = [[] for _ in range(n)]
    for i in range(n):
        if i % 7:
            docs[i] = [i % 4] * (1 + i % 3)
            if i % 3 == 0:
                docs[i].append(4)
        if i % 11 == 0:
            docs[i] += [0, 5, 5]
    return docs


def index_from_documents(docs, vocab, k1=1.2, b=0.75):
    """CSC index with the reference builder's formulas (cpu_index_search_tests.cpp:37-64,
    standalone/large_corpus.py:148,174): idf and norm in double, stored as FP32."""
    offsets, ids, tf = [0], [], []
    counts = [Counter(d) for d in docs]
    for t in range(vocab):
        for d, c in enumerate(counts):
            if c.get(t):
                ids.append(d)
                tf.append(float(c[t]))
        offsets.append(len(ids))
    n = len(docs)
    avg = sum(len(d) for d in docs) / n
    df = np.diff(offsets)
    idf = [np.float32(math.log1p((n - x + .5) / (x + .5))) for x in df]
    norm = [np.float32(k1 * (1 - b + b * len(d) / avg)) for d in docs]
    return bm25mi.Index(np.array(offsets, np.int64), np.array(ids, np.int32), np.array(tf, np.float32),
                        np.array(norm, np.float32), np.array(idf, np.float32), k1=k1, b=b)


def independent_scores(docs, queries, vocab, k1=1.2, b=0.75):
    """Per-document FP32 oracle from the original documents
    (cpu_index_search_tests.cpp:65-93): sum in query-term order, repeats included."""
    n = len(docs)
    avg = sum(len(d) for d in docs) / n
    f32 = np.float32
    idf = []
    for t in range(vocab):
        df = sum(1 for d in docs if t in d)
        idf.append(f32(math.log1p((n - df + .5) / (df + .5))))
    kp1 = f32(k1 + 1.0)
    out = np.zeros((len(queries), n), np.float32)
    for q, query in enumerate(queries):
        for d, doc in enumerate(docs):
            norm = f32(k1 * (1 - b + b * len(doc) / avg))
            score = f32(0)
            for t in query:
                if t < vocab:
                    tf = f32(doc.count(t))
                    if tf > 0:
                        score = f32(score + f32(f32(idf[t] * f32(tf * kp1)) / f32(tf + norm)))
            out[q, d] = score
    return out


def strict_top(scores_row, k):
    """Top-k doc IDs by (score desc, id asc) from a full sort."""
    order = np.lexsort((np.arange(len(scores_row)), -scores_row.astype(np.float64)))
    return order[:k]


def uniform_documents(n, vocab, seed):
    """The reference's uniform synthetic stream (standalone/large_corpus.py:58-62), written
    with Python's own random.Random: length randrange(16, 49), tokens randrange(vocab)."""
    rnd = random.Random(seed)
    for _ in range(n):
        length = rnd.randrange(16, 49)
        yield [rnd.randrange(vocab) for _ in range(length)]


def uniform_index(n, vocab, seed, k1=1.2):
    docs = list(uniform_documents(n, vocab, seed))
    return index_from_documents_fast(docs, vocab, k1=k1)


def index_from_documents_fast(docs, vocab, k1=1.2, b=0.75):
    """Same as index_from_documents, vectorised for larger corpora."""
    n = len(docs)
    rows, cols, vals = [], [], []
    for d, doc in enumerate(docs):
        for t, c in Counter(doc).items():
            rows.append(t)
            cols.append(d)
            vals.append(c)
    rows, cols, vals = np.array(rows, np.int64), np.array(cols, np.int64), np.array(vals, np.float32)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    df = np.bincount(rows, minlength=vocab)
    offsets = np.zeros(vocab + 1, np.int64)
    offsets[1:] = np.cumsum(df)
    lengths = np.array([len(d) for d in docs], np.int32)
    total = float(lengths.sum())
    avg = total / n
    idf = np.array([np.float32(math.log1p((n - x + .5) / (x + .5))) for x in df], np.float32)
    norm = np.array([np.float32(k1 * (1 - b + b * int(L) / avg)) for L in lengths], np.float32)
    return bm25mi.Index(offsets, cols.astype(np.int32), vals, norm, idf, k1=k1, b=b)


QUERY_A = [[0, 0], [], [6], [7], [4, 5, 0], [2147483646, 2], [0, 1, 2, 3]]


def golden_index(name):
    """Corpora the golden vectors were computed on (all at the reference's k1 = 1.2)."""
    if name == 'uniform1024_8192':
        # The reference's published small synthetic workload: 8192 docs, vocab 1024,
        # seed 20261014 (evidence/synthetic8192_regression_a5_20261105.json).
        return uniform_index(8192, 1024, 20261014, k1=1.2)
    if name == 'synthetic8k_k12':
        return bm25mi.generate('synthetic8k').index(k1=1.2)
    if name == 'tiny129':
        return index_from_documents(tiny_documents(129), 7)
    raise KeyError(name)
