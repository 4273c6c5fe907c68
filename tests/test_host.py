"""Host-side logic on the CPU: the C ABI library loads and exports every declared
symbol, the corpus generator and index builder, index validation known answers
(FR/ut/cases/bm25/index_data_tests.cpp:27-44), impacts, file formats, and the
request/facade checks that need no GPU."""

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import bm25mi
from bm25mi import _lib
from oracle import oracle as orc

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / 'include' / 'bm25mi.h').read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(bm25mi_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), 'ctypes prototypes and header disagree'


def test_generator_deterministic_and_thread_independent():
    cfg = bm25mi.CorpusConfig(20000, 3000, 1.1, 40.0, seed=9)
    a = bm25mi.generate(cfg, threads=1)
    b = bm25mi.generate(cfg, threads=7)
    for x, y in ((a.offsets, b.offsets), (a.docs, b.docs), (a.tf, b.tf), (a.doc_len, b.doc_len)):
        assert np.array_equal(x, y)
    c = bm25mi.generate(bm25mi.CorpusConfig(20000, 3000, 1.1, 40.0, seed=10))
    assert not np.array_equal(a.docs, c.docs)


def test_generator_csc_invariants_and_statistics():
    corpus = bm25mi.generate('synthetic8k')
    off, docs, tf, lengths = corpus.offsets, corpus.docs, corpus.tf, corpus.doc_len
    assert off[0] == 0 and off[-1] == len(docs) and np.all(np.diff(off) >= 0)
    for t in range(corpus.vocab):
        seg = docs[off[t]:off[t + 1]]
        assert np.all(np.diff(seg) > 0)
    # tf sums to the document lengths; every tf is a positive integer.
    assert np.all(tf >= 1) and np.all(np.floor(tf) == tf)
    assert np.bincount(docs, weights=tf, minlength=corpus.num_docs).astype(np.int64).tolist() == lengths.tolist()
    assert abs(lengths.mean() - 60.0) < 1.5
    # Zipf: term 0 is the most frequent token, and frequency falls with rank.
    freq = np.bincount(np.repeat(np.arange(corpus.vocab), np.diff(off)), weights=tf, minlength=corpus.vocab)
    assert freq[0] == freq.max() and freq[:10].sum() > freq[100:110].sum() * 5


def test_weights_and_impacts_match_oracle_bitwise():
    corpus = bm25mi.generate('synthetic8k')
    for k1 in (1.2, 1.5):
        index = corpus.index(k1=k1)
        idf, norm = orc.weights(corpus.offsets, corpus.doc_len, k1, 0.75)
        assert np.array_equal(idf.view(np.uint32), index.idf.view(np.uint32))
        assert np.array_equal(norm.view(np.uint32), index.norm.view(np.uint32))
        impacts = index.impacts()
        term = np.repeat(np.arange(index.vocab), np.diff(index.offsets))
        kp1 = np.float32(k1 + 1.0)
        want = (index.idf[term] * (index.tf * kp1)) / (index.tf + index.norm[index.docs])
        assert want.dtype == np.float32
        assert np.array_equal(impacts.view(np.uint32), want.view(np.uint32))


def test_index_validation_known_answers():
    """index_data_tests.cpp:27-44."""
    valid = bm25mi.Index([0, 2], [0, 2], [1, 3], [1, 2, 1], [0.5])
    assert valid.postings == 2 and valid.num_docs == 3
    bm25mi.Index([0, 0], [], [], [1], [0])  # empty posting list
    bad = [
        ([], [], [], [], []),
        ([0, 1], [0], [1], [], [1]),
        ([0, 2], [0, 0], [1, 1], [1], [1]),
        ([0, 2], [1, 0], [1, 1], [1, 1], [1]),
        ([0, 1], [1], [1], [1], [1]),
        ([0, 1], [0], [1.5], [1], [1]),
        ([0, 0], [], [], [0], [1]),
        ([0, 0], [], [], [1], [-1]),
        ([0, 0], [], [], [1], [float('nan')]),
        ([1, 1], [], [], [1], [1]),
        ([0, 1, 0], [], [], [1], [1, 1]),
        ([0, 1], [0], [16777218.0], [1], [1]),
        ([0, 1], [0], [float('inf')], [1], [1]),
    ]
    for args in bad:
        with pytest.raises(ValueError):
            bm25mi.Index(*args)


def test_query_files_and_packing():
    qa = bm25mi.read_queries(ROOT / 'tests' / 'golden' / 'reference_queries_A.txt')
    assert len(qa) == 8 and all(len(r) == 3 for r in qa)
    assert qa[0] == [783, 343, 370]
    offsets, terms = bm25mi.pack(qa + [[], [5, 5]])
    assert offsets.tolist()[-3:] == [24, 24, 26] and terms.dtype == np.int32
    # Edge cases of the packing: no queries, empty rows only, numpy rows, and IDs that do not
    # fit int32 (rejected rather than wrapped).
    offsets, terms = bm25mi.pack([])
    assert offsets.tolist() == [0] and offsets.dtype == np.int64 and terms.shape == (0,)
    offsets, terms = bm25mi.pack([[], []])
    assert offsets.tolist() == [0, 0, 0] and terms.shape == (0,) and terms.dtype == np.int32
    offsets, terms = bm25mi.pack([np.array([7, 8], np.int64), [9]])
    assert offsets.tolist() == [0, 2, 3] and terms.tolist() == [7, 8, 9]
    with pytest.raises(OverflowError):
        bm25mi.pack([[2**31]])


def test_query_file_round_trip(tmp_path):
    rows = [[1, 2, 3], [], [2147483646, 0, 0]]
    bm25mi.write_queries(tmp_path / 'q.txt', rows)
    assert bm25mi.read_queries(tmp_path / 'q.txt') == rows
    for bad in ([], [[-1]], [[2**31 - 1]], [[True]]):
        with pytest.raises(ValueError):
            bm25mi.write_queries(tmp_path / 'bad.txt', bad)


def test_index_file_round_trip(tmp_path):
    index = bm25mi.generate('synthetic8k').index()
    record = bm25mi.save_index(index, tmp_path / 'index')
    assert record['posting_count'] == index.postings and len(record['index_id']) == 64
    back = bm25mi.load_index(tmp_path / 'index')
    for a, b in ((back.offsets, index.offsets), (back.docs, index.docs), (back.tf, index.tf),
                 (back.norm, index.norm), (back.idf, index.idf)):
        assert np.array_equal(a, b)
    raw = bytearray((tmp_path / 'index' / 'idf.bin').read_bytes())
    raw[0] ^= 1
    (tmp_path / 'index' / 'idf.bin').write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        bm25mi.load_index(tmp_path / 'index')


def test_draw_queries_distinct_and_df_weighted():
    corpus = bm25mi.generate('synthetic8k')
    qs = bm25mi.draw_queries(corpus.offsets, 2000, 3, seed=1)
    assert all(len(set(q)) == 3 for q in qs)
    flat = np.array([t for q in qs for t in q])
    df = corpus.df()
    # Drawn terms are much more frequent than the vocabulary average.
    assert df[flat].mean() > 5 * df.mean()
    lengths = bm25mi.corpus.geometric_lengths(5000, 6, 1, 32, seed=2)
    assert lengths.min() >= 1 and lengths.max() <= 32 and abs(lengths.mean() - 6) < 0.5


def test_create_without_gpu_reports_error_and_facade_state_machine():
    index = bm25mi.Index([0, 2], [0, 2], [1, 3], [1, 2, 1], [0.5])
    facade = bm25mi.AscendIndexBm25()
    result = bm25mi.Bm25SearchResult()
    assert facade.Search([[0]], result) == bm25mi.search.APP_ERR_ILLEGAL_OPERATION
    assert 'not initialized' in facade.GetLastError().message
    assert facade.Close() == bm25mi.search.APP_ERR_OK
    assert facade.Close() == bm25mi.search.APP_ERR_OK  # idempotent
    # Invalid index -> INVALID_PARAM before any device work.
    broken = bm25mi.Index([0, 2], [0, 2], [1, 3], [1, 2, 1], [0.5], validate=False)
    broken.tf[0] = -1
    assert facade.Init(broken, bm25mi.Bm25Config(topK=1)) == bm25mi.search.APP_ERR_INVALID_PARAM
    import tests.conftest as cf
    if not cf.gpu_available():
        code = facade.Init(index, bm25mi.Bm25Config(topK=1))
        assert code != bm25mi.search.APP_ERR_OK
        assert facade.GetLastError().message


def test_last_error_truncates_to_1023_chars():
    lib = _lib.load()
    huge = 'x' * 5000
    # An invalid k1 carries a short message; the buffer contract still NUL-terminates.
    buf = C.create_string_buffer(8)
    rc = lib.bm25mi_impacts(None, 1, None, None, None, None, 1.2, None, 0)
    assert rc == _lib.INVALID_PARAM
    lib.bm25mi_last_error(None, buf, len(buf))
    assert len(buf.value) == 7
    assert len(huge[:1023]) == 1023


def test_config_defaults():
    """Documented defaults of bm25mi_config_default: direct launches (no graph replay), 2048-doc
    tiles, the full grid and the default task cost chosen by the library."""
    import ctypes as C
    from bm25mi import _lib
    cfg = _lib.Config()
    _lib.load().bm25mi_config_default(C.byref(cfg))
    assert cfg.use_graph == 0 and cfg.block_docs == 2048
    assert cfg.grid_blocks == 0 and cfg.task_cost == 0 and cfg.slots == 1 and cfg.pool_entries == 0
    assert cfg.max_queries == 8 and cfg.max_words == 24 and cfg.max_k == 10 and cfg.k1 == 1.2


def test_mlp_param_count():
    lib = _lib.load()
    assert lib.bm25mi_mlp_param_count(16) == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1
