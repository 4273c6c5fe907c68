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
    """ErrorInfo contract (FR/ut/cases/bm25/sdk_error_adapter_tests.cpp:34-39): a 2048-character
    message reaches bm25mi_last_error cut to 1023 characters and NUL-terminated, the device code
    is kept, and the next successful call clears both. The message is injected at the create
    stage, which fails before any device work."""
    lib = _lib.load()
    long_message = b'x' * 2048
    assert lib.bm25mi_inject_failure(b'create', 161001, long_message) == _lib.OK
    cfg = _lib.Config()
    lib.bm25mi_config_default(C.byref(cfg))
    handle = C.c_void_p()
    index = bm25mi.Index([0, 2], [0, 2], [1, 3], [1, 2, 1], [0.5])
    rc = lib.bm25mi_create(_lib.ptr(index.offsets), index.vocab, _lib.ptr(index.docs), _lib.ptr(index.tf),
                           index.postings, _lib.ptr(index.norm), index.num_docs, _lib.ptr(index.idf), C.byref(cfg),
                           C.byref(handle))
    assert rc == _lib.INNER and not handle.value
    buf = C.create_string_buffer(4096)
    assert lib.bm25mi_last_error(None, buf, len(buf)) == 161001
    assert buf.value == b'x' * 1023
    small = C.create_string_buffer(8)
    lib.bm25mi_last_error(None, small, len(small))
    assert small.value == b'x' * 7  # cut to the caller's buffer, NUL-terminated
    # A success clears the message and the code.
    assert lib.bm25mi_validate_index(_lib.ptr(index.offsets), index.vocab, _lib.ptr(index.docs), _lib.ptr(index.tf),
                                     index.postings, _lib.ptr(index.norm), index.num_docs,
                                     _lib.ptr(index.idf)) == _lib.OK
    assert lib.bm25mi_last_error(None, buf, len(buf)) == 0 and buf.value == b''
    # Unknown stages are rejected; a NULL stage clears pending injections.
    assert lib.bm25mi_inject_failure(b'launch', 1, b'') == _lib.INVALID_PARAM
    assert lib.bm25mi_inject_failure(b'search', 1, b'pending') == _lib.OK
    assert lib.bm25mi_inject_failure(None, 0, None) == _lib.OK


def test_facade_init_failure_keeps_device_code():
    """The facade's Init failure keeps the device code (sdk_facade_tests.cpp:38-40) and leaves
    the object uninitialised (a later Search is an ILLEGAL_OPERATION)."""
    lib = _lib.load()
    index = bm25mi.Index([0, 2], [0, 2], [1, 3], [1, 2, 1], [0.5])
    facade = bm25mi.AscendIndexBm25()
    assert lib.bm25mi_inject_failure(b'create', 42, b'injected init failure') == _lib.OK
    assert facade.Init(index, bm25mi.Bm25Config(queryCount=1, maxQueryWords=1, topK=1)) == \
        bm25mi.search.APP_ERR_INNER_ERROR
    err = facade.GetLastError()
    assert err.aclCode == 42 and err.message == 'injected init failure'
    result = bm25mi.Bm25SearchResult()
    assert facade.Search([[0]], result) == bm25mi.search.APP_ERR_ILLEGAL_OPERATION


def test_config_defaults():
    """Documented defaults of bm25mi_config_default: direct launches (no graph replay), 2048-doc
    tiles, the full grid and the default task cost chosen by the library."""
    import ctypes as C
    from bm25mi import _lib
    cfg = _lib.Config()
    _lib.load().bm25mi_config_default(C.byref(cfg))
    assert cfg.use_graph == 0 and cfg.block_docs == 2048
    assert cfg.grid_blocks == 0 and cfg.task_cost == 0 and cfg.slots == 1 and cfg.pool_entries == 0
    assert cfg.live_max_k == 0 and cfg.keep_cap == 0
    assert cfg.max_queries == 8 and cfg.max_words == 24 and cfg.max_k == 10 and cfg.k1 == 1.2


def test_mlp_param_count():
    lib = _lib.load()
    assert lib.bm25mi_mlp_param_count(16) == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1


def test_search_session_request_and_response_checks():
    """The raw-ID Search surface's checks (standalone/run_bm25s_index_search.py:38-46, 56-75),
    without a GPU: accepted and rejected requests, and responses that break the contract."""
    from bm25mi.search import check_request, check_response
    check_request([[1, 2], [], [2**31 - 2]], 3, 3)
    check_request([[]], 1, 0)
    for bad in ([], [[1]] * 4, ([1],), [(1,)], [[True]], [[-1]], [[2**31 - 1]], [[1.0]], [[np.int32(1)]],
                [[1, 2], [3, 4]], 'abc', None):
        with pytest.raises(ValueError):
            check_request(bad, 3, 3)
    scores = np.array([[3.0, 2.0, 2.0], [0.0, 0.0, 0.0]], np.float32)
    ids = np.array([[5, 1, 4], [0, 1, 2]], np.int32)
    check_response(scores, ids, [[1], [2]], 3, 10)
    broken = [
        (scores.astype(np.float64), ids), (scores, ids.astype(np.int64)), (scores[:, :2], ids[:, :2]),
        (np.where(scores == 3.0, np.nan, scores).astype(np.float32), ids),
        (np.where(scores == 3.0, -1.0, scores).astype(np.float32), ids),
        (scores, np.array([[5, 1, 10], [0, 1, 2]], np.int32)),        # ID out of range
        (scores, np.array([[5, 1, 1], [0, 1, 2]], np.int32)),         # duplicate ID
        (scores, np.array([[5, 4, 1], [0, 1, 2]], np.int32)),         # tie not by ascending ID
        (np.array([[2.0, 3.0, 2.0], [0, 0, 0]], np.float32), ids),   # not descending
    ]
    for s, i in broken:
        with pytest.raises(RuntimeError):
            check_response(s, i, [[1], [2]], 3, 10)


def test_array_address_cache():
    """_lib.ptr caches the addresses of small arrays it has seen (request arrays and output
    buffers a caller reuses): the address equals numpy's, the array cannot be resized in place
    while cached, non-contiguous arrays are rejected, large arrays are not cached and the cache
    stays bounded."""
    a = np.arange(64, dtype=np.int64)
    assert _lib.ptr(a) == a.ctypes.data and _lib.ptr(a) == a.ctypes.data
    with pytest.raises(ValueError):
        a.resize(128)  # the cached address would go stale
    with pytest.raises(ValueError):
        _lib.ptr(np.arange(64, dtype=np.int32)[::2])
    assert _lib.ptr(None) is None
    big = np.zeros(_lib._ADDR_BYTES // 4 + 1, np.int32)
    assert _lib.ptr(big) == big.ctypes.data and id(big) not in _lib._addr
    keep = [np.zeros(4, np.int32) for _ in range(3 * _lib._ADDR_MAX)]
    for x in keep:
        assert _lib.ptr(x) == x.ctypes.data
    assert len(_lib._addr) <= _lib._ADDR_MAX
