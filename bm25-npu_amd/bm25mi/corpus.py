"""Synthetic corpora, query sets and the on-disk index format.

The corpus generator runs in the library (C++, multithreaded, deterministic
per seed). Index files follow the reference's resident-index contract
(standalone/resident_index.py:12-18): term_offsets.bin int64[V+1],
posting_doc_ids.bin int32[M], posting_tf.bin float32[M], doc_norm.bin
float32[N], idf.bin float32[V], plus metadata.json. Query files use the
reference's BM25_QUERY_IDS_V1 text format (standalone/resident_index.py:97-106).
"""

import ctypes as C
import hashlib
import itertools
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib
from .index import Index

INDEX_FILES = {
    'term_offsets.bin': np.int64,
    'posting_doc_ids.bin': np.int32,
    'posting_tf.bin': np.float32,
    'doc_norm.bin': np.float32,
    'idf.bin': np.float32,
}
LIMIT = 2**31 - 1


@dataclass(frozen=True)
class CorpusConfig:
    num_docs: int
    vocab: int
    zipf_s: float
    mean_len: float
    sigma: float = 0.5
    seed: int = 0
    k1: float = 1.5
    b: float = 0.75


# The project's workloads: a CPU-runnable 8K-doc corpus, the 10M-doc corpus of
# the headline Search, and the 8.84M-passage corpus of the batched runs.
CONFIGS = {
    'synthetic8k': CorpusConfig(8192, 5000, 1.1, 60.0),
    'synthetic10m': CorpusConfig(10_000_000, 1_000_000, 1.07, 56.0),
    'passages8p8m': CorpusConfig(8_841_823, 2_600_000, 1.05, 56.0),
}


class Corpus:
    """Generated raw corpus: CSC term offsets, doc IDs, tf and doc lengths (numpy)."""

    def __init__(self, offsets, docs, tf, doc_len, config=None):
        self.offsets, self.docs, self.tf, self.doc_len = offsets, docs, tf, doc_len
        self.config = config

    @property
    def num_docs(self):
        return len(self.doc_len)

    @property
    def vocab(self):
        return len(self.offsets) - 1

    def df(self):
        return np.diff(self.offsets)

    def index(self, k1=None, b=None, threads=0):
        """BM25 index arrays (idf, norm in double -> FP32) as an Index."""
        k1 = self.config.k1 if k1 is None else k1
        b = self.config.b if b is None else b
        lib = _lib.load()
        idf = np.empty(self.vocab, np.float32)
        norm = np.empty(self.num_docs, np.float32)
        _lib.check(lib.bm25mi_bm25_weights(_lib.ptr(self.offsets), self.vocab, _lib.ptr(self.doc_len),
                                           self.num_docs, float(k1), float(b), _lib.ptr(idf), _lib.ptr(norm),
                                           threads))
        return Index(self.offsets, self.docs, self.tf, norm, idf, k1=k1, b=b, validate=False)


def generate(config, threads=0):
    """Generates a corpus for a CorpusConfig (or a CONFIGS name)."""
    if isinstance(config, str):
        config = CONFIGS[config]
    lib = _lib.load()
    spec = _lib.CorpusSpec(config.num_docs, config.vocab, config.zipf_s, config.mean_len, config.sigma,
                           config.seed, threads)
    handle = C.c_void_p()
    _lib.check(lib.bm25mi_corpus_generate(C.byref(spec), C.byref(handle)))
    try:
        n, v, m = C.c_int64(), C.c_int64(), C.c_int64()
        po, pd, pt, pl = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        _lib.check(lib.bm25mi_corpus_view(handle, C.byref(n), C.byref(v), C.byref(m), C.byref(po), C.byref(pd),
                                          C.byref(pt), C.byref(pl)))

        def copy(p, count, ctype, dtype):
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), shape=(count,)).astype(dtype, copy=True)

        offsets = copy(po, v.value + 1, C.c_int64, np.int64)
        docs = copy(pd, m.value, C.c_int32, np.int32)
        tf = copy(pt, m.value, C.c_float, np.float32)
        doc_len = copy(pl, n.value, C.c_int32, np.int32)
    finally:
        lib.bm25mi_corpus_free(handle)
    return Corpus(offsets, docs, tf, doc_len, config)


def draw_queries(offsets, count, terms, seed, distinct=True):
    """Queries of term IDs drawn with probability proportional to df.

    terms: an int (every query has that many terms) or an int array of per-query
    lengths. Terms inside one query are distinct when `distinct`.
    """
    rng = np.random.default_rng(seed)
    offsets = np.asarray(offsets, np.int64)
    m = int(offsets[-1])
    lengths = np.full(count, terms, np.int64) if np.isscalar(terms) else np.asarray(terms, np.int64)
    nonempty = int(np.count_nonzero(np.diff(offsets)))
    out = []
    for length in lengths:
        length = int(min(length, nonempty)) if distinct else int(length)
        row = []
        while len(row) < length:
            pos = rng.integers(0, m, size=2 * (length - len(row)) + 1)
            for term in np.searchsorted(offsets, pos, side='right') - 1:
                term = int(term)
                if distinct and term in row:
                    continue
                row.append(term)
                if len(row) == length:
                    break
        out.append(row)
    return out


def geometric_lengths(count, mean, low, high, seed):
    """Per-query term counts: geometric with the given mean, truncated to [low, high]."""
    rng = np.random.default_rng(seed)
    p = 1.0 / (mean - low + 1)
    out = np.empty(count, np.int64)
    for i in range(count):
        while True:
            x = low - 1 + int(rng.geometric(p))
            if x <= high:
                out[i] = x
                break
    return out


def pack(queries):
    """List-of-lists -> (offsets int64[Q+1], terms int32[W])."""
    ends = [0]
    ends.extend(itertools.accumulate(map(len, queries)))
    return np.array(ends, np.int64), np.array(list(itertools.chain.from_iterable(queries)), np.int32)


def write_queries(path, queries):
    """BM25_QUERY_IDS_V1 text file (standalone/resident_index.py:97-106)."""
    if not (isinstance(queries, list) and 1 <= len(queries) <= 10000 and all(
            isinstance(r, list) and all(type(w) is int and 0 <= w < LIMIT for w in r) for r in queries)):
        raise ValueError('queries must be 1..10000 lists of non-negative INT32 word IDs')
    lines = ['BM25_QUERY_IDS_V1', str(len(queries))]
    lines += [' '.join([str(len(r))] + [str(w) for w in r]) for r in queries]
    Path(path).write_bytes(('\n'.join(lines) + '\n').encode('ascii'))


def read_queries(path):
    """Parses BM25_QUERY_IDS_V1 (standalone/large_corpus.py:41-55)."""
    parts = Path(path).read_text(encoding='utf-8').split()
    if not parts or parts[0] != 'BM25_QUERY_IDS_V1':
        raise ValueError('Raw grouped query contract missing')
    count = int(parts[1])
    if not 1 <= count <= 10000:
        raise ValueError('Raw query count exceeds declared range')
    pos, rows = 2, []
    for _ in range(count):
        length = int(parts[pos])
        pos += 1
        if not 0 <= length <= 64:
            raise ValueError('Query length exceeds 64')
        rows.append([int(x) for x in parts[pos:pos + length]])
        pos += length
    if pos != len(parts) or not all(0 <= w < LIMIT for r in rows for w in r):
        raise ValueError('Invalid query packet')
    return rows


def save_index(index, directory):
    """Writes the five index arrays and metadata.json with SHA-256 identities."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=False)
    arrays = dict(zip(INDEX_FILES, (index.offsets, index.docs, index.tf, index.norm, index.idf)))
    hashes = {}
    for name, dtype in INDEX_FILES.items():
        data = np.ascontiguousarray(arrays[name], dtype=dtype).astype(np.dtype(dtype).newbyteorder('<'))
        raw = data.tobytes()
        (d / name).write_bytes(raw)
        hashes[name] = hashlib.sha256(raw).hexdigest()
    record = {
        'contract': 'bm25_resident_postings_fp32_v1',
        'num_docs': int(index.num_docs),
        'vocab_size': int(index.vocab),
        'posting_count': int(index.postings),
        'k1': float(index.k1),
        'b': float(index.b),
        'idf_formula': 'log1p((N-df+0.5)/(df+0.5))',
        'files': hashes,
    }
    keys = ('contract', 'num_docs', 'vocab_size', 'posting_count', 'k1', 'b', 'idf_formula', 'files')
    record['index_id'] = hashlib.sha256(json.dumps({k: record[k] for k in keys}, sort_keys=True,
                                                   separators=(',', ':')).encode()).hexdigest()
    (d / 'metadata.json').write_text(json.dumps(record, indent=2) + '\n')
    return record


def load_index(directory, validate=True):
    d = Path(directory)
    meta = json.loads((d / 'metadata.json').read_text())
    arrays = {}
    for name, dtype in INDEX_FILES.items():
        raw = (d / name).read_bytes()
        if hashlib.sha256(raw).hexdigest() != meta['files'][name]:
            raise ValueError(f'{name} does not match metadata.json')
        arrays[name] = np.frombuffer(raw, dtype=np.dtype(dtype).newbyteorder('<')).astype(dtype)
    return Index(arrays['term_offsets.bin'], arrays['posting_doc_ids.bin'], arrays['posting_tf.bin'],
                 arrays['doc_norm.bin'], arrays['idf.bin'], k1=meta.get('k1', 1.2), b=meta.get('b', 0.75),
                 validate=validate)
