"""The in-memory BM25 index: ascend::bm25::Index (FR/common/bm25/index_data.h:9-42)."""

import numpy as np

from . import _lib


class Index:
    """offsets int64[V+1], docs int32[M], tf float32[M], norm float32[N], idf float32[V].

    Construction validates like the reference constructor (index_data.h:13-41):
    sorted unique in-range doc IDs per term, integral tf in [1, 2^24], finite
    positive norms, finite non-negative idf. Invalid input raises ValueError.
    """

    def __init__(self, offsets, docs, tf, norm, idf, k1=1.2, b=0.75, validate=True):
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.docs = np.ascontiguousarray(docs, np.int32)
        self.tf = np.ascontiguousarray(tf, np.float32)
        self.norm = np.ascontiguousarray(norm, np.float32)
        self.idf = np.ascontiguousarray(idf, np.float32)
        self.k1, self.b = float(k1), float(b)
        if validate:
            self.validate()

    @property
    def num_docs(self):
        return len(self.norm)

    @property
    def vocab(self):
        return len(self.idf)

    @property
    def postings(self):
        return len(self.docs)

    def df(self):
        return np.diff(self.offsets)

    def validate(self):
        if self.norm.size == 0 or self.idf.size == 0 or self.offsets.size != self.idf.size + 1 \
                or self.docs.size != self.tf.size or self.offsets[0] != 0 or self.offsets[-1] != self.docs.size:
            raise ValueError('Index dimensions/endpoints differ')
        lib = _lib.load()
        rc = lib.bm25mi_validate_index(_lib.ptr(self.offsets), self.vocab, _lib.ptr(self.docs), _lib.ptr(self.tf),
                                       self.postings, _lib.ptr(self.norm), self.num_docs, _lib.ptr(self.idf))
        if rc != _lib.OK:
            raise ValueError(_lib.last_error()[0])

    def impacts(self, threads=0):
        """FP32 impacts idf*(tf*(k1+1)) / (tf+norm) per posting, as the device stores them."""
        out = np.empty(self.postings, np.float32)
        _lib.check(_lib.load().bm25mi_impacts(_lib.ptr(self.offsets), self.vocab, _lib.ptr(self.docs),
                                              _lib.ptr(self.tf), _lib.ptr(self.norm), _lib.ptr(self.idf),
                                              float(self.k1), _lib.ptr(out), threads))
        return out

    def copy(self):
        return Index(self.offsets.copy(), self.docs.copy(), self.tf.copy(), self.norm.copy(), self.idf.copy(),
                     self.k1, self.b, validate=False)
