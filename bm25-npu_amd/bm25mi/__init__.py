"""bm25mi: BM25 full Search on one MI355X (gfx950) behind a C-ABI library.

Search(query term IDs, K) -> (doc IDs, FP32 scores) on a resident index, exact
under strict ranking (score descending, then doc ID ascending).
"""

from . import _lib
from ._lib import Bm25Error
from .corpus import CONFIGS, Corpus, CorpusConfig, draw_queries, generate, load_index, pack, read_queries, \
    save_index, write_queries
from .index import Index
from .search import AscendIndexBm25, Bm25Config, Bm25SearchResult, SearchSession, Session
from .serving import Server

__all__ = [
    'AscendIndexBm25', 'Bm25Config', 'Bm25Error', 'Bm25SearchResult', 'CONFIGS', 'Corpus', 'CorpusConfig',
    'Index', 'SearchSession', 'Server', 'Session', 'draw_queries', 'generate', 'load_index', 'pack', 'read_queries',
    'save_index', 'write_queries', '_lib',
]
