"""ctypes binding of libbm25mi.so (the C ABI declared in include/bm25mi.h).

The library is built in-tree by bm25-npu_amd/build.py. Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parents[1] / 'lib' / 'libbm25mi.so'

OK, INVALID_PARAM, ILLEGAL_OPERATION, INNER = 0, 1, 2, 3
MAX_K = 1024
MAX_TERMS = 64
FEATURES = 16
MLP_HIDDEN = 64


class Config(C.Structure):
    _fields_ = [
        ('device_id', C.c_int32),
        ('max_queries', C.c_int32),
        ('max_words', C.c_int64),
        ('max_k', C.c_int32),
        ('block_docs', C.c_int32),
        ('k1', C.c_double),
        ('use_graph', C.c_int32),
        ('grid_blocks', C.c_int32),
        ('task_cost', C.c_int32),
        ('slots', C.c_int32),
        ('pool_entries', C.c_int64),
        ('live_max_k', C.c_int32),
        ('keep_cap', C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ('device_ms', C.c_double),
        ('score_kernel_ms', C.c_double),
        ('host_ms', C.c_double),
        ('postings_scanned', C.c_int64),
        ('postings_total', C.c_int64),
        ('blocks_scored', C.c_int64),
        ('blocks_skipped', C.c_int64),
        ('candidates', C.c_int64),
        ('fallbacks', C.c_int32),
        ('queries', C.c_int32),
        ('final_ms', C.c_double),
        ('postings_in_range', C.c_int64),
        ('emitting_blocks', C.c_int64),
        ('exact_blocks', C.c_int64),
        ('score_grid', C.c_int32),
        ('block_docs', C.c_int32),
        ('phase_cycles', C.c_int64 * 4),
        ('wave_busy_max', C.c_int64),
        ('wave_span', C.c_int64),
        ('wave_start_spread', C.c_int64),
    ]

    def as_dict(self):
        out = {name: getattr(self, name) for name, _ in self._fields_}
        out['phase_cycles'] = list(self.phase_cycles)
        return out


class CorpusSpec(C.Structure):
    _fields_ = [
        ('num_docs', C.c_int64),
        ('vocab', C.c_int64),
        ('zipf_s', C.c_double),
        ('mean_len', C.c_double),
        ('sigma', C.c_double),
        ('seed', C.c_uint64),
        ('threads', C.c_int32),
    ]


P = C.c_void_p
i32, i64, f32, f64 = C.c_int32, C.c_int64, C.c_float, C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    'bm25mi_config_default': (None, [P]),
    'bm25mi_create': (C.c_int, [P, i64, P, P, i64, P, i64, P, P, P]),
    'bm25mi_search': (C.c_int, [P, P, P, i32, i32, P, P]),
    'bm25mi_search_batch': (C.c_int, [P, P, P, i32, P, P, P]),
    'bm25mi_search_batch_noalloc': (C.c_int, [P, P, P, i32, P, P, P]),
    'bm25mi_search_pruned': (C.c_int, [P, P, P, i32, P, P, P, P, P]),
    'bm25mi_search_timed': (C.c_int, [P, P, P, i32, P, P, P, P, P]),
    'bm25mi_set_threshold_model': (C.c_int, [P, P, i64, f32, f32, f32]),
    'bm25mi_search_learned': (C.c_int, [P, P, P, i32, P, i32, P, P, P]),
    'bm25mi_learned_thresholds': (C.c_int, [P, P, P, i32, P, P]),
    'bm25mi_query_features': (C.c_int, [P, P, P, i32, P, P]),
    'bm25mi_query_bounds': (C.c_int, [P, P, P, i32, P, P, P]),
    'bm25mi_server_start': (C.c_int, [P, i32, P]),
    'bm25mi_server_submit': (C.c_int, [P, P, i32, i32, P]),
    'bm25mi_server_wait': (C.c_int, [P, C.c_uint64, P, P, P]),
    'bm25mi_server_submit_pruned': (C.c_int, [P, P, i32, i32, f32, P]),
    'bm25mi_server_run_poisson': (C.c_int, [P, P, P, i32, P, P, i32, f64, C.c_uint64, P, P, P, P]),
    'bm25mi_server_submit_learned': (C.c_int, [P, P, i32, i32, P]),
    'bm25mi_server_counters': (C.c_int, [P, P, P, P, P]),
    'bm25mi_server_stop': (C.c_int, [P]),
    'bm25mi_set_launch': (C.c_int, [P, i32, i32]),
    'bm25mi_last_stats': (C.c_int, [P, P]),
    'bm25mi_destroy': (C.c_int, [P]),
    'bm25mi_last_error': (i64, [P, C.c_char_p, C.c_size_t]),
    'bm25mi_inject_failure': (C.c_int, [C.c_char_p, i64, C.c_char_p]),
    'bm25mi_device_info': (C.c_int, [i32, P, P, P]),
    'bm25mi_corpus_generate': (C.c_int, [P, P]),
    'bm25mi_corpus_view': (C.c_int, [P, P, P, P, P, P, P, P]),
    'bm25mi_corpus_free': (None, [P]),
    'bm25mi_bm25_weights': (C.c_int, [P, i64, P, i64, f64, f64, P, P, i32]),
    'bm25mi_impacts': (C.c_int, [P, i64, P, P, P, P, f64, P, i32]),
    'bm25mi_validate_index': (C.c_int, [P, i64, P, P, i64, P, i64, P]),
    'bm25mi_mlp_param_count': (i64, [i32]),
    'bm25mi_mlp_forward': (C.c_int, [P, P, i32, i32, P, P]),
    'bm25mi_mlp_backward': (C.c_int, [P, P, P, i32, i32, P, P, P]),
    'bm25mi_mlp_adam': (C.c_int, [P, P, P, P, i64, f32, f32, f32, f32, i32, P]),
    'bm25mi_mlp_init': (C.c_int, [i32, C.c_uint64, P]),
    'bm25mi_mlp_gradient': (C.c_int, [P, P, P, i32, i32, P, P]),
    'bm25mi_mlp_predict': (C.c_int, [P, P, i32, i32, P]),
    'bm25mi_mlp_train': (C.c_int, [P, P, P, i32, i32, i32, i32, f32, C.c_uint64, P]),
}

_lib = None


def load():
    """Returns the loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get('BM25MI_LIB', str(LIB_PATH)))
        if not path.exists():
            raise RuntimeError(f'{path} is missing: build it with python bm25-npu_amd/build.py')
        lib = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class Bm25Error(RuntimeError):
    """A non-OK status from the library; .code mirrors the APP_ERR class."""

    def __init__(self, code, message, device_code=0):
        super().__init__(f'[{code}] {message}')
        self.code = code
        self.message = message
        self.device_code = device_code


def last_error(handle=None):
    buf = C.create_string_buffer(1024)
    device_code = load().bm25mi_last_error(handle, buf, len(buf))
    return buf.value.decode('utf-8', 'replace'), device_code


def check(rc, handle=None):
    if rc != OK:
        message, device_code = last_error(handle)
        raise Bm25Error(rc, message, device_code)


# Addresses of small arrays passed again (request arrays and output buffers a caller reuses):
# numpy's ctypes.data costs about 2 us per array, a few percent of a Search. The cache holds each
# array itself, so its id cannot be reused by another object and numpy refuses to resize it in
# place (refcheck) while it is cached: the address stays valid. At most _ADDR_MAX entries of at
# most _ADDR_BYTES each; the cache is emptied when full.
_ADDR_MAX = 64
_ADDR_BYTES = 1 << 16
_addr = {}


def ptr(array):
    """Raw data address of a C-contiguous numpy array (or None), passed as a void* argument."""
    if array is None:
        return None
    hit = _addr.get(id(array))
    if hit is not None and hit[0] is array:
        return hit[1]
    if not array.flags.c_contiguous:
        raise ValueError('array must be C-contiguous')
    address = array.ctypes.data
    if array.nbytes <= _ADDR_BYTES:
        if len(_addr) >= _ADDR_MAX:
            _addr.clear()
        _addr[id(array)] = (array, address)
    return address
