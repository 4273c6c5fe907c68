"""The C++ facade (include/bm25mi_index.hpp) through its check program, and the Python
facade's fixed query count rule (FR/common/bm25/fixed_query_packet.h:28)."""

import subprocess
from pathlib import Path

import pytest

import bm25mi

CHECK = Path(__file__).resolve().parents[1] / 'bm25-npu_amd' / 'lib' / 'facade_check'


def run(mode):
    r = subprocess.run([str(CHECK), mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert 'PASS' in r.stdout


def test_cpp_facade_state_machine_without_gpu():
    run('cpu')


@pytest.mark.gpu
def test_cpp_facade_full_contract(need_gpu):
    run('gpu')


@pytest.mark.gpu
def test_python_facade_fixed_query_count(need_gpu):
    index = bm25mi.Index([0, 3, 5, 6], [0, 2, 4, 1, 2, 3], [1, 2, 1, 3, 1, 1], [1.0, 0.9, 1.1, 1.3, 0.8],
                         [0.5, 1.5, 2.0])
    facade = bm25mi.AscendIndexBm25()
    cfg = bm25mi.Bm25Config(queryCount=2, maxQueryWords=8, topK=3)
    assert facade.Init(index, cfg) == bm25mi.search.APP_ERR_OK
    result = bm25mi.Bm25SearchResult()
    assert facade.Search([[0]], result) == bm25mi.search.APP_ERR_INVALID_PARAM
    assert 'declared query count' in facade.GetLastError().message
    assert facade.Search([[0, 1], [2]], result) == bm25mi.search.APP_ERR_OK
    assert result.labels.tolist()[:3] == [1, 2, 4] and result.labels.dtype.name == 'int64'
    assert facade.Close() == bm25mi.search.APP_ERR_OK
