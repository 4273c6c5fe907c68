"""The launch autotuner: every configuration gives the same results, and the chosen one is
the fastest measured."""

import numpy as np
import pytest

import bm25mi
from bm25mi import autotune


@pytest.mark.gpu
def test_autotune_picks_fastest_and_launch_never_changes_results(need_gpu):
    corpus = bm25mi.generate(bm25mi.CorpusConfig(400_000, 200_000, 1.07, 56.0, seed=2))
    index = corpus.index()
    batches = [bm25mi.draw_queries(corpus.offsets, 8, 4, seed=s) for s in (1, 2)]
    r = autotune.tune(index, [(b, 10) for b in batches], 10, block_docs=(2048, 8192), keep_caps=(0, 1024),
                      wave_scales=(0.5, 1.0), task_costs=(1024, 8192), repeats=2, metric='device_ms')
    assert len(r['table']) == 2 * 2 * 2 * 2
    assert r['ms'] == min(row['ms'] for row in r['table'])
    best = r['best']
    with bm25mi.Session(index, max_queries=8, max_words=32, max_k=10, **best) as s:
        got = s.search(batches[0], 10)
        assert s.stats()['score_grid'] == best['grid_blocks']
    with bm25mi.Session(index, max_queries=8, max_words=32, max_k=10) as s:
        want = s.search(batches[0], 10)
        with pytest.raises(bm25mi.Bm25Error):
            s.set_launch(10**7, 0)  # beyond the waves the packet holds
    assert np.array_equal(got[1], want[1]) and np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))


class _FakeSession:
    """Stands in for a Session on the CPU: time depends on the launch, results do not
    (unless `drift` names a launch whose results differ)."""

    def __init__(self, bd, cap, drift=None):
        self.bd, self.cap, self.waves, self.cost, self.drift = bd, cap, 100, 4096, drift

    def search(self, queries, k, timed=False):
        ids = np.arange(len(queries) * k, dtype=np.int32).reshape(len(queries), k)
        if (self.bd, self.cap, self.waves, self.cost) == self.drift:
            ids = ids[:, ::-1].copy()
        return np.ones(ids.shape, np.float32), ids, np.zeros(len(queries), np.int32)

    def stats(self):
        ms = abs(self.waves - 75) * 0.01 + abs(self.cost - 2048) * 1e-5 + (0.0 if self.bd == 4096 else 0.5) + \
            (0.0 if self.cap == 1024 else 0.25)
        return {'score_grid': self.waves, 'score_kernel_ms': ms, 'device_ms': ms + 0.01}

    def set_launch(self, waves, cost):
        self.waves, self.cost = waves, cost

    def close(self):
        pass


def test_autotune_selection_on_cpu():
    batches = [([[1, 2], [3]], 5), ([[4]], 5)]
    r = autotune.tune(None, batches, 5, keep_caps=autotune.KEEP_CAPS, repeats=1, session_factory=_FakeSession)
    assert r['best'] == {'block_docs': 4096, 'keep_cap': 1024, 'grid_blocks': 75, 'task_cost': 2048}
    assert len(r['table']) == (len(autotune.BLOCK_DOCS) * len(autotune.KEEP_CAPS) * len(autotune.WAVE_SCALES) *
                               len(autotune.TASK_COSTS))
    assert 8192 in autotune.BLOCK_DOCS and {1024, 4096, 8192} <= set(autotune.KEEP_CAPS)


def test_autotune_rejects_a_launch_that_changes_results():
    batches = [([[1, 2], [3]], 5)]
    with pytest.raises(AssertionError, match='changed the results'):
        autotune.tune(None, batches, 5, keep_caps=(0, 4096), repeats=1,
                      session_factory=lambda bd, cap: _FakeSession(bd, cap, (4096, 4096, 150, 8192)))
