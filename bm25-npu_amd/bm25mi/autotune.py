"""Autotuner of the scoring launch (no reference counterpart; the reference tuned its
Ascend kernel by hand-edited source variants, tools/capacity_sources.py:13-60, whose
candidate capacities 1024/2048/4096 per leaf are the keep_cap axis here).

Searches the tile size (docs per scoring task: one wave's LDS accumulator row), the candidate
capacity of a task (keep_cap: how many candidates a task may append before it cuts to its
top max(K, keep_cap); 0 = the library's max(2K, 64)), the number of persistent scoring waves
and the per-task cost the host schedule uses to give every wave equal work, by timing the
device work of a Search (HIP events) on a sample workload. Every configuration's results are
checked to equal the first one's (the launch never changes results), and the fastest is
returned as Session keyword arguments.

    python -m bm25mi.autotune --workload headline --out tuned.json
    python -m bm25mi.autotune --workload passages --out tuned.json   # uneven batches, Q1-256
"""

import argparse
import itertools
import json
import statistics
import sys
from pathlib import Path

import numpy as np

BLOCK_DOCS = (2048, 4096, 8192)
KEEP_CAPS = (0, 1024, 4096, 8192)
WAVE_SCALES = (0.5, 0.75, 1.0, 1.5)
TASK_COSTS = (1024, 2048, 4096, 8192, 16384)


def measure(session, batches, repeats, metric):
    """Mean per-pass time (the sum over the batches of `metric`, ms) over repeats, and the results
    of the last pass. batches: (queries, K or per-query Ks) pairs."""
    passes, results = [], []
    for _ in range(repeats):
        results, total = [], 0.0
        for queries, ks in batches:
            scores, ids, _ = session.search(queries, ks, timed=True)
            total += session.stats()[metric]
            results.append((scores, ids))
        passes.append(total)
    return statistics.mean(passes), results


def tune(index, batches, max_k, block_docs=BLOCK_DOCS, keep_caps=(0,), wave_scales=WAVE_SCALES,
         task_costs=TASK_COSTS, repeats=5, metric='score_kernel_ms', session_factory=None, log=None):
    """Returns {'best': Session kwargs, 'ms': best time per pass over the batches, 'metric',
    'table': all timings}. batches: (queries, K or per-query Ks) pairs."""
    from .search import Session

    def make(bd, cap):
        return Session(index, max_queries=max(len(q) for q, _ in batches),
                       max_words=max(sum(len(x) for x in q) for q, _ in batches), max_k=max_k, block_docs=bd,
                       keep_cap=cap)

    factory = session_factory or make
    table, reference, best = [], None, None
    for bd, cap in itertools.product(block_docs, keep_caps):
        s = factory(bd, cap)
        try:
            s.search(*batches[0])
            full = s.stats()['score_grid']
            for scale, cost in itertools.product(wave_scales, task_costs):
                waves = max(1, int(round(full * scale)))
                s.set_launch(waves, cost)
                ms, results = measure(s, batches, repeats, metric)
                if reference is None:
                    reference = results
                for (s0, i0), (s1, i1) in zip(reference, results):
                    if not (np.array_equal(i0, i1) and np.array_equal(s0.view(np.uint32), s1.view(np.uint32))):
                        raise AssertionError(f'launch {bd}/{cap}/{waves}/{cost} changed the results')
                row = {'block_docs': bd, 'keep_cap': cap, 'grid_blocks': waves, 'task_cost': cost, 'ms': ms}
                table.append(row)
                if log:
                    log(json.dumps(row))
                if best is None or ms < best['ms']:
                    best = row
        finally:
            s.close()
    return {'best': {k_: best[k_] for k_ in ('block_docs', 'keep_cap', 'grid_blocks', 'task_cost')},
            'ms': best['ms'], 'metric': metric, 'table': table}


def workload(name, seed=11):
    """(index, batches, max_k, description) of a named sample workload: 'headline' (10M docs, two
    Q8 x 3-term batches, K10) or 'passages' (8.84M passages, Q in {1, 8, 64, 256}, geometric 1-32
    terms of mean 6, K drawn from {10, 100, 1000} per query, two batches each)."""
    from . import corpus as corpus_mod
    if name == 'headline':
        c = corpus_mod.generate('synthetic10m')
        batches = [(corpus_mod.draw_queries(c.offsets, 8, 3, seed=s), 10) for s in (seed, seed + 1)]
        return c.index(), batches, 10, 'synthetic10m, two Q8 x 3-term batches, K10'
    if name == 'passages':
        c = corpus_mod.generate('passages8p8m')
        rng = np.random.default_rng(seed)
        batches = []
        for nq in (1, 8, 64, 256):
            for j in (0, 1):
                lengths = corpus_mod.geometric_lengths(nq, 6, 1, 32, seed=seed * 1000 + nq + j)
                ks = rng.choice(np.array([10, 100, 1000], np.int32), size=nq).astype(np.int32)
                batches.append((corpus_mod.draw_queries(c.offsets, nq, lengths, seed=seed * 2000 + nq + j), ks))
        return c.index(), batches, 1000, ('passages8p8m, Q in {1, 8, 64, 256} x 2 batches, geometric 1-32 terms '
                                          '(mean 6), K from {10, 100, 1000}')
    raise ValueError(f'unknown workload {name}')


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument('--workload', choices=('headline', 'passages'), default='headline')
    ap.add_argument('--repeats', type=int, default=3)
    ap.add_argument('--metric', choices=('score_kernel_ms', 'device_ms'), default='device_ms')
    ap.add_argument('--block-docs', type=int, nargs='*', default=list(BLOCK_DOCS))
    ap.add_argument('--keep-caps', type=int, nargs='*', default=list(KEEP_CAPS))
    ap.add_argument('--wave-scales', type=float, nargs='*', default=[0.75, 1.0])
    ap.add_argument('--task-costs', type=int, nargs='*', default=[1024, 4096, 16384])
    ap.add_argument('--out', default=None)
    args = ap.parse_args()
    index, batches, max_k, desc = workload(args.workload)
    result = tune(index, batches, max_k, block_docs=args.block_docs, keep_caps=args.keep_caps,
                  wave_scales=args.wave_scales, task_costs=args.task_costs, repeats=args.repeats,
                  metric=args.metric, log=lambda m: print(m, file=sys.stderr, flush=True))
    result['workload'] = desc
    text = json.dumps(result)
    if args.out:
        Path(args.out).write_text(text + '\n')
    print(text)


if __name__ == '__main__':
    main()
