"""Autotuner of the scoring launch (no reference counterpart; the reference tuned its
Ascend kernel by hand-edited source variants, tools/capacity_sources.py).

Searches the tile size (docs per scoring task: one wave's LDS accumulator row), the number
of persistent scoring waves, and the per-task cost the host schedule uses to give every wave
equal work, by timing the scoring kernel with HIP events on a sample workload. Candidate
capacity follows from the tile (BlockKeepCap) and K. Every configuration's results are
checked to equal the first one's (the launch never changes results), and the fastest is
returned as Session keyword arguments.

    python -m bm25mi.autotune --corpus synthetic10m --out tuned.json
"""

import argparse
import itertools
import json
import statistics
import sys
from pathlib import Path

import numpy as np

BLOCK_DOCS = (2048, 4096)
WAVE_SCALES = (0.5, 0.75, 1.0, 1.5)
TASK_COSTS = (1024, 2048, 4096, 8192, 16384)


def measure(session, batches, k, repeats):
    """Mean scoring-kernel ms over repeats of every batch, and the results of the last pass."""
    ms, results = [], []
    for _ in range(repeats):
        results = []
        for b in batches:
            scores, ids, _ = session.search(b, k, timed=True)
            ms.append(session.stats()['score_kernel_ms'])
            results.append((scores, ids))
    return statistics.mean(ms), results


def tune(index, batches, k, block_docs=BLOCK_DOCS, wave_scales=WAVE_SCALES, task_costs=TASK_COSTS, repeats=5,
         session_factory=None, log=None):
    """Returns {'best': Session kwargs, 'score_ms': best time, 'table': all timings}."""
    from .search import Session
    make = session_factory or (lambda bd: Session(index, max_queries=max(len(b) for b in batches),
                                                  max_words=max(sum(len(q) for q in b) for b in batches),
                                                  max_k=k, block_docs=bd))
    table, reference, best = [], None, None
    for bd in block_docs:
        s = make(bd)
        try:
            s.search(batches[0], k)
            full = s.stats()['score_grid']
            for scale, cost in itertools.product(wave_scales, task_costs):
                waves = max(1, int(round(full * scale)))
                s.set_launch(waves, cost)
                ms, results = measure(s, batches, k, repeats)
                if reference is None:
                    reference = results
                for (s0, i0), (s1, i1) in zip(reference, results):
                    if not (np.array_equal(i0, i1) and np.array_equal(s0.view(np.uint32), s1.view(np.uint32))):
                        raise AssertionError(f'launch {bd}/{waves}/{cost} changed the results')
                row = {'block_docs': bd, 'grid_blocks': waves, 'task_cost': cost, 'score_ms': ms}
                table.append(row)
                if log:
                    log(json.dumps(row))
                if best is None or ms < best['score_ms']:
                    best = row
        finally:
            s.close()
    return {'best': {k_: best[k_] for k_ in ('block_docs', 'grid_blocks', 'task_cost')},
            'score_ms': best['score_ms'], 'table': table}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument('--corpus', default='synthetic10m')
    ap.add_argument('--queries', type=int, default=8)
    ap.add_argument('--terms', type=int, default=3)
    ap.add_argument('--k', type=int, default=10)
    ap.add_argument('--repeats', type=int, default=5)
    ap.add_argument('--out', default=None)
    args = ap.parse_args()
    from . import corpus as corpus_mod
    c = corpus_mod.generate(args.corpus)
    index = c.index()
    batches = [corpus_mod.draw_queries(c.offsets, args.queries, args.terms, seed=s) for s in (11, 12)]
    result = tune(index, batches, args.k, repeats=args.repeats, log=lambda m: print(m, file=sys.stderr))
    text = json.dumps(result)
    if args.out:
        Path(args.out).write_text(text + '\n')
    print(text)


if __name__ == '__main__':
    main()
