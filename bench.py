"""BM25 full-Search benchmark on one MI355X.

Headline workload: the 10,000,000-doc synthetic corpus (vocabulary 1,000,000, Zipf 1.07,
lognormal lengths of mean 56, k1 1.5, b 0.75, seed 0), Q=8 queries of 3 distinct df-weighted
terms, K=10. One step is one full Search of one Q8 batch: host term IDs in, host top-K (doc IDs +
FP32 scores) out, index resident in HBM. The timed steps cycle forward and backward over a
rotation of 8 batches (seeds 1..8, about 880 MB of distinct postings, well above the 256 MB
Infinity Cache); the reference's A/B protocol (batches A and B in AB, BA cycles,
FR/test/ascendhost/bm25/benchmark_common.h:51-107) is timed beside it. Every timed call writes
a poisoned output buffer whose bits are checked against the batch's warm-up result.

Prints one JSON line: QPS (value), ms per Search, the scoring kernel's HBM roofline fraction,
the reference's own CPU Search timed on this host's cores, parity against the CPU oracle, and
(unless disabled) the learned-threshold path against the plain path and the serving loop under
Poisson load. --batched adds the uneven-batch runs on the 8.84M-passage corpus.
"""

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / 'bm25-npu_amd')]

import bm25mi  # noqa: E402
from bm25mi import learned  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def nearest_rank(values, p):
    v = sorted(values)
    return v[max(0, int(np.ceil(p * len(v))) - 1)]


def summarize(ms, queries):
    """median / nearest-rank P95 / QPS = Q*1000/mean (standalone/final_support.py:78-90)."""
    return {'ms_median': statistics.median(ms), 'ms_p95': nearest_rank(ms, 0.95), 'ms_mean': statistics.mean(ms),
            'qps': queries * 1e3 / statistics.mean(ms), 'samples': len(ms)}


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in Path('/proc/cpuinfo').read_text().splitlines():
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def algorithmic_bytes(offsets, terms, k, stats):
    """Bytes a Search must move: 8 B per posting of the query terms ({doc ID, impact}), the
    query packet (8 B per query offset, 4 B per term) and the top-K output (8 B per row)."""
    nq = len(offsets) - 1
    return 8 * stats['postings_total'] + 8 * (nq + 1) + 4 * len(terms) + 8 * nq * k


def measured_traffic(kernel, args):
    """HBM bytes per launch of the scoring kernel from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy_calibration program's readings), when
    they cover this kernel and workload; else None."""
    path = ROOT / 'profiles' / 'kscore_synthetic10m_q8x3_k10.json'
    if not path.exists() or (args.corpus, args.queries, args.terms, args.k) != ('synthetic10m', 8, 3, 10):
        return None, None
    rec = json.loads(path.read_text())
    if rec.get('kernel') != kernel or rec.get('rotation') != args.rotation:
        return None, None
    return rec['traffic_bytes_per_launch'], rec['source']


def ab_protocol(call, batches, warmup=5, cycles=30):
    """The reference's measurement protocol: warm-up A/B cycles, then timed A/B cycles
    (each cycle calls A then B, the next B then A). Returns per-call ms and last results."""
    out = {}
    for c in range(warmup):
        for i in ((0, 1) if c % 2 == 0 else (1, 0)):
            call(batches[i])
    ms = []
    for c in range(cycles):
        for i in ((0, 1) if c % 2 == 0 else (1, 0)):
            t = time.perf_counter()
            out[i] = call(batches[i])
            ms.append((time.perf_counter() - t) * 1e3)
    return ms, out


POISON = 0x7FC0DEAD  # a NaN as FP32 score bits, an invalid doc ID as int32


def cycle_order(n, steps):
    """Batch index per step: forward then backward cycles over n batches (A B, B A, ...)."""
    order = []
    c = 0
    while len(order) < steps:
        order.extend(range(n) if c % 2 == 0 else range(n - 1, -1, -1))
        c += 1
    return order[:steps]


def headline_protocol(session, batches, args):
    """args.warmup untimed Searches, then args.steps timed end-to-end Searches in forward/backward
    cycles over `batches`. Each timed call writes its own output buffer, poisoned beforehand;
    after the timed region every call's doc IDs and FP32 score bits are compared with its
    batch's warm-up result."""
    nq, k = args.queries, args.k
    width = 2 * nq * k + nq
    ref = {}
    for step, b in enumerate(cycle_order(len(batches), max(args.warmup, len(batches)))):
        s, i, _ = session.search(batches[b], k)
        ref.setdefault(b, (s.copy(), i.copy()))
    order = cycle_order(len(batches), args.steps)
    bufs = np.full((args.steps, width), POISON, np.int32)
    samples = []
    t_start = time.perf_counter()
    for step, b in enumerate(order):
        t = time.perf_counter()
        session.search(batches[b], k, out=bufs[step])
        samples.append((time.perf_counter() - t) * 1e3)
    total_s = time.perf_counter() - t_start
    rows = nq * k
    mismatches = 0
    for step, b in enumerate(order):
        got = (bufs[step, :rows].view(np.float32), bufs[step, rows:2 * rows])
        mismatches += same(got, (ref[b][0].ravel(), ref[b][1].ravel()))
    out = summarize(samples, nq)
    out.update({'batches': len(batches), 'bit_checks': len(order), 'mismatches': mismatches, 'total_s': total_s,
                'results': ref})
    return out


def posting_sets(offsets, batches):
    """Posting bytes (8 B per posting of each distinct term) of the batches: their union, the mean
    per batch, A and B's shared bytes and the mean shared between consecutive batches."""
    df = np.diff(offsets)
    sets = [set(np.unique(terms).tolist()) for _, terms in batches]
    mb = lambda ts: 8 * int(df[sorted(ts)].sum()) / 1e6  # noqa: E731
    union = set().union(*sets)
    return {'union_mb': mb(union), 'mean_batch_mb': statistics.mean(mb(x) for x in sets),
            'a_and_b_mb': mb(sets[0] & sets[1]),
            'consecutive_shared_mb': statistics.mean(mb(sets[i] & sets[(i + 1) % len(sets)]) for i in range(len(sets))),
            'infinity_cache_mb': 256}


def same(a, b):
    """Mismatching entries between two (scores, ids) results: doc IDs or FP32 score bits."""
    return int((np.asarray(a[1]) != np.asarray(b[1])).sum()) + int(
        (np.asarray(a[0], np.float32).view(np.uint32) != np.asarray(b[0], np.float32).view(np.uint32)).sum())


def cpu_legs(args, index, packed, gpu_results, rotation):
    from oracle import oracle
    cores = host_cores()
    threads = min(32, cores)
    out = {'host': {'cores': cores, 'model': cpu_model()}}
    # The oracle (restated reference CPU Search at this index's k1): its timing on A/B, and parity
    # of every batch of the rotation.
    port_ms, _ = ab_protocol(
        lambda b: oracle.search(index, b[0], b[1], args.k, shards=threads, threads=threads, chunk=args.queries),
        packed, warmup=2, cycles=args.cpu_cycles)
    out['port'] = summarize(port_ms, args.queries)
    mismatches = sum(same(gpu_results[i], oracle.search(index, b[0], b[1], args.k, shards=threads, threads=threads,
                                                        chunk=args.queries))
                     for i, b in enumerate(rotation))
    out['parity'] = {'mismatches_vs_oracle': mismatches, 'checked_batches': len(rotation),
                     'rule': 'doc IDs and FP32 score bits equal'}
    # The reference's own CPU Search (compiled from its sources into oracle/_ref): both variants.
    if oracle.ref_available() and args.k <= 32:
        for variant in ('plain', 'directory'):
            ref = oracle.RefSession(index, args.k, threads, variant=variant, max_queries=args.queries,
                                    max_words=args.queries * args.terms)
            ms, _ = ab_protocol(lambda b: ref.search(b[0], b[1]), packed, warmup=5, cycles=args.cpu_cycles)
            ref.close()
            out[f'reference_{variant}'] = summarize(ms, args.queries)
    # The bm25s-numba retrieve algorithm restated (dense row per query + heap top-k).
    impacts = index.impacts()
    best = None
    for workers in sorted({w for w in (1, 8, 32) if w <= cores} | {min(8, cores)}):
        ms, _ = ab_protocol(lambda b: oracle.bm25s_search(index, impacts, b[0], b[1], args.k, threads=workers),
                            packed, warmup=1, cycles=max(2, args.cpu_cycles // 6))
        r = summarize(ms, args.queries)
        r['workers'] = workers
        out.setdefault('bm25s_restated', []).append(r)
        best = r if best is None or r['ms_mean'] < best['ms_mean'] else best
    del impacts
    ref = min((out[k] for k in ('reference_plain', 'reference_directory') if k in out),
              key=lambda r: r['ms_mean'], default=None)
    if ref is not None:
        kind, base = 'reference', ref
        sample = (f'{args.cpu_cycles} A/B cycles ({2 * args.cpu_cycles} Q{args.queries} Searches) after 5 warm-up '
                  f'cycles of the same workload; the reference CpuSearchSession / DirectoryCpuSearchSession '
                  f'(faster one) compiled from its sources (oracle/_ref), {threads} OpenMP threads = shards, '
                  f'{cpu_model()}; its arithmetic uses the reference constant k1=1.2, so only its timing is '
                  f'compared; parity is checked against oracle/bm25_oracle.c')
    else:
        kind, base = 'port', out['port']
        sample = (f'{args.cpu_cycles} A/B cycles of Q{args.queries} Searches of the same workload; '
                  f'oracle/bm25_oracle.c (restated reference CPU Search), {threads} threads, {cpu_model()}')
    cpu_baseline = {'value': base['qps'], 'unit': 'queries/s', 'cores': threads, 'kind': kind,
                    'ms_median': base['ms_median'], 'sample': sample}
    return cpu_baseline, out


ROLES = ('gpu', 'cpu_reference', 'cpu_reference_single', 'bm25s_restated')


def role_main(args):
    """One role of the round protocol, in a process of its own: the A and B headline batches in
    AB/BA cycles (args.warmup warm-up cycles, args.cycles timed), each case's median and nearest-rank
    P95 of the per-call time, and the result bits of both cases, written as JSON to args.role_out."""
    corpus = bm25mi.generate(args.corpus)
    index = corpus.index()
    packed = [bm25mi.pack(bm25mi.draw_queries(corpus.offsets, args.queries, args.terms, seed=s)) for s in (1, 2)]
    del corpus
    cores = host_cores()
    closer = None
    comparable = True  # same arithmetic as this library (k1 of the index)
    if args.role == 'gpu':
        sess = bm25mi.Session(index, max_queries=args.queries, max_words=args.queries * args.terms, max_k=args.k,
                              block_docs=args.block_docs)
        call, closer, threads = (lambda b: sess.search(b, args.k)[:2]), sess.close, 0
    else:
        from oracle import oracle
        if args.role == 'bm25s_restated':
            impacts = index.impacts()
            threads = min(8, cores)
            call = lambda b: oracle.bm25s_search(index, impacts, b[0], b[1], args.k, threads=threads)  # noqa: E731
        else:
            threads = 1 if args.role == 'cpu_reference_single' else min(32, cores)
            if oracle.ref_available():
                ref = oracle.RefSession(index, args.k, threads, variant='plain', max_queries=args.queries,
                                        max_words=args.queries * args.terms)
                call, closer = (lambda b: ref.search(b[0], b[1])), ref.close
                comparable = False  # the reference's own constant k1 = 1.2
            else:
                call = lambda b: oracle.search(index, b[0], b[1], args.k, shards=threads, threads=threads,  # noqa: E731
                                               chunk=args.queries)
    ms, res = [], {}
    per = {0: [], 1: []}
    for c in range(args.warmup + args.cycles):
        for i in ((0, 1) if c % 2 == 0 else (1, 0)):
            t = time.perf_counter()
            res[i] = call(packed[i])
            if c >= args.warmup:
                per[i].append((time.perf_counter() - t) * 1e3)
    if closer:
        closer()
    summary = {case: {'median_batch_ms': statistics.median(per[i]), 'p95_batch_ms': nearest_rank(per[i], 0.95),
                      'samples': len(per[i])} for i, case in enumerate('AB')}
    bits = {case: [np.asarray(res[i][0], np.float32).view(np.uint32).ravel().tolist(),
                   np.asarray(res[i][1]).ravel().tolist()] for i, case in enumerate('AB')}
    Path(args.role_out).write_text(json.dumps({'role': args.role, 'threads': threads, 'summary': summary,
                                               'results': bits if comparable else None}))


def gate(rows, candidate, roles):
    """The reference's decision rule (standalone/final_support.py:336-358): per case, the
    candidate's median beats the best of the baseline roles' medians in every round, and the
    median over rounds of its P95 is below theirs."""
    cases = {}
    for case in 'AB':
        base = [min(row[r]['summary'][case]['median_batch_ms'] for r in roles) for row in rows]
        cand = [row[candidate]['summary'][case]['median_batch_ms'] for row in rows]
        bp95 = [min(row[r]['summary'][case]['p95_batch_ms'] for r in roles) for row in rows]
        cp95 = [row[candidate]['summary'][case]['p95_batch_ms'] for row in rows]
        wins = sum(a < b for a, b in zip(cand, base))
        cases[case] = {'baseline_medians_ms': base, 'candidate_medians_ms': cand, 'baseline_p95_ms': bp95,
                       'candidate_p95_ms': cp95, 'wins': wins,
                       'baseline_over_candidate': statistics.median(base) / statistics.median(cand),
                       'passed': wins == len(rows) and statistics.median(cp95) < statistics.median(bp95)}
    return {'target_met': all(v['passed'] for v in cases.values()), 'baseline_roles': list(roles), 'cases': cases}


def rounds_leg(args):
    """The reference's round protocol (standalone/benchmark_final.py:196-285): args.rounds rounds,
    each running every role in a process of its own, in order on even rounds and in reverse order
    on odd ones; then the CPU gate over the rounds and the result bits of every role checked equal
    to the GPU's."""
    import subprocess
    import tempfile
    rows, mismatches = [], 0
    with tempfile.TemporaryDirectory() as tmp:
        for r in range(args.rounds):
            order = list(ROLES) if r % 2 == 0 else list(reversed(ROLES))
            row = {'round': r + 1, 'process_order': order}
            for role in order:
                out = Path(tmp) / f'{r}_{role}.json'
                cmd = [sys.executable, str(ROOT / 'bench.py'), '--role', role, '--role-out', str(out),
                       '--corpus', args.corpus, '--queries', str(args.queries), '--terms', str(args.terms),
                       '--k', str(args.k), '--block-docs', str(args.block_docs), '--warmup', '5',
                       '--cycles', str(args.round_cycles)]
                t0 = time.time()
                subprocess.run(cmd, check=True, timeout=900)
                rec = json.loads(out.read_text())
                row[role] = {'summary': rec['summary'], 'threads': rec['threads'], 'wall_s': round(time.time() - t0, 1),
                             'results': rec.pop('results')}
                log(f'round {r + 1} {role}: ' + json.dumps(rec['summary']))
            rows.append(row)
    compared = set()
    for row in rows:
        for role in ROLES:
            if role != 'gpu' and row[role]['results'] is not None:
                compared.add(role)
                mismatches += sum(a != b for case in 'AB' for part in (0, 1)
                                  for a, b in zip(row[role]['results'][case][part], row['gpu']['results'][case][part]))
    for row in rows:
        for role in ROLES:
            row[role].pop('results')
    return {'rounds': rows, 'cpu_decision': gate(rows, 'gpu', ('cpu_reference', 'cpu_reference_single')),
            'bm25s_decision': gate(rows, 'gpu', ('bm25s_restated',)),
            'results_vs_gpu_mismatches': mismatches, 'roles_compared_with_gpu': sorted(compared),
            'note': 'the reference CPU Search (oracle/_ref) scores with its own k1 = 1.2, so its results are '
                    'left out of the bit comparison; roles with the index\'s k1 are compared'}


def learned_leg(args, index, offsets):
    """Trains the K-th-score MLP on the GPU from exact plain-path outputs, makes it resident on
    the Search sessions and compares learned Searches (thresholds predicted on the GPU inside the
    timed call) with plain Searches of the same held-out batches."""
    t0 = time.time()
    s = bm25mi.Session(index, max_queries=256, max_words=256 * 12, max_k=100, block_docs=args.block_docs)
    qs, ks = learned.make_queries(offsets, args.train_queries, seed=100)
    t_draw = time.time() - t0
    data = learned.collect(s, qs, ks, batch=256)
    del qs
    t_collect = time.time() - t0
    # Selection queries (to pick among the seeds' models) and calibration queries (for the kept
    # model's safety factor): both held out from training and from the timed batches.
    sq, sk = learned.make_queries(offsets, 4096, seed=300)
    sdata = learned.collect(s, sq, sk, batch=256)
    vq, vk = learned.make_queries(offsets, 4096, seed=200)
    vdata = learned.collect(s, vq, vk, batch=256)
    t1 = time.time()
    model, candidates = learned.fit_best(data, sdata, seeds=tuple(range(args.train_models)),
                                         fallback_rate=args.fallback_target, epochs=args.train_epochs,
                                         batch=args.train_batch, lr=2e-3)
    t_fit = time.time() - t1
    model.calibrate(vdata, fallback_rate=args.fallback_target)
    s.close()
    out = {'train_pairs': int(len(data.kth)), 'train_terms': '2-12 df-weighted', 'train_k': [10, 100],
           'draw_s': round(t_draw, 2), 'collect_s': round(t_collect, 2), 'fit_s': round(t_fit, 2),
           'epochs': args.train_epochs, 'batch': args.train_batch,
           'candidates': [{'seed': c[0], 'selection_safety': c[1], 'loss_last': c[2]} for c in candidates],
           'loss_first_last': [float(model.losses[0]), float(model.losses[-1])], 'safety': model.safety,
           'inference': 'on the GPU inside every timed learned Search (threshold kernel before scoring)',
           'cases': []}
    del data

    def run_case(name, batches, k, sess):
        plain_ms, plain_res = ab_protocol(lambda i: sess.search(batches[i], k), [0, 1], warmup=3, cycles=args.cycles)
        prune_ms, prune_res = ab_protocol(lambda i: sess.search(batches[i], k, learned=True), [0, 1], warmup=3,
                                          cycles=args.cycles)
        mism = sum(same(plain_res[i][:2], prune_res[i][:2]) for i in (0, 1))
        fb = sum(int(prune_res[i][2].sum()) for i in (0, 1))
        scan = {'plain': 0, 'learned': 0}
        dev = {'plain': [], 'learned': []}
        kern = {'plain': [], 'learned': []}  # timed scoring kernels: KScore / KThreshold .. KLive
        for rep in range(5):
            for i in (0, 1):
                sess.search(batches[i], k, timed=True)
                st = sess.stats()
                scan['plain'] += st['postings_scanned'] if rep == 0 else 0
                dev['plain'].append(st['device_ms'])
                kern['plain'].append(st['score_kernel_ms'])
                sess.search(batches[i], k, learned=True, timed=True)
                st = sess.stats()
                scan['learned'] += st['postings_scanned'] if rep == 0 else 0
                dev['learned'].append(st['device_ms'])
                kern['learned'].append(st['score_kernel_ms'])
        nq = len(batches[0])
        p, lr = summarize(plain_ms, nq), summarize(prune_ms, nq)
        kp, kl = statistics.median(kern['plain']), statistics.median(kern['learned'])
        out['cases'].append({
            'case': name, 'queries': nq, 'k': k, 'plain_ms_median': p['ms_median'],
            'learned_ms_median': lr['ms_median'], 'speedup': p['ms_median'] / lr['ms_median'],
            'plain_qps': p['qps'], 'learned_qps': lr['qps'],
            'plain_device_ms': statistics.median(dev['plain']), 'learned_device_ms': statistics.median(dev['learned']),
            'mismatches': mism, 'fallback_rate': fb / (2 * nq),
            'postings_plain': scan['plain'], 'postings_learned': scan['learned'],
            'bytes_skipped_frac': 1 - scan['learned'] / max(1, scan['plain']),
            # HBM roofline of the scoring kernels: 8 B per posting streamed (the two A/B batches)
            # over their timed duration; 'effective' counts the plain path's postings instead.
            'roofline': {
                'plain_kernels_ms': kp, 'learned_kernels_ms': kl,
                'plain_frac': 8 * scan['plain'] / 2 / (kp * 1e-3) / 1e9 / HBM_PEAK_GBS,
                'learned_frac': 8 * scan['learned'] / 2 / (kl * 1e-3) / 1e9 / HBM_PEAK_GBS,
                'learned_effective_frac': 8 * scan['plain'] / 2 / (kl * 1e-3) / 1e9 / HBM_PEAK_GBS}})
        log(f'learned {name}: plain {p["ms_median"]:.4f} ms, learned {lr["ms_median"]:.4f} ms, '
            f'{p["ms_median"] / lr["ms_median"]:.2f}x, device {kp:.4f}/{kl:.4f} ms, mismatches {mism}, fb {fb}')

    for nq in (8, 64):
        sess = bm25mi.Session(index, max_queries=nq, max_words=nq * 12, max_k=100, block_docs=args.block_docs,
                              live_max_k=args.live_max_k)
        model.upload(sess)
        if nq == 8:
            run_case('Q8x3 K10 (headline batches)', [bm25mi.draw_queries(offsets, 8, 3, seed=1),
                                                    bm25mi.draw_queries(offsets, 8, 3, seed=2)], 10, sess)
        for k in (10, 100):
            batches = [learned.make_queries(offsets, nq, seed=900 + 10 * nq + k + j, k_choices=(k,))[0]
                       for j in (0, 1)]
            run_case(f'Q{nq} x 2-12 terms K{k} (held out)', batches, k, sess)
        sess.close()
    return out, model


def serving_leg(args, index, offsets, model):
    """Poisson arrivals through the continuous-batching server, plain and learned side by side.
    Learned requests have their thresholds predicted on the GPU in the batch they join; every
    learned request's rows are compared with the plain run's."""
    s = bm25mi.Session(index, max_queries=256, max_words=256 * 32, max_k=1000, slots=4,
                       pool_entries=args.pool_entries, block_docs=args.block_docs, live_max_k=args.live_max_k)
    if model is not None:
        model.upload(s)
    out = {'slots': 4, 'max_batch': 256, 'pool_entries_per_slot': s.config.pool_entries or 'full',
           'terms': 'uniform 1-32, df-weighted', 'k_mix': {'10': 0.7, '100': 0.25, '1000': 0.05}, 'runs': []}
    rng = np.random.default_rng(55)
    for qps in args.serving_qps:
        n = int(min(args.serving_requests, max(500, qps * args.serving_seconds)))
        lengths = rng.integers(1, 33, size=n)
        ks = rng.choice(np.array([10, 100, 1000], np.int32), size=n, p=[0.7, 0.25, 0.05]).astype(np.int32)
        queries = bm25mi.draw_queries(offsets, n, lengths, seed=int(qps))
        # The plain path streams every posting of a request's terms (8 B each): the HBM traffic
        # the offered rate asks for, against the 8 TB/s peak.
        df = np.diff(offsets)
        post = float(np.mean([df[q].sum() for q in queries]))
        plain = None
        for mode in (['plain', 'learned'] if model is not None else ['plain']):
            with bm25mi.Server(s, max_batch=256) as srv:
                r = srv.run_poisson(queries, ks, qps=qps, seed=int(qps) + 1, keep_results=True,
                                    learned=(mode == 'learned'))
                c = srv.counters()
            rec = {'mode': mode, 'offered_qps': qps, 'requests': n, 'sustained_qps': r['sustained_qps'],
                   'p50_ms': r['p50_ms'], 'p95_ms': r['p95_ms'], 'p99_ms': r['p99_ms'], 'batches': c['batches'],
                   'mean_batch': c['queries'] / max(1, c['batches']), 'max_inflight': c['max_inflight'],
                   'fallbacks': c['fallbacks'], 'postings_per_request': post,
                   'plain_gbs_at_offered': post * 8 * qps / 1e9,
                   'plain_gbs_at_sustained': post * 8 * r['sustained_qps'] / 1e9,
                   'hbm_peak_gbs': HBM_PEAK_GBS}
            if mode == 'plain':
                plain = (r['scores'], r['ids'])
            else:
                rec['mismatches_vs_plain'] = same((r['scores'], r['ids']), plain)
                rec['requests_checked'] = n
            out['runs'].append(rec)
            log(f'serving {mode} {qps} QPS: sustained {r["sustained_qps"]:.0f}, p50 {r["p50_ms"]:.2f} ms, '
                f'p99 {r["p99_ms"]:.2f} ms')
    s.close()
    return out


def batched_leg(args):
    """Uneven batches on the 8.84M-passage corpus: Q in {1, 8, 64, 256}, geometric 1-32 terms
    (mean 6), K drawn from {10, 100, 1000}; parity against the oracle for one batch per Q."""
    from oracle import oracle
    corpus = bm25mi.generate('passages8p8m')
    index = corpus.index()
    offsets = corpus.offsets
    del corpus
    s = bm25mi.Session(index, max_queries=256, max_words=256 * 32, max_k=1000, block_docs=args.block_docs)
    out = {'corpus': 'passages8p8m', 'docs': index.num_docs, 'postings': index.postings, 'runs': []}
    rng = np.random.default_rng(77)
    for nq in (1, 8, 64, 256):
        batches = []
        for j in (0, 1):
            lengths = bm25mi.corpus.geometric_lengths(nq, 6, 1, 32, seed=1000 * nq + j)
            ks = rng.choice(np.array([10, 100, 1000], np.int32), size=nq).astype(np.int32)
            batches.append((bm25mi.draw_queries(offsets, nq, lengths, seed=2000 * nq + j), ks))
        ms, res = ab_protocol(lambda i: s.search(batches[i][0], batches[i][1]), [0, 1], warmup=3,
                              cycles=args.cycles)
        want = oracle.search(index, *bm25mi.pack(batches[0][0]), batches[0][1], threads=host_cores())
        r = summarize(ms, nq)
        r.update({'queries': nq, 'mismatches_vs_oracle': same(res[0][:2], want)})
        out['runs'].append(r)
        log(f'batched Q{nq}: median {r["ms_median"]:.3f} ms, {r["qps"]:.0f} QPS')
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--corpus', default='synthetic10m')
    ap.add_argument('--queries', type=int, default=8)
    ap.add_argument('--terms', type=int, default=3)
    ap.add_argument('--k', type=int, default=10)
    ap.add_argument('--rotation', type=int, default=8, help='headline batches timed in rotation (seeds 1..N)')
    ap.add_argument('--block-docs', type=int, default=0, help='docs per scoring task (0 = library default)')
    ap.add_argument('--grid', type=int, default=0, help='persistent scoring waves (0 = fill the GPU)')
    ap.add_argument('--task-cost', type=int, default=0, help='per-task cost of the wave schedule (0 = default)')
    ap.add_argument('--cpu-cycles', type=int, default=30, help='timed A/B cycles of each CPU baseline')
    ap.add_argument('--cycles', type=int, default=30, help='timed A/B cycles of the learned and batched legs')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-learned', action='store_true')
    ap.add_argument('--no-serving', action='store_true')
    ap.add_argument('--batched', action='store_true', help='add the 8.84M-passage uneven-batch runs')
    ap.add_argument('--train-queries', type=int, default=1_000_000, help='exact (query, K-th score) training pairs')
    ap.add_argument('--train-epochs', type=int, default=20)
    ap.add_argument('--train-models', type=int, default=4,
                    help='models trained from seeds 0..N-1; the tightest on held-out selection queries is kept')
    ap.add_argument('--train-batch', type=int, default=1024)
    ap.add_argument('--live-max-k', type=int, default=0, help='largest K scored on the live path (0 = library default)')
    ap.add_argument('--fallback-target', type=float, default=0.005,
                    help='held-out rate of queries whose threshold exceeds the K-th score (plain reruns)')
    ap.add_argument('--serving-qps', type=float, nargs='*', default=[2000.0, 10000.0, 40000.0])
    ap.add_argument('--serving-requests', type=int, default=6000)
    ap.add_argument('--serving-seconds', type=float, default=1.0)
    ap.add_argument('--pool-entries', type=int, default=1 << 26, help='candidate-pool entries per serving slot')
    ap.add_argument('--profile', action='store_true', help='short run for rocprofv3: the rotation only')
    ap.add_argument('--rounds', type=int, default=0,
                    help='add the round protocol: this many rounds of the GPU and CPU roles, one process each')
    ap.add_argument('--round-cycles', type=int, default=30, help='timed A/B cycles per role and round')
    ap.add_argument('--role', choices=ROLES, help=argparse.SUPPRESS)
    ap.add_argument('--role-out', help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.role:
        role_main(args)
        return
    if args.gpus != 1:
        print(json.dumps({'error': 'the library runs on one GPU; sharding a Search across GPUs is not implemented',
                          'n_gpus': args.gpus}))
        sys.exit(2)

    t0 = time.time()
    corpus = bm25mi.generate(args.corpus)
    index = corpus.index()
    offsets = corpus.offsets
    log(f'corpus {args.corpus}: N={index.num_docs} V={index.vocab} M={index.postings} ({time.time() - t0:.1f}s)')
    # Headline batches: A and B (seeds 1, 2) and the rotation of args.rotation batches (seeds 1..R,
    # A and B among them) whose union of postings is well above the 256 MB Infinity Cache.
    rotation = [bm25mi.pack(bm25mi.draw_queries(offsets, args.queries, args.terms, seed=s))
                for s in range(1, max(2, args.rotation) + 1)]
    packed = rotation[:2]
    data = posting_sets(offsets, rotation)
    del corpus

    t0 = time.time()
    session = bm25mi.Session(index, max_queries=args.queries, max_words=args.queries * args.terms, max_k=args.k,
                             block_docs=args.block_docs, grid_blocks=args.grid, task_cost=args.task_cost)
    log(f'session create {time.time() - t0:.1f}s')

    # Reference protocol on A/B, then the rotation (whose figures are the headline's): warm-up,
    # then K timed end-to-end Searches (host IDs -> host top-K) in forward/backward cycles, every
    # call into its own poisoned output buffer, all bit-checked after the timed region.
    # A profiling run (--profile) times the rotation only, so that a trace's dispatches are its.
    ab = None
    if not args.profile:
        ab = headline_protocol(session, packed, args)
        log(f'A/B: {args.steps} Searches, median {ab["ms_median"]:.4f} ms, bit checks {ab["bit_checks"]}, '
            f'mismatches {ab["mismatches"]}')
        ab.pop('results')
        ab.pop('total_s')
    rot = headline_protocol(session, rotation, args)
    log(f'rotation of {len(rotation)}: {args.steps} Searches, median {rot["ms_median"]:.4f} ms, '
        f'bit checks {rot["bit_checks"]}, mismatches {rot["mismatches"]}')
    results = rot.pop('results')
    total_s = rot.pop('total_s')

    # Scoring-kernel time with HIP events on the launch stream (direct launches), per batch of the
    # rotation (and of A/B for comparison).
    def kernel_times(batches, calls):
        kernel_ms, bytes_per, final_ms, device_ms = [], [], [], []
        for step in range(calls):
            offs, terms = batches[step % len(batches)]
            session.search(batches[step % len(batches)], args.k, timed=True)
            st = session.stats()
            kernel_ms.append(st['score_kernel_ms'])
            final_ms.append(st['final_ms'])
            device_ms.append(st['device_ms'])
            bytes_per.append(algorithmic_bytes(offs, terms, args.k, st))
        return kernel_ms, bytes_per, final_ms, device_ms

    calls = max(4 * len(rotation), min(args.steps, 100))
    kernel_ms, bytes_per, final_ms, device_ms = kernel_times(rotation, calls)
    k_ms = statistics.mean(kernel_ms)
    b_avg = statistics.mean(bytes_per)
    achieved = b_avg / (k_ms * 1e-3) / 1e9
    ab_frac = ab_kernel_ms = None
    if not args.profile:
        ab_kernel, ab_bytes, _, _ = kernel_times(packed, max(20, min(args.steps, 100)))
        ab_kernel_ms = statistics.mean(ab_kernel)
        ab_frac = statistics.mean(ab_bytes) / (ab_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    session.search(packed[0], args.k, timed=True)
    st = session.stats()
    call_stats = []
    for step in range(20):
        session.search(packed[step % 2], args.k)
        call_stats.append(session.stats())
    import ctypes as C
    cus, lds_block, lds_cu = C.c_int32(), C.c_int32(), C.c_int32()
    bm25mi._lib.check(bm25mi._lib.load().bm25mi_device_info(0, C.byref(cus), C.byref(lds_block), C.byref(lds_cu)))

    traffic, traffic_note = measured_traffic(f'KScore<{st["block_docs"]}>', args)
    record = {
        'metric': 'Q8/K10 full-Search ms @10M docs; QPS; % HBM roofline; learned vs plain',
        'value': args.queries * args.steps / total_s,  # the rotation's timed Searches
        'unit': 'queries/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * total_s / args.steps,
        'ms_median': rot['ms_median'],
        'ms_p95': rot['ms_p95'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': f'{args.corpus} full Search, Q{args.queries}x{args.terms} terms, K{args.k}',
                   'docs': index.num_docs, 'vocab': index.vocab, 'postings': index.postings,
                   'queries': args.queries, 'terms_per_query': args.terms, 'k': args.k,
                   'block_docs': st['block_docs'], 'k1': index.k1, 'b': index.b,
                   'batches': f'rotation of {len(rotation)} Q{args.queries} batches (seeds 1..{len(rotation)})'},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic, 'traffic_source': traffic_note,
                     'kernel': f'KScore<{st["block_docs"]}>', 'kernel_ms': k_ms, 'algorithmic_bytes': b_avg,
                     'launch_waves': st['score_grid'], 'batches': 'rotation',
                     'ab_frac': ab_frac, 'ab_kernel_ms': ab_kernel_ms},
        'protocol': {'rotation': rot, 'ab': ab, 'postings': data,
                     'rule': 'forward/backward cycles over the batches (AB, BA, ... for A/B); every timed '
                             'call writes its own output buffer, pre-filled with a NaN/invalid-ID pattern, '
                             'and its doc IDs and FP32 score bits are checked equal to that batch\'s '
                             'warm-up result after the timed region'},
        'postings_per_search': st['postings_total'],
        'search_breakdown': {
            'device_ms': statistics.mean(device_ms),  # timed (direct) launches
            'host_call_ms': statistics.mean(g['host_ms'] for g in call_stats),
            'candidates_per_search': statistics.mean(g['candidates'] for g in call_stats),
            'tasks_per_search': st['blocks_scored'] + st['blocks_skipped'],
            'final_kernel_ms': statistics.mean(final_ms),
            'postings_streamed': st['postings_scanned'],
            'postings_in_range': st['postings_in_range'],
            'emitting_tasks': st['emitting_blocks'],
            'exact_tasks': st['exact_blocks'],
        },
        'device': {'compute_units': cus.value, 'lds_per_block': lds_block.value, 'lds_per_cu': lds_cu.value},
    }
    session.close()

    if not args.profile and not args.no_cpu:
        t0 = time.time()
        record['cpu_baseline'], cpu = cpu_legs(args, index, packed, results, rotation)
        record['parity'] = cpu.pop('parity')
        record['cpu_baselines'] = cpu
        log(f'CPU legs {time.time() - t0:.1f}s')
    model = None
    if not args.profile and not args.no_learned:
        t0 = time.time()
        record['learned'], model = learned_leg(args, index, offsets)
        log(f'learned leg {time.time() - t0:.1f}s')
    if not args.profile and not args.no_serving:
        t0 = time.time()
        record['serving'] = serving_leg(args, index, offsets, model)
        log(f'serving leg {time.time() - t0:.1f}s')
    if args.rounds and not args.profile:
        t0 = time.time()
        record['rounds'] = rounds_leg(args)
        log(f'round protocol {time.time() - t0:.1f}s')
    if args.batched and not args.profile:
        del index
        t0 = time.time()
        record['batched'] = batched_leg(args)
        log(f'batched leg {time.time() - t0:.1f}s')
    print(json.dumps(record), flush=True)


if __name__ == '__main__':
    main()
