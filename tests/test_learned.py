"""The learned K-th-score model: the float64 CPU oracle pinned by finite differences, the
GPU kernels (forward, gradient, Adam training) checked against it, and the learned pruned
Search checked against the plain path (identical results, fewer postings streamed)."""

import numpy as np
import pytest

import bm25mi
from bm25mi import learned
from oracle import mlp_oracle as mo
from oracle import oracle as orc


def random_case(seed, n=300, n_in=16):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(learned.param_count(n_in)) * 0.2).astype(np.float32)
    x = rng.standard_normal((n, n_in)).astype(np.float32)
    target = rng.standard_normal(n).astype(np.float32)
    return params, x, target


def test_oracle_gradient_matches_finite_differences():
    params, x, target = random_case(1, n=40)
    grad, _ = mo.gradient(params, x, target)
    p = params.astype(np.float64)
    rng = np.random.default_rng(2)
    for i in rng.choice(len(p), size=60, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mo.loss(p + e, x, target) - mo.loss(p - e, x, target)) / 2e-6
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), i
    # A planted mistake (a dropped ReLU mask on layer 2) is caught by the same check.
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, x.shape[1])
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    dy = 2.0 * ((np.maximum(h1 @ W2.T + b2, 0) @ W3 + b3) - target) / len(target)
    wrong_gb2 = np.outer(dy, W3).sum(axis=0)
    b2 = 64 * 16 + 64 + 64 * 64  # offset of b2 in the flat layout
    assert not np.allclose(wrong_gb2, grad[b2:b2 + 64])


def test_param_layout_and_init_are_seeded():
    n = learned.param_count()
    assert n == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1
    a, b, c = learned.init_params(3), learned.init_params(3), learned.init_params(4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    W1 = a[:64 * 16]
    assert np.abs(W1).max() <= np.sqrt(6 / 16) + 1e-6 and np.all(a[64 * 16:64 * 17] == 0)


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(need_gpu):
    # Partial last tiles, feature counts other than 16 (up to the kernels' 32), and more tiles
    # than the backward kernel's 512 workgroups (each then accumulates several tiles).
    for seed, n, n_in in ((1, 300, 16), (2, 300, 16), (3, 64, 32), (4, 77, 5), (5, 40_000, 16)):
        params, x, target = random_case(seed, n, n_in)
        y = learned.forward(params, x)
        assert np.allclose(y, mo.forward(params, x), rtol=1e-4, atol=1e-5)
        g, loss = learned.gradient(params, x, target)
        want, want_loss = mo.gradient(params, x, target)
        # FP32 sums of ~300 products against float64: relative error per entry well below 1e-3.
        assert np.allclose(g, want, rtol=1e-3, atol=1e-5 * np.abs(want).max())
        assert abs(loss - want_loss) <= 1e-4 * want_loss


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_function(need_gpu):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((4096, 16)).astype(np.float32)
    target = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
    params, losses = learned.train(learned.init_params(0), x, target, epochs=40, batch=256, lr=3e-3, seed=1)
    assert losses[-1] < 0.1 * losses[0]
    assert mo.loss(params, x, target) < 0.05
    # Same seed, same result: the GPU reduction order is fixed.
    params2, _ = learned.train(learned.init_params(0), x, target, epochs=40, batch=256, lr=3e-3, seed=1)
    assert np.array_equal(params, params2)


@pytest.fixture(scope='module')
def small_corpus():
    corpus = bm25mi.generate(bm25mi.CorpusConfig(300_000, 200_000, 1.07, 56.0, seed=11))
    return corpus, corpus.index()


@pytest.mark.gpu
def test_learned_threshold_prunes_and_keeps_results(small_corpus, need_gpu):
    corpus, index = small_corpus
    with bm25mi.Session(index, max_queries=64, max_words=64 * 12, max_k=100) as s:
        # K values the live path takes (the library scores queries with K above 64 on the
        # plain path, where a threshold changes nothing).
        qs, ks = learned.make_queries(corpus.offsets, 1536, seed=100, k_choices=(10, 50))
        data = learned.collect(s, qs, ks, batch=64)
        model = learned.ThresholdModel(seed=0).fit(data, epochs=60, batch=128, lr=2e-3, seed=0)
        assert model.losses[-1] < model.losses[0]
        vq, vk = learned.make_queries(corpus.offsets, 512, seed=200, k_choices=(10, 50))
        model.calibrate(learned.collect(s, vq, vk, batch=64), fallback_rate=0.02)
        scanned = {'plain': 0, 'learned': 0, 'lower': 0}
        fallbacks = 0
        for k in (10, 50):
            for seed in range(3):
                # Short queries like the headline workload's (3 terms): their learned thresholds
                # clear the live-path bar, where the lower bound alone mostly does not.
                queries = learned.make_queries(corpus.offsets, 64, seed=300 + seed, terms=(3, 3), k_choices=(k,))[0]
                feats = s.features(queries, k)
                lower, upper = s.bounds(queries, k)
                thr = model.thresholds(feats, lower, upper)
                mism, fb, plain, pruned = learned.compare(s, queries, k, thr)
                assert mism == 0
                fallbacks += int(fb.sum())
                scanned['plain'] += plain['postings_scanned']
                scanned['learned'] += pruned['postings_scanned']
                # The method without the network: the provable lower bound alone as threshold.
                mism, fb0, _, base = learned.compare(s, queries, k, lower.astype(np.float32))
                assert mism == 0 and not fb0.any()
                scanned['lower'] += base['postings_scanned']
        assert fallbacks <= 0.1 * 6 * 64
        # The trained threshold streams fewer postings than the plain path and than the
        # provable lower bound used as the threshold without the network.
        assert scanned['learned'] < min(scanned['lower'], scanned['plain']), scanned


@pytest.mark.gpu
def test_fit_best_keeps_the_tightest_model_on_selection_queries(small_corpus, need_gpu):
    """Model selection over training seeds: the kept model is the candidate with the largest
    safety factor on the selection queries, identical to a lone fit from that seed."""
    corpus, index = small_corpus
    with bm25mi.Session(index, max_queries=64, max_words=64 * 12, max_k=100) as s:
        qs, ks = learned.make_queries(corpus.offsets, 1024, seed=100, k_choices=(10, 50))
        data = learned.collect(s, qs, ks, batch=64)
        sq, sk = learned.make_queries(corpus.offsets, 512, seed=300, k_choices=(10, 50))
        select = learned.collect(s, sq, sk, batch=64)
    model, rows = learned.fit_best(data, select, seeds=(0, 1, 2), fallback_rate=0.02, epochs=20, batch=128)
    assert [r[0] for r in rows] == [0, 1, 2]
    assert len({r[1] for r in rows}) > 1  # the seeds give different models
    best = max(rows, key=lambda r: r[1])
    assert model.safety == best[1]
    lone = learned.ThresholdModel(seed=best[0]).fit(data, epochs=20, batch=128, seed=best[0])
    assert np.array_equal(lone.params, model.params)
    assert lone.calibrate(select, 0.02) == best[1]


@pytest.fixture(scope='module')
def trained_small(small_corpus):
    """A model trained on exact K-th scores of the 300K-doc corpus, calibrated on held-out
    queries and made resident on a session."""
    corpus, index = small_corpus
    s = bm25mi.Session(index, max_queries=64, max_words=64 * 16, max_k=1000)
    qs, ks = learned.make_queries(corpus.offsets, 2048, seed=100, k_choices=(10, 50, 100))
    model = learned.ThresholdModel(seed=0).fit(learned.collect(s, qs, ks, batch=64), epochs=40, batch=256, lr=2e-3,
                                               seed=0)
    vq, vk = learned.make_queries(corpus.offsets, 512, seed=200, k_choices=(10, 50, 100))
    model.calibrate(learned.collect(s, vq, vk, batch=64), fallback_rate=0.02)
    model.upload(s)
    yield corpus, index, s, model
    s.close()


@pytest.mark.gpu
def test_device_thresholds_match_float64_oracle(trained_small, need_gpu):
    """The threshold kernel's MLP (inside every learned Search) against the float64 oracle of the
    same model: threshold = clamp(safety * upper * exp(z * y_scale + y_mean), lower, upper), and 0
    for queries routed to the plain path (threshold below 0.3 of the upper bound, or K above 64)."""
    corpus, index, s, model = trained_small
    queries, ks = learned.make_queries(corpus.offsets, 64, seed=300, terms=(1, 12), k_choices=(1, 10, 64, 65, 100))
    queries[0], queries[1] = [], [index.vocab + 5]
    got = s.learned_thresholds(queries, ks)
    feats = s.features(queries, ks)
    lower, upper = s.bounds(queries, ks)
    x = (feats.astype(np.float64) - model.mean) / model.scale
    z = mo.forward(model.params, x)
    want = np.clip(model.safety * upper * np.exp(z * model.y_scale + model.y_mean), lower, upper)
    live = (upper > 0) & (ks <= 64) & (want >= 0.3 * upper)
    margin = np.abs(want - 0.3 * upper) > 1e-3 * upper  # routing decisions far from the boundary
    assert not got[:2].any()
    assert (got[margin & ~live] == 0).all()
    assert np.allclose(got[margin & live], want[margin & live], rtol=1e-4, atol=0)
    assert (live & margin).sum() >= 8 and (~live & margin).sum() >= 8


@pytest.mark.gpu
def test_learned_search_equals_plain_and_prunes(trained_small, need_gpu):
    """Learned Searches (thresholds predicted on the GPU inside the Search) return the plain
    path's results bit for bit, on batches mixing live and plain routing, empty and unknown-term
    queries and K from 1 to 1000, timed and untimed; they stream fewer postings than the plain
    path, and a query rerun on the plain path is flagged."""
    corpus, index, s, model = trained_small
    scanned = {'plain': 0, 'learned': 0}
    for seed in range(4):
        queries, ks = learned.make_queries(corpus.offsets, 48, seed=400 + seed, terms=(1, 12),
                                           k_choices=(1, 10, 50, 100, 1000))
        queries[3] = []
        queries[5] = [index.vocab + 1, queries[5][0]]
        want_s, want_i = orc.search(index, *bm25mi.pack(queries), ks, threads=8)
        for timed in (False, True):
            got_s, got_i, fb = s.search(queries, ks, learned=True, timed=timed)
            assert np.array_equal(got_i, want_i), f'seed {seed} timed={timed}'
            assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32)), f'seed {seed} timed={timed}'
            assert set(np.unique(fb)) <= {0, 1}
        scanned['learned'] += s.stats()['postings_scanned']
        s.search(queries, ks, timed=True)
        scanned['plain'] += s.stats()['postings_scanned']
    assert scanned['learned'] < scanned['plain'], scanned
    # Short queries at K10 (the headline shape): every query takes the live path.
    queries = learned.make_queries(corpus.offsets, 32, seed=500, terms=(3, 3), k_choices=(10,))[0]
    want_s, want_i = orc.search(index, *bm25mi.pack(queries), 10, threads=8)
    got_s, got_i, fb = s.search(queries, 10, learned=True)
    assert np.array_equal(got_i, want_i) and np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32))
    assert (s.learned_thresholds(queries, 10) > 0).mean() > 0.8


@pytest.mark.gpu
def test_learned_search_needs_a_model(small_corpus, need_gpu):
    _, index = small_corpus
    with bm25mi.Session(index, max_queries=2, max_words=8, max_k=10) as s:
        with pytest.raises(bm25mi.Bm25Error) as err:
            s.search([[1, 2]], 10, learned=True)
        assert err.value.code == bm25mi._lib.ILLEGAL_OPERATION
        with pytest.raises(bm25mi.Bm25Error) as err:
            s.set_threshold_model(np.zeros(10, np.float32), np.zeros(16), np.ones(16), 0.0, 1.0, 1.0)
        assert err.value.code == bm25mi._lib.INVALID_PARAM
        s.search([[1, 2]], 10)  # the session stays usable
