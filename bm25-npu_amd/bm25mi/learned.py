"""Learned K-th-score threshold for safe block pruning (no reference counterpart).

A 3-layer MLP (16 -> 64 -> 64 -> 1, ReLU, FP32) predicts, from per-query statistics of the
query's terms (bm25mi_query_features), log(K-th score / upper), where upper is the FP32 sum of
the terms' largest impacts (an upper bound of every score). Training and inference run on the
GPU through the library's own HIP kernels (bm25mi_mlp_train / bm25mi_mlp_predict); training
pairs are exact Search outputs of the plain GPU path over seeded queries.

The prediction, scaled by a safety factor calibrated on held-out queries and clamped to the
query's provable bounds [lower, upper] (bm25mi_query_bounds), is the query's threshold. Uploaded
to a session (ThresholdModel.upload), the model predicts thresholds on the device inside every
learned Search (Session.search(..., learned=True)). Granules whose upper bound is below the
threshold are skipped and only docs scoring at least it are kept; a query with fewer than K such
docs is rerun on the plain path, so results always equal the plain path.
"""

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .corpus import draw_queries, pack

FEATURES = _lib.FEATURES


def param_count(in_features=FEATURES):
    return int(_lib.load().bm25mi_mlp_param_count(in_features))


def init_params(seed=0, in_features=FEATURES):
    out = np.empty(param_count(in_features), np.float32)
    _lib.check(_lib.load().bm25mi_mlp_init(in_features, C.c_uint64(seed), _lib.ptr(out)))
    return out


def gradient(params, x, target):
    """GPU gradient of the mean squared error (1/n) sum (MLP(x) - target)^2 and its value."""
    x = np.ascontiguousarray(x, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    grad = np.empty_like(np.ascontiguousarray(params, np.float32))
    loss = C.c_float()
    _lib.check(_lib.load().bm25mi_mlp_gradient(_lib.ptr(np.ascontiguousarray(params, np.float32)), _lib.ptr(x),
                                               _lib.ptr(target), len(target), x.shape[1], _lib.ptr(grad),
                                               C.byref(loss)))
    return grad, loss.value


def forward(params, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(len(x), np.float32)
    _lib.check(_lib.load().bm25mi_mlp_predict(_lib.ptr(np.ascontiguousarray(params, np.float32)), _lib.ptr(x),
                                              len(x), x.shape[1], _lib.ptr(y)))
    return y


def train(params, x, target, epochs, batch, lr, seed=0):
    """Adam on the GPU; returns (params, per-epoch mean loss)."""
    params = np.ascontiguousarray(params, np.float32).copy()
    x = np.ascontiguousarray(x, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    losses = np.empty(epochs, np.float32)
    _lib.check(_lib.load().bm25mi_mlp_train(_lib.ptr(params), _lib.ptr(x), _lib.ptr(target), len(target), x.shape[1],
                                            epochs, batch, float(lr), C.c_uint64(seed), _lib.ptr(losses)))
    return params, losses


@dataclass
class TrainingSet:
    queries: list
    ks: np.ndarray          # int32[n]
    features: np.ndarray    # float32[n, FEATURES]
    kth: np.ndarray         # float32[n]: exact K-th score from the plain GPU path
    lower: np.ndarray       # float32[n]: provable lower bound of the K-th score
    upper: np.ndarray       # float32[n]: upper bound of any score


def make_queries(offsets, count, seed, terms=(2, 12), k_choices=(10, 100)):
    """Seeded queries of terms[0]..terms[1] distinct df-weighted terms, with K drawn from k_choices."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(terms[0], terms[1] + 1, size=count)
    ks = rng.choice(np.asarray(k_choices, np.int32), size=count).astype(np.int32)
    return draw_queries(offsets, count, lengths, seed=seed + 1), ks


def collect(session, queries, ks, batch):
    """Exact K-th scores of queries from the plain GPU path, with features and bounds."""
    kth = np.empty(len(queries), np.float32)
    for b0 in range(0, len(queries), batch):
        qs, kb = queries[b0:b0 + batch], ks[b0:b0 + batch]
        scores, _, _ = session.search(qs, kb)
        ends = np.cumsum(kb) - 1
        kth[b0:b0 + len(qs)] = scores[ends]
    feats = np.concatenate([session.features(queries[b0:b0 + batch], ks[b0:b0 + batch])
                            for b0 in range(0, len(queries), batch)])
    lower, upper = zip(*[session.bounds(queries[b0:b0 + batch], ks[b0:b0 + batch])
                         for b0 in range(0, len(queries), batch)])
    return TrainingSet(queries, np.asarray(ks, np.int32), feats, kth, np.concatenate(lower), np.concatenate(upper))


class ThresholdModel:
    """Predicts log(K-th score / upper); thresholds = clamp(safety * upper * exp(prediction), lower, upper)."""

    def __init__(self, seed=0):
        self.params = init_params(seed)
        self.mean = np.zeros(FEATURES, np.float32)
        self.scale = np.ones(FEATURES, np.float32)
        self.y_mean, self.y_scale = 0.0, 1.0
        self.safety = 1.0
        self.losses = np.empty(0, np.float32)

    def _x(self, features):
        return ((np.asarray(features, np.float32) - self.mean) / self.scale).astype(np.float32)

    @staticmethod
    def _usable(data):
        return (data.kth > 0) & (data.upper > 0)

    def fit(self, data, epochs=60, batch=256, lr=2e-3, seed=0):
        """Trains on the queries whose K-th score is positive (the others need no threshold)."""
        keep = self._usable(data)
        x = data.features[keep]
        self.mean = x.mean(axis=0).astype(np.float32)
        self.scale = np.maximum(x.std(axis=0), 1e-6).astype(np.float32)
        y = np.log(data.kth[keep].astype(np.float64) / data.upper[keep].astype(np.float64))
        self.y_mean, self.y_scale = float(y.mean()), float(max(y.std(), 1e-6))
        target = ((y - self.y_mean) / self.y_scale).astype(np.float32)
        self.params, self.losses = train(self.params, self._x(x), target, epochs, min(batch, len(target)), lr, seed)
        return self

    def predict_ratio(self, features):
        """Predicted K-th score / upper bound per query (float64)."""
        z = forward(self.params, self._x(features)).astype(np.float64)
        return np.exp(z * self.y_scale + self.y_mean)

    def thresholds(self, features, lower, upper, safety=None):
        s = self.safety if safety is None else safety
        upper = np.asarray(upper, np.float64)
        t = np.clip(s * upper * self.predict_ratio(features), np.asarray(lower, np.float64), upper)
        return t.astype(np.float32)

    def calibrate(self, data, fallback_rate=0.02):
        """Largest safety factor whose thresholds exceed the true K-th score (forcing a
        plain-path rerun) on at most `fallback_rate` of the held-out queries."""
        keep = self._usable(data)
        ratio = data.kth[keep].astype(np.float64) / (data.upper[keep] * self.predict_ratio(data.features[keep]))
        self.safety = float(np.quantile(ratio, fallback_rate)) if len(ratio) else 1.0
        return self.safety

    def upload(self, session):
        """Makes the model resident on the session's GPU for learned Searches."""
        session.set_threshold_model(self.params, self.mean, self.scale, self.y_mean, self.y_scale, self.safety)


def fit_best(data, select, seeds=(0, 1, 2, 3), fallback_rate=0.005, **fit_args):
    """Trains one model per seed and keeps the one whose thresholds are tightest on held-out
    selection queries: the largest safety factor at the target rerun rate (the thresholds'
    average distance below the K-th scores, at the same share of reruns). The selection queries
    are not the ones the kept model is calibrated on afterwards. Returns the model and one
    (seed, selection safety, last epoch loss) row per candidate."""
    rows, best = [], None
    for seed in seeds:
        model = ThresholdModel(seed=seed).fit(data, seed=seed, **fit_args)
        safety = model.calibrate(select, fallback_rate)
        rows.append((int(seed), float(safety), float(model.losses[-1])))
        if best is None or safety > best[0]:
            best = (safety, model)
    return best[1], rows


def compare(session, queries, k, thresholds):
    """Plain and pruned Search of the same queries: (mismatches, fallback flags, plain stats,
    pruned stats). Mismatches count differing doc IDs or FP32 score bits."""
    offsets, terms = pack(queries)
    s0, i0, _ = session.search((offsets, terms), k, timed=True)
    plain = session.stats()
    s1, i1, fb = session.search((offsets, terms), k, thresholds=thresholds, timed=True)
    pruned = session.stats()
    mismatches = int((i0 != i1).sum()) + int((s0.view(np.uint32) != s1.view(np.uint32)).sum())
    return mismatches, fb, plain, pruned
