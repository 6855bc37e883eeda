"""scikit-learn adapters over the GPU exact search, checked against sklearn's brute force.

Covers the behaviours the reference tests for its adapters (pynear/tests/test_sklearn_adapter.py:
52-236): indices and distances equal to brute force, return_distance=False, self exclusion when
X is None, n_neighbors override, kneighbors_graph shape / values, metric aliases, unfitted and
unsupported-metric errors, classifier and regressor predictions (uniform and distance weights,
multi-output), probabilities summing to one.
"""

import numpy as np
import pytest

sklearn = pytest.importorskip("sklearn")
from sklearn.exceptions import NotFittedError  # noqa: E402
from sklearn.neighbors import KNeighborsClassifier, KNeighborsRegressor, NearestNeighbors  # noqa: E402

from pynear_amd.sklearn_adapter import (PyNearKNeighborsClassifier, PyNearKNeighborsRegressor,  # noqa: E402
                                        PyNearNearestNeighbors)


def _data(n=600, nq=80, d=12, seed=0):
    rng = np.random.default_rng(seed)
    return rng.random((n, d), dtype=np.float32), rng.random((nq, d), dtype=np.float32)


# ------------------------------------------------------------------ host-only checks
def test_unsupported_metric_raises():
    X, _ = _data(20, 2, 4)
    with pytest.raises(ValueError):
        PyNearNearestNeighbors(metric="cosine").fit(X)


def test_not_fitted_raises():
    X, Q = _data(20, 2, 4)
    with pytest.raises(NotFittedError):
        PyNearNearestNeighbors().kneighbors(Q)


def test_invalid_weights_raises_before_search():
    from pynear_amd.sklearn_adapter import _weights
    with pytest.raises(ValueError):
        _weights(np.ones((2, 3)), "gaussian")


def test_distance_weights_exact_match_takes_all():
    from pynear_amd.sklearn_adapter import _weights
    w = _weights(np.array([[0.0, 1.0, 2.0], [1.0, 2.0, 4.0]]), "distance")
    assert np.array_equal(w[0], [1.0, 0.0, 0.0])
    assert np.allclose(w[1], [1.0, 0.5, 0.25])


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("metric,skm", [("euclidean", "euclidean"), ("l2", "euclidean"), ("manhattan", "manhattan"),
                                        ("l1", "manhattan"), ("chebyshev", "chebyshev"), ("linf", "chebyshev")])
def test_kneighbors_match_sklearn_brute(metric, skm):
    X, Q = _data()
    ours = PyNearNearestNeighbors(n_neighbors=7, metric=metric).fit(X)
    ref = NearestNeighbors(n_neighbors=7, metric=skm, algorithm="brute").fit(X.astype(np.float64))
    d0, i0 = ours.kneighbors(Q)
    d1, i1 = ref.kneighbors(Q.astype(np.float64))
    assert i0.shape == (len(Q), 7) and d0.dtype == np.float64
    assert np.array_equal(i0, i1)  # continuous data: no ties
    np.testing.assert_allclose(d0, d1, rtol=1e-5, atol=1e-6)
    assert np.all(np.diff(d0, axis=1) >= 0)


@pytest.mark.gpu
def test_return_distance_false_and_override():
    X, Q = _data()
    nn = PyNearNearestNeighbors(n_neighbors=4).fit(X)
    ind = nn.kneighbors(Q, return_distance=False)
    assert ind.shape == (len(Q), 4)
    assert nn.kneighbors(Q, n_neighbors=9)[1].shape == (len(Q), 9)


@pytest.mark.gpu
def test_kneighbors_without_query_excludes_self():
    X, _ = _data(300, 1)
    nn = PyNearNearestNeighbors(n_neighbors=3).fit(X)
    d, i = nn.kneighbors()
    assert i.shape == (300, 3)
    assert not np.any(i == np.arange(300)[:, None])
    assert np.all(d > 0)


@pytest.mark.gpu
def test_kneighbors_graph():
    X, Q = _data()
    nn = PyNearNearestNeighbors(n_neighbors=5).fit(X)
    G = nn.kneighbors_graph(Q)
    assert G.shape == (len(Q), len(X)) and G.nnz == len(Q) * 5
    assert np.all(G.data == 1.0)
    Gd = nn.kneighbors_graph(Q, mode="distance")
    assert np.all(Gd.data > 0)
    d, i = nn.kneighbors(Q)
    np.testing.assert_allclose(Gd[np.arange(len(Q)), i[:, 0]].A1, d[:, 0])


@pytest.mark.gpu
def test_hamming_metric_matches_oracle():
    import oracle
    rng = np.random.default_rng(3)
    X = rng.integers(0, 256, (500, 32), dtype=np.uint8)
    Q = rng.integers(0, 256, (40, 32), dtype=np.uint8)
    d, i = PyNearNearestNeighbors(n_neighbors=6, metric="hamming").fit(X).kneighbors(Q)
    ri, rd = oracle.knn_hamming(X, Q, 6)
    assert np.array_equal(i, ri)
    assert np.array_equal(d, rd.astype(np.float64))


@pytest.mark.gpu
@pytest.mark.parametrize("weights", ["uniform", "distance"])
def test_classifier_matches_sklearn(weights):
    X, Q = _data(800, 100, 6, seed=4)
    y = (X[:, 0] + X[:, 1] > 1.0).astype(int) + 2 * (X[:, 2] > 0.5)
    ours = PyNearKNeighborsClassifier(n_neighbors=5, weights=weights).fit(X, y)
    ref = KNeighborsClassifier(n_neighbors=5, weights=weights, algorithm="brute").fit(X.astype(np.float64), y)
    assert np.array_equal(ours.classes_, ref.classes_)
    p0 = ours.predict_proba(Q)
    np.testing.assert_allclose(p0, ref.predict_proba(Q.astype(np.float64)), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(p0.sum(axis=1), 1.0)
    assert np.array_equal(ours.predict(Q), ref.predict(Q.astype(np.float64)))
    assert ours.score(X, y) > 0.8


@pytest.mark.gpu
@pytest.mark.parametrize("weights", ["uniform", "distance"])
def test_regressor_matches_sklearn(weights):
    X, Q = _data(800, 100, 6, seed=5)
    y = np.sin(4 * X[:, 0]) + X[:, 1]
    ours = PyNearKNeighborsRegressor(n_neighbors=6, weights=weights).fit(X, y)
    ref = KNeighborsRegressor(n_neighbors=6, weights=weights, algorithm="brute").fit(X.astype(np.float64), y)
    np.testing.assert_allclose(ours.predict(Q), ref.predict(Q.astype(np.float64)), rtol=1e-5, atol=1e-7)
    assert ours.score(X, y) > 0.5


@pytest.mark.gpu
def test_regressor_multi_output():
    X, Q = _data(400, 30, 5, seed=6)
    Y = np.stack([X[:, 0], 2 * X[:, 1]], axis=1)
    ours = PyNearKNeighborsRegressor(n_neighbors=4).fit(X, Y)
    ref = KNeighborsRegressor(n_neighbors=4, algorithm="brute").fit(X.astype(np.float64), Y)
    out = ours.predict(Q)
    assert out.shape == (30, 2)
    np.testing.assert_allclose(out, ref.predict(Q.astype(np.float64)), rtol=1e-5, atol=1e-7)
