"""scikit-learn estimators over the exact GPU search (the reference's pynear/sklearn_adapter.py).

Public names, constructor parameters and method contracts follow the reference
(sklearn_adapter.py:102-400): ``PyNearNearestNeighbors`` (``fit``, ``kneighbors``,
``kneighbors_graph``), ``PyNearKNeighborsClassifier`` (``fit``, ``predict``, ``predict_proba``,
``score``) and ``PyNearKNeighborsRegressor`` (``fit``, ``predict``, ``score``), with the metric
names of sklearn_adapter.py:38-57.

The reference asks a VP-tree for lists of lists and re-sorts them (sklearn_adapter.py:64-76);
here the estimators hold the native exact index directly and read its dense, already ascending
``(distance, index)`` arrays, so no per-query Python work remains on the search path.
"""

import numpy as np
from sklearn.base import BaseEstimator, ClassifierMixin, RegressorMixin
from sklearn.utils.multiclass import unique_labels
from sklearn.utils.validation import check_array, check_is_fitted

from . import _native as N

# metric name -> (native metric, input dtype); the aliases of sklearn_adapter.py:38-57
_METRICS = {
    "euclidean": (N.METRIC_L2, np.float32),
    "l2": (N.METRIC_L2, np.float32),
    "manhattan": (N.METRIC_L1, np.float32),
    "l1": (N.METRIC_L1, np.float32),
    "chebyshev": (N.METRIC_CHEBYSHEV, np.float32),
    "linf": (N.METRIC_CHEBYSHEV, np.float32),
    "hamming": (N.METRIC_HAMMING, np.uint8),
}


def _metric_spec(metric):
    spec = _METRICS.get(str(metric).lower())
    if spec is None:
        raise ValueError(f"Unsupported metric '{metric}'. Supported values: {sorted(_METRICS)}")
    return spec


def _weights(dist, weights):
    """Neighbour weights (n_queries, k): uniform, or 1/d with exact matches taking all the weight."""
    if weights == "uniform":
        return np.ones_like(dist)
    if weights != "distance":
        raise ValueError(f"weights must be 'uniform' or 'distance', got '{weights}'")
    exact = dist == 0.0
    with np.errstate(divide="ignore"):
        w = 1.0 / dist
    hit = exact.any(axis=1)
    w[hit] = exact[hit].astype(np.float64)
    return w


class _ExactNeighbors(BaseEstimator):
    """Holds the fitted native index; subclasses add the estimator API."""

    def _build(self, X):
        native_metric, dtype = _metric_spec(self.metric)
        X = np.ascontiguousarray(check_array(X, dtype=dtype))
        index = N.ExactHandle(native_metric)
        index.set(X)
        self._index = index
        self._dtype = dtype
        self.n_samples_fit_ = X.shape[0]
        self.n_features_in_ = X.shape[1]
        return X

    def _query(self, X, k):
        """Ascending (distances float64, indices intp), each (n_queries, min(k, n_samples_fit_))."""
        X = np.ascontiguousarray(check_array(X, dtype=self._dtype))
        idx, dist = self._index.search(X, k)
        return dist.astype(np.float64), idx.astype(np.intp)


class PyNearNearestNeighbors(_ExactNeighbors):
    """Unsupervised exact neighbour lookup (drop-in for sklearn.neighbors.NearestNeighbors)."""

    def __init__(self, n_neighbors: int = 5, metric: str = "euclidean"):
        self.n_neighbors = n_neighbors
        self.metric = metric

    def fit(self, X, y=None):
        self._fit_X = self._build(X)
        return self

    def kneighbors(self, X=None, n_neighbors=None, return_distance=True):
        """(distances, indices) nearest-first; with X=None the training points, self excluded."""
        check_is_fitted(self)
        k = self.n_neighbors if n_neighbors is None else n_neighbors
        if X is None:
            dist, ind = self._query(self._fit_X, k + 1)
            dist, ind = dist[:, 1:], ind[:, 1:]
        else:
            dist, ind = self._query(X, k)
        return (dist, ind) if return_distance else ind

    def kneighbors_graph(self, X=None, n_neighbors=None, mode="connectivity"):
        """CSR matrix (n_queries, n_samples_fit_): 1 per edge, or the distance with mode='distance'."""
        from scipy.sparse import csr_matrix

        dist, ind = self.kneighbors(X, n_neighbors=n_neighbors, return_distance=True)
        nq, k = ind.shape
        rows = np.repeat(np.arange(nq), k)
        vals = np.ones(nq * k) if mode == "connectivity" else dist.ravel()
        return csr_matrix((vals, (rows, ind.ravel())), shape=(nq, self.n_samples_fit_))


class PyNearKNeighborsClassifier(ClassifierMixin, _ExactNeighbors):
    """k-NN classifier (drop-in for sklearn.neighbors.KNeighborsClassifier)."""

    def __init__(self, n_neighbors: int = 5, metric: str = "euclidean", weights: str = "uniform"):
        self.n_neighbors = n_neighbors
        self.metric = metric
        self.weights = weights

    def fit(self, X, y):
        self._y = np.asarray(y)
        self.classes_ = unique_labels(y)
        self._y_code = np.searchsorted(self.classes_, self._y)
        self._build(X)
        return self

    def predict_proba(self, X):
        """(n_queries, n_classes) vote shares, columns in ``classes_`` order, rows summing to 1."""
        check_is_fitted(self)
        dist, ind = self._query(X, self.n_neighbors)
        w = _weights(dist, self.weights)
        nq = ind.shape[0]
        proba = np.zeros((nq, len(self.classes_)))
        np.add.at(proba, (np.repeat(np.arange(nq), ind.shape[1]), self._y_code[ind].ravel()), w.ravel())
        s = proba.sum(axis=1, keepdims=True)
        return proba / np.where(s == 0, 1.0, s)

    def predict(self, X):
        return self.classes_[np.argmax(self.predict_proba(X), axis=1)]


class PyNearKNeighborsRegressor(RegressorMixin, _ExactNeighbors):
    """k-NN regressor (drop-in for sklearn.neighbors.KNeighborsRegressor)."""

    def __init__(self, n_neighbors: int = 5, metric: str = "euclidean", weights: str = "uniform"):
        self.n_neighbors = n_neighbors
        self.metric = metric
        self.weights = weights

    def fit(self, X, y):
        self._y = np.asarray(y, dtype=np.float64)
        self._build(X)
        return self

    def predict(self, X):
        """Weighted mean of the neighbours' targets: (n_queries,) or (n_queries, n_outputs)."""
        check_is_fitted(self)
        dist, ind = self._query(X, self.n_neighbors)
        w = _weights(dist, self.weights)
        ny = self._y[ind]
        if ny.ndim == 2:
            return (w * ny).sum(axis=1) / w.sum(axis=1)
        return (w[:, :, None] * ny).sum(axis=1) / w.sum(axis=1)[:, None]
