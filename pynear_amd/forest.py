"""IVFFlatL2Index: the interface of pynear's IVF index over L2 (pynear/forest.py:25-199), with its
k-means and its search on the GPU.

The reference keeps per-cluster numpy arrays and searches them with a Python loop of BLAS
GEMVs (forest.py:63-111). This class keeps the same public surface and the same host-side
attributes (``_centroids``, ``_cluster_data``, ``_cluster_norms``, ``_orig_indices``, so default
pickling works as in the reference) and hands search to the device index ``pn_ivfl2``:
coarse top-n_probe clusters, a list-major MFMA scan of the probed clusters and an exact rerank.
Results are ascending squared L2 distances with (distance, id) order.

The k-means is ``kmeans_l2(data, n_clusters, 100, 42)``, the reference's fast path
(forest.py:136-143), computed on the GPU with results identical to the reference's. Empty clusters
are left out of the search, as the reference leaves them out of its per-cluster arrays.
"""

import numpy as np

from . import _native as N


class _IVFFlatIndex:
    def __init__(self, n_clusters: int, n_probe: int):
        self._n_clusters = n_clusters
        self._n_probe = min(n_probe, n_clusters)
        self._centroids = None       # (C, D) float32, every k-means centroid
        self._cluster_data = []      # list of (n_c, D) float32 arrays, non-empty clusters
        self._cluster_norms = []     # list of (n_c,) float32 squared norms
        self._orig_indices = []      # list of (n_c,) global row ids
        self._list_centroids = None  # centroid id of each non-empty cluster
        self.precision = "bf16"      # candidate distances: "bf16", "bf16x3" or "fp8"; rerank is fp32
        self._dev = None

    # ── Build ──────────────────────────────────────────────────────────────
    def set(self, data: np.ndarray) -> None:
        data = np.ascontiguousarray(np.asarray(data, dtype=np.float32))
        if data.ndim != 2:
            raise ValueError("data must be a 2-D array of shape (N, D)")
        n = len(data)
        n_clusters = min(self._n_clusters, n)
        dev = N.IvfL2Handle()
        labels, centroids = dev.build(data, n_clusters, 100, 42)
        self._centroids = centroids
        self._cluster_data, self._cluster_norms, self._orig_indices = [], [], []
        keep = []
        for c in range(len(centroids)):
            idx = np.where(labels == c)[0]
            if not len(idx):
                continue
            pts = data[idx]
            keep.append(c)
            self._cluster_data.append(pts)
            self._cluster_norms.append((pts * pts).sum(axis=1))
            self._orig_indices.append(idx)
        self._list_centroids = np.asarray(keep, np.int64)
        self._dev = dev

    # ── Search ─────────────────────────────────────────────────────────────
    def _device(self):
        if self._dev is None:  # after unpickling: rebuild the device index from the host arrays
            n = sum(len(i) for i in self._orig_indices)
            d = self._centroids.shape[1]
            rows = np.empty((n, d), np.float32)
            labels = np.empty(n, np.int32)
            for j, (pts, ids) in enumerate(zip(self._cluster_data, self._orig_indices)):
                rows[ids] = pts
                labels[ids] = j
            cent = np.ascontiguousarray(self._centroids[self._list_centroids], dtype=np.float32)
            dev = N.IvfL2Handle()
            dev.set(rows, cent, labels)
            self._dev = dev
        return self._dev

    def searchKNN(self, queries: np.ndarray, k: int):
        """Return ``(indices, distances)`` for the k nearest neighbours of each query."""
        queries = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
        if queries.ndim == 1:
            queries = queries[np.newaxis]
        if self._centroids is None:
            raise RuntimeError("Index is empty — call set() first")
        n_probe = min(self._n_probe, len(self._cluster_data))
        idx, dist, cnt = self._device().search(queries, n_probe, k, self.precision)
        return ([idx[i, :c].tolist() for i, c in enumerate(cnt)],
                [dist[i, :c].tolist() for i, c in enumerate(cnt)])

    def search1NN(self, queries: np.ndarray):
        """Shortcut for k=1 nearest neighbour search."""
        return self.searchKNN(queries, 1)

    # ── Info ───────────────────────────────────────────────────────────────
    @property
    def n_clusters(self) -> int:
        """Number of clusters actually built (may be less than requested)."""
        return len(self._cluster_data)

    @property
    def n_probe(self) -> int:
        return self._n_probe

    def __getstate__(self):
        state = dict(self.__dict__)
        state["_dev"] = None
        return state


class IVFFlatL2Index(_IVFFlatIndex):
    """IVF-style approximate index over L2 (Euclidean) distance, searched on the GPU.

    Parameters
    ----------
    n_clusters : int
        Number of Voronoi cells (``int(sqrt(N))`` is a common start).
    n_probe : int
        Cells probed per query; ``n_probe == n_clusters`` is exact.
    """

    def __init__(self, n_clusters: int = 100, n_probe: int = 10):
        super().__init__(n_clusters, n_probe)
