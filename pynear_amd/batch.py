"""Batched exact L2 search over many independent indices (one graph replay per call).

``BatchSearch([(index, queries, k), ...])`` plans every sub-search at once: the queries are
uploaded once, the device outputs are allocated once, and each ``run()`` replays the plan (one
split, one persistent scan over all tiles of all indices, one merge / rerank / fallback launch).
Each sub-search returns exactly what ``index.searchKNN(queries, k)`` returns (same neighbours, same
distance bits), ascending instead of farthest-first when read through ``results()``.
"""

import numpy as np

from . import _native as N
from . import hipdev
from ._pynear import VPTreeL2Index


class BatchSearch:
    def __init__(self, searches, tune=False, table_path=None):
        self._dev = []
        handles, qptrs, nqs, ks, oi, od = [], [], [], [], [], []
        self._shapes = []
        for index, queries, k in searches:
            if not isinstance(index, VPTreeL2Index):
                raise TypeError("BatchSearch serves VPTreeL2Index sub-searches")
            q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
            if q.ndim != 2:
                raise ValueError("queries must be a 2-D array")
            n, d = index._h.info()
            if n == 0:
                raise RuntimeError("index must be first initialized with .set() function and non empty dataset")
            if q.shape[1] != d:
                raise ValueError("query dimension does not match the index")
            kk = min(int(k), n)
            dq = hipdev.DeviceArray.from_numpy(q)
            di = hipdev.DeviceArray((q.shape[0], kk), np.int32)
            dd = hipdev.DeviceArray((q.shape[0], kk), np.float32)
            self._dev += [dq, di, dd]
            handles.append(index._h)
            qptrs.append(dq.ptr)
            nqs.append(q.shape[0])
            ks.append(int(k))
            oi.append(di.ptr)
            od.append(dd.ptr)
            self._shapes.append((di, dd))
        self._plan = N.BatchHandle(handles, qptrs, nqs, ks, oi, od)
        if tune:
            self._plan.tune(table_path)

    def run(self, stream=None):
        """Enqueue one replay of the whole batch (device outputs are overwritten)."""
        self._plan.run(stream.handle if stream is not None else None)

    def results(self):
        """Per sub-search (indices int64, distances float32), ascending by (distance, index)."""
        hipdev.synchronize()
        return [(di.numpy().astype(np.int64), dd.numpy()) for di, dd in self._shapes]

    def info(self):
        return self._plan.info()
