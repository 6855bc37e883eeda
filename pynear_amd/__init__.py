"""pynear_amd: PyNear's nearest-neighbour search path on AMD Instinct MI355X (gfx950).

The public surface follows ``pynear`` (pynear/__init__.py:1-160): the exact ``VPTree*Index``
classes, ``VPTreeBinaryIndex`` with its width dispatch and input validation, ``IVFFlatBinaryIndex``,
``MIHBinaryIndex``, the ``BKTreeBinaryIndex*`` range-search classes, ``IVFFlatL2Index``, the
distance free functions, ``kmeans_l2`` and the scikit-learn adapters.
Every search runs in HIP kernels behind the C-ABI of ``include/pynear_amd.h``; there is no CPU
fallback, and a missing library raises ``NativeLibraryError`` at first use.
"""

from typing import List, Tuple

import numpy as np

from ._native import NativeLibraryError  # noqa: F401
from ._pynear import IVFFlatBinaryIndex, MIHBinaryIndex
from ._pynear import BKTreeBinaryIndex as BKTreeBinaryIndexN
from ._pynear import BKTreeBinaryIndex64, BKTreeBinaryIndex128, BKTreeBinaryIndex256, BKTreeBinaryIndex512
from ._pynear import VPTreeBinaryIndex as VPTreeBinaryIndexN
from ._pynear import VPTreeBinaryIndex64, VPTreeBinaryIndex128, VPTreeBinaryIndex256, VPTreeBinaryIndex512
from ._pynear import VPTreeChebyshevIndex, VPTreeL1Index, VPTreeL2Index
from ._pynear import dist_chebyshev, dist_hamming_64, dist_hamming_128, dist_hamming_256, dist_hamming_512
from ._pynear import dist_l1, dist_l2, kmeans_l2
from ._version import __version__  # noqa: F401
from .forest import IVFFlatL2Index

try:
    from .sklearn_adapter import PyNearKNeighborsClassifier, PyNearKNeighborsRegressor, PyNearNearestNeighbors
except ImportError:  # scikit-learn absent
    pass


def dist_hamming(a: List, b: List):
    """Hamming distance dispatched on byte width (pynear/__init__.py:40-53)."""
    if len(a) != len(b):
        raise ValueError("invalid data dimension: a and b dimensions must agree.")
    dim = len(a)
    fn = {64: dist_hamming_512, 32: dist_hamming_256, 16: dist_hamming_128, 8: dist_hamming_64}.get(dim)
    if fn is None:
        raise ValueError("invalid data dimension: hamming distance only supports 64, 32, 16 or 8 bytes of data")
    return fn(a, b)


class VPTreeBinaryIndex:
    """Exact Hamming index over uint8 rows with width dispatch (pynear/__init__.py:56-104)."""

    _BY_WIDTH = {64: VPTreeBinaryIndex512, 32: VPTreeBinaryIndex256, 16: VPTreeBinaryIndex128, 8: VPTreeBinaryIndex64}

    def __init__(self) -> None:
        self._index = None
        self._dimension = None

    def set(self, data: np.ndarray) -> None:
        self._validate(data)
        dim = data.shape[1]
        self._index = self._BY_WIDTH.get(dim, VPTreeBinaryIndexN)()
        self._dimension = dim
        self._index.set(data)

    def searchKNN(self, queries: np.ndarray, k: int) -> Tuple[list, list]:
        dim = queries.shape[1]
        if dim != self._dimension:
            raise ValueError(
                "invalid data dimension: index built data and query data dimensions must agree, "
                f"index built data dimension is {dim}")
        if self._index is None:
            return [], []
        self._validate(queries)
        return self._index.searchKNN(queries, k)

    def search1NN(self, queries: np.ndarray) -> Tuple[list, list]:
        if self._index is None:
            return [], []
        self._validate(queries)
        return self._index.search1NN(queries)

    def _validate(self, data: np.ndarray) -> None:
        if len(data.shape) != 2:
            raise ValueError("invalid data shape: binary indexes must be 2D")
        if data.dtype != "uint8":
            raise TypeError("invalid data type: binary indexes must be uint8")


__all__ = [
    "IVFFlatBinaryIndex", "MIHBinaryIndex", "BKTreeBinaryIndexN", "BKTreeBinaryIndex64", "BKTreeBinaryIndex128",
    "BKTreeBinaryIndex256", "BKTreeBinaryIndex512", "IVFFlatL2Index", "VPTreeL2Index", "VPTreeL1Index", "VPTreeChebyshevIndex", "VPTreeBinaryIndex", "VPTreeBinaryIndexN",
    "VPTreeBinaryIndex64", "VPTreeBinaryIndex128", "VPTreeBinaryIndex256", "VPTreeBinaryIndex512",
    "dist_l2", "dist_l1", "dist_chebyshev", "dist_hamming", "dist_hamming_64", "dist_hamming_128",
    "dist_hamming_256", "dist_hamming_512", "kmeans_l2", "NativeLibraryError",
]
