"""ctypes binding of ``libpynear_amd.so`` (the C-ABI declared in ``include/pynear_amd.h``).

This module is the Python side of the drop-in boundary: it plays the role of the reference's
pybind11 module ``_pynear`` (pynear/src/PythonBindings.cpp:331-481). Status codes are mapped to
the exception types pybind11 raises for the reference's C++ throws (runtime_error ->
RuntimeError, invalid_argument -> ValueError).

There is no fallback: if the library is missing or no HIP device is usable, calls raise.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PYNEAR_AMD_LIB") or os.path.join(HERE, "libpynear_amd.so")

PN_OK, PN_EINVAL, PN_ERUNTIME, PN_ETYPE, PN_EDEVICE, PN_ENOMEM = range(6)
METRIC_L2, METRIC_L1, METRIC_CHEBYSHEV, METRIC_HAMMING = range(4)

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32

# name -> (restype, argtypes); every symbol declared in include/pynear_amd.h
SIGNATURES = {
    "pn_last_error": (ctypes.c_char_p, []),
    "pn_version": (ctypes.c_char_p, []),
    "pn_device_count": (_i32, [ctypes.POINTER(_i32)]),
    "pn_stage_timing": (_i32, [_i32]),
    "pn_stage_times": (_i32, [_vp, _vp]),
    "pn_tune_set": (_i32, [ctypes.c_char_p, _i64]),
    "pn_tune_get": (_i32, [ctypes.c_char_p, ctypes.POINTER(_i64)]),
    "pn_pairwise_distance": (_i32, [_i32, _vp, _vp, _i64, _i64, _vp]),
    "pn_exact_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "pn_exact_destroy": (None, [_vp]),
    "pn_exact_set": (_i32, [_vp, _vp, _i64, _i64]),
    "pn_exact_info": (_i32, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "pn_exact_search": (_i32, [_vp, _vp, _i64, _i64, _i32, _vp, _vp]),
    "pn_exact_search_device": (_i32, [_vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "pn_exact_search_device_noalloc": (_i32, [_vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "pn_exact_range": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _i64, ctypes.POINTER(_i64)]),
    "pn_mih_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "pn_mih_destroy": (None, [_vp]),
    "pn_mih_set": (_i32, [_vp, _vp, _i64, _i64]),
    "pn_mih_info": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "pn_mih_search": (_i32, [_vp, _vp, _i64, _i64, _i32, _i32, _vp, _vp, _vp]),
    "pn_kmeans_l2": (_i32, [_vp, _i64, _i64, _i64, _i64, _u32, _vp, _vp]),
    "pn_ivfl2_create": (_i32, [ctypes.POINTER(_vp)]),
    "pn_ivfl2_destroy": (None, [_vp]),
    "pn_ivfl2_build": (_i32, [_vp, _vp, _i64, _i64, _i64, _i64, _u32, _vp, _vp]),
    "pn_ivfl2_set": (_i32, [_vp, _vp, _i64, _i64, _vp, _i64, _vp]),
    "pn_ivfl2_info": (_i32, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i32)]),
    "pn_ivfl2_scan_work": (_i32, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "pn_ivfl2_search": (_i32, [_vp, _vp, _i64, _i64, _i32, _i32, _i32, _vp, _vp, _vp]),
    "pn_ivfl2_search_device": (_i32, [_vp, _vp, _i64, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "pn_ivfbin_create": (_i32, [_i32, _i32, _i32, _u32, ctypes.POINTER(_vp)]),
    "pn_ivfbin_destroy": (None, [_vp]),
    "pn_ivfbin_set": (_i32, [_vp, _vp, _i64, _i64]),
    "pn_ivfbin_info": (_i32, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i32)]),
    "pn_ivfbin_get_params": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "pn_ivfbin_set_nprobe": (_i32, [_vp, _i32]),
    "pn_ivfbin_export": (_i32, [_vp, _vp, _vp]),
    "pn_ivfbin_search": (_i32, [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp]),
    "pn_ivfbin_search_device": (_i32, [_vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp]),
    "pn_ivfl2_search_probes_device": (_i32, [_vp, _vp, _i64, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "pn_probe_create": (_i32, [_i32, _i32, _i32, _u32, ctypes.POINTER(_vp)]),
    "pn_probe_destroy": (None, [_vp]),
    "pn_probe_get_params": (_i32, [_vp, _vp]),
    "pn_probe_set_params": (_i32, [_vp, _vp]),
    "pn_probe_grads_device": (_i32, [_vp, _vp, _vp, _i32, _i64, _vp, _vp, _vp]),
    "pn_probe_train_step_device": (_i32, [_vp, _vp, _vp, _i32, _i64, ctypes.c_float, _vp, _vp]),
    "pn_probe_forward_device": (_i32, [_vp, _vp, _i64, _vp, _vp]),
    "pn_probe_select_device": (_i32, [_vp, _vp, _i64, ctypes.c_float, _i32, _i32, _vp, _vp, _vp]),
    "pn_batch_create": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "pn_batch_destroy": (None, [_vp]),
    "pn_batch_run": (_i32, [_vp, _vp]),
    "pn_batch_tune": (_i32, [_vp, ctypes.c_char_p, _i32]),
    "pn_batch_info": (_i32, [_vp, ctypes.POINTER(_i64), _vp]),
}

_LIB = None


class NativeLibraryError(RuntimeError):
    """The HIP library could not be loaded; the GPU path has no CPU fallback."""


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} not found: build it with `python -m pynear_amd.build` (hipcc, gfx950)")
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(status):
    if status == PN_OK:
        return
    msg = lib().pn_last_error().decode(errors="replace")
    if status == PN_EINVAL:
        raise ValueError(msg)
    if status == PN_ETYPE:
        raise TypeError(msg)
    raise RuntimeError(msg)


def ptr(a):
    """Address of a numpy array's data (None for an empty array)."""
    return ctypes.c_void_p(a.ctypes.data) if a.size else None


def device_count():
    c = _i32(0)
    check(lib().pn_device_count(ctypes.byref(c)))
    return c.value


class ExactHandle:
    """Owner of a ``pn_exact`` handle (brute-force exact index on the current device)."""

    def __init__(self, metric):
        h = _vp()
        check(lib().pn_exact_create(metric, ctypes.byref(h)))
        self._h = h
        self.metric = metric

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.pn_exact_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set(self, rows):
        n, d = rows.shape
        check(lib().pn_exact_set(self._h, ptr(rows), n, d))

    def info(self):
        n, d = _i64(0), _i64(0)
        check(lib().pn_exact_info(self._h, ctypes.byref(n), ctypes.byref(d)))
        return n.value, d.value

    def search(self, queries, k):
        """Ascending (dist, idx) arrays of shape (nq, min(k, n))."""
        nq, d = queries.shape
        n, _ = self.info()
        kk = max(0, min(int(k), n))
        idx = np.empty((nq, kk), np.int64)
        dist = np.empty((nq, kk), np.int64 if self.metric == METRIC_HAMMING else np.float32)
        check(lib().pn_exact_search(self._h, ptr(queries), nq, d, int(k), ptr(idx), ptr(dist)))
        return idx, dist

    def search_device(self, q_ptr, nq, k, out_idx_ptr, out_dist_ptr, stream_ptr=None):
        check(lib().pn_exact_search_device(self._h, _vp(q_ptr), nq, int(k), _vp(out_idx_ptr),
                                           _vp(out_dist_ptr), _vp(stream_ptr) if stream_ptr else None))

    def search_device_noalloc(self, q_ptr, nq, k, out_idx_ptr, out_dist_ptr, stream_ptr=None):
        """search_device through the allocation-free entry point (fixed per-GPU scratch)."""
        check(lib().pn_exact_search_device_noalloc(self._h, _vp(q_ptr), nq, int(k), _vp(out_idx_ptr),
                                                   _vp(out_dist_ptr), _vp(stream_ptr) if stream_ptr else None))

    def range(self, queries, threshold, capacity=None):
        """Hamming range search: (offsets int64[nq + 1], rows int32[total], distances int32[total]),
        the matches of query q at [offsets[q], offsets[q + 1]) in ascending row order."""
        nq, d = queries.shape
        offsets = np.zeros(nq + 1, np.int64)
        total = _i64(0)
        cap = int(capacity) if capacity is not None else min(64 * nq, 1 << 22)
        while True:
            idx = np.empty(cap, np.int32)
            dist = np.empty(cap, np.int32)
            check(lib().pn_exact_range(self._h, ptr(queries), nq, d, int(threshold), ptr(offsets), ptr(idx), ptr(dist),
                                       cap, ctypes.byref(total)))
            if total.value <= cap:
                return offsets, idx[:total.value], dist[:total.value]
            cap = total.value  # too many matches for the first guess: once more with room for all


class MihHandle:
    """Owner of a ``pn_mih`` handle (MIHBinaryIndex on the current device)."""

    def __init__(self, m):
        h = _vp()
        check(lib().pn_mih_create(int(m), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.pn_mih_destroy(h)
            self._h = None

    def set(self, rows):
        n, nb = rows.shape
        check(lib().pn_mih_set(self._h, ptr(rows), n, nb))

    def info(self):
        m, n, nb = _i32(0), _i64(0), _i64(0)
        check(lib().pn_mih_info(self._h, ctypes.byref(m), ctypes.byref(n), ctypes.byref(nb)))
        return m.value, n.value, nb.value

    def search(self, queries, k, radius):
        """(idx int64[nq, kk], dist int64[nq, kk], counts int64[nq]); kk = min(k, n), slots past
        the count are -1."""
        nq, nb = queries.shape
        _, n, _ = self.info()
        kk = max(0, min(int(k), n))
        idx = np.full((nq, kk), -1, np.int64)
        dist = np.full((nq, kk), -1, np.int64)
        cnt = np.zeros(nq, np.int64)
        check(lib().pn_mih_search(self._h, ptr(queries), nq, nb, int(k), int(radius), ptr(idx), ptr(dist), ptr(cnt)))
        return idx, dist, cnt


def tune_set(name, value):
    check(lib().pn_tune_set(name.encode(), int(value)))


def tune_get(name):
    v = _i64(0)
    check(lib().pn_tune_get(name.encode(), ctypes.byref(v)))
    return v.value


class IvfBinHandle:
    """Owner of a ``pn_ivfbin`` handle (IVFFlatBinaryIndex on the current device)."""

    def __init__(self, nlist, nprobe, max_iter, seed):
        h = _vp()
        check(lib().pn_ivfbin_create(int(nlist), int(nprobe), int(max_iter), int(seed), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.pn_ivfbin_destroy(h)
            self._h = None

    def set(self, rows):
        n, nb = rows.shape
        check(lib().pn_ivfbin_set(self._h, ptr(rows), n, nb))

    def info(self):
        n, nb, nc = _i64(0), _i64(0), _i32(0)
        check(lib().pn_ivfbin_info(self._h, ctypes.byref(n), ctypes.byref(nb), ctypes.byref(nc)))
        return n.value, nb.value, nc.value

    def params(self):
        a, b = _i32(0), _i32(0)
        check(lib().pn_ivfbin_get_params(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_nprobe(self, nprobe):
        check(lib().pn_ivfbin_set_nprobe(self._h, int(nprobe)))

    def export(self):
        n, nb, nc = self.info()
        cent = np.zeros((nc, nb), np.uint8)
        labels = np.zeros(n, np.int32)
        check(lib().pn_ivfbin_export(self._h, ptr(cent), ptr(labels)))
        return cent, labels

    def search(self, queries, k):
        nq, nb = queries.shape
        k = int(k)
        idx = np.full((nq, max(k, 0)), -1, np.int64)
        dist = np.full((nq, max(k, 0)), -1, np.int64)
        cnt = np.zeros(nq, np.int32)
        check(lib().pn_ivfbin_search(self._h, ptr(queries), nq, nb, k, ptr(idx), ptr(dist), ptr(cnt)))
        return idx, dist, cnt

    def search_device(self, q_ptr, nq, k, out_idx_ptr, out_dist_ptr, out_cnt_ptr, stream_ptr=None):
        check(lib().pn_ivfbin_search_device(self._h, _vp(q_ptr), nq, int(k), _vp(out_idx_ptr), _vp(out_dist_ptr),
                                            _vp(out_cnt_ptr), _vp(stream_ptr) if stream_ptr else None))


class IvfL2Handle:
    """Owner of a ``pn_ivfl2`` handle (the IVFFlatL2Index search on the current device)."""

    PRECISION = {"bf16": 0, "bf16x3": 1, "fp8": 2}

    def __init__(self):
        h = _vp()
        check(lib().pn_ivfl2_create(ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.pn_ivfl2_destroy(h)
            self._h = None

    def build(self, rows, n_clusters, max_iter, seed):
        n, d = rows.shape
        labels = np.empty(n, np.int32)
        cent = np.empty((int(n_clusters), d), np.float32)
        check(lib().pn_ivfl2_build(self._h, ptr(rows), n, d, int(n_clusters), int(max_iter), int(seed),
                                   ptr(labels), ptr(cent)))
        return labels, cent

    def set(self, rows, centroids, labels):
        n, d = rows.shape
        check(lib().pn_ivfl2_set(self._h, ptr(rows), n, d, ptr(centroids), len(centroids), ptr(labels)))

    def info(self):
        n, d, nl = _i64(0), _i64(0), _i32(0)
        check(lib().pn_ivfl2_info(self._h, ctypes.byref(n), ctypes.byref(d), ctypes.byref(nl)))
        return n.value, d.value, nl.value

    def scan_work(self):
        """(MFMA instructions, fetched fp8 row bytes) of the last fp8 list scan."""
        m, b = _i64(0), _i64(0)
        check(lib().pn_ivfl2_scan_work(self._h, ctypes.byref(m), ctypes.byref(b)))
        return m.value, b.value

    def search(self, queries, nprobe, k, precision="bf16"):
        nq, d = queries.shape
        k = int(k)
        idx = np.full((nq, k), -1, np.int64)
        dist = np.full((nq, k), -1, np.float32)
        cnt = np.zeros(nq, np.int32)
        check(lib().pn_ivfl2_search(self._h, ptr(queries), nq, d, int(nprobe), k, self.PRECISION[precision],
                                    ptr(idx), ptr(dist), ptr(cnt)))
        return idx, dist, cnt

    def search_probes_device(self, q_ptr, nq, probes_ptr, counts_ptr, np_max, k, out_idx_ptr, out_dist_ptr,
                             out_cnt_ptr, precision="bf16", stream_ptr=None):
        check(lib().pn_ivfl2_search_probes_device(self._h, _vp(q_ptr), nq, _vp(probes_ptr), _vp(counts_ptr),
                                                  int(np_max), int(k), self.PRECISION[precision], _vp(out_idx_ptr),
                                                  _vp(out_dist_ptr), _vp(out_cnt_ptr),
                                                  _vp(stream_ptr) if stream_ptr else None))

    def search_device(self, q_ptr, nq, nprobe, k, out_idx_ptr, out_dist_ptr, out_cnt_ptr, precision="bf16",
                      stream_ptr=None):
        check(lib().pn_ivfl2_search_device(self._h, _vp(q_ptr), nq, int(nprobe), int(k), self.PRECISION[precision],
                                           _vp(out_idx_ptr), _vp(out_dist_ptr), _vp(out_cnt_ptr),
                                           _vp(stream_ptr) if stream_ptr else None))


def kmeans_l2(data, k, max_iter, seed):
    """(labels int32[n], centroids float32[k, d]) from the GPU k-means."""
    n, d = data.shape
    labels = np.empty(n, np.int32)
    cent = np.empty((max(int(k), 0), d), np.float32)
    check(lib().pn_kmeans_l2(ptr(data), n, d, int(k), int(max_iter), int(seed), ptr(labels), ptr(cent)))
    return labels, cent


STAGES = ("scan", "select", "coarse", "rerank", "other")


def stage_timing(enable):
    check(lib().pn_stage_timing(1 if enable else 0))


def stage_times():
    """{stage: (total_ms, launches)} since the previous call (events resolved here)."""
    ms = np.zeros(5, np.float64)
    n = np.zeros(5, np.int64)
    check(lib().pn_stage_times(ptr(ms), ptr(n)))
    return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(STAGES)}


def pairwise_distance(metric, a, b):
    """Row-wise distances between equal-shape 2-D arrays a and b."""
    n, d = a.shape
    out = np.empty(n, np.int64 if metric == METRIC_HAMMING else np.float32)
    check(lib().pn_pairwise_distance(metric, ptr(a), ptr(b), n, d, ptr(out)))
    return out


class BatchHandle:
    """Owner of a ``pn_batch`` plan: exact L2 searches over many indices, replayed as one graph.

    ``indices`` are ExactHandle objects (L2); ``queries``/``out_idx``/``out_dist`` are device
    pointers (ints), one per index; the plan binds them."""

    def __init__(self, indices, queries, nq, k, out_idx, out_dist):
        n = len(indices)
        P = _vp * n
        self._keep = indices
        hs = P(*[h.handle for h in indices])
        qs = P(*[_vp(int(p)) for p in queries])
        oi = P(*[_vp(int(p)) for p in out_idx])
        od = P(*[_vp(int(p)) for p in out_dist])
        nqa = (_i64 * n)(*[int(v) for v in nq])
        ka = (_i32 * n)(*[int(v) for v in k])
        h = _vp()
        check(lib().pn_batch_create(n, hs, qs, nqa, ka, oi, od, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.pn_batch_destroy(h)
            self._h = None

    def run(self, stream_ptr=None):
        check(lib().pn_batch_run(self._h, _vp(stream_ptr) if stream_ptr else None))

    def tune(self, table_path=None, reps=3):
        check(lib().pn_batch_tune(self._h, table_path.encode() if table_path else None, int(reps)))

    def info(self):
        t = _i64(0)
        cfg = (_i32 * 3)()
        check(lib().pn_batch_info(self._h, ctypes.byref(t), cfg))
        return t.value, {"slice_rows": cfg[0], "wg_per_cu": cfg[1], "order": cfg[2]}


class ProbeHandle:
    """Owner of a ``pn_probe`` (learned probe MLP d -> hidden -> nlist on the current device)."""

    def __init__(self, d, hidden, nlist, seed=0):
        h = _vp()
        check(lib().pn_probe_create(int(d), int(hidden), int(nlist), int(seed), ctypes.byref(h)))
        self._h = h
        self.d, self.hidden, self.nlist = int(d), int(hidden), int(nlist)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _LIB is not None:
            _LIB.pn_probe_destroy(h)
            self._h = None

    @property
    def n_params(self):
        return self.hidden * self.d + self.hidden + self.nlist * self.hidden + self.nlist

    def get_params(self):
        out = np.empty(self.n_params, np.float32)
        check(lib().pn_probe_get_params(self._h, ptr(out)))
        return out

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == self.n_params
        check(lib().pn_probe_set_params(self._h, ptr(params)))

    def grads_device(self, x_ptr, labels_ptr, L, B, loss_ptr, grads_ptr, stream_ptr=None):
        check(lib().pn_probe_grads_device(self._h, _vp(x_ptr), _vp(labels_ptr), int(L), int(B), _vp(loss_ptr),
                                          _vp(grads_ptr), _vp(stream_ptr) if stream_ptr else None))

    def train_step_device(self, x_ptr, labels_ptr, L, B, lr, loss_ptr=None, stream_ptr=None):
        check(lib().pn_probe_train_step_device(self._h, _vp(x_ptr), _vp(labels_ptr), int(L), int(B), float(lr),
                                               _vp(loss_ptr) if loss_ptr else None,
                                               _vp(stream_ptr) if stream_ptr else None))

    def forward_device(self, x_ptr, B, logits_ptr, stream_ptr=None):
        check(lib().pn_probe_forward_device(self._h, _vp(x_ptr), int(B), _vp(logits_ptr),
                                            _vp(stream_ptr) if stream_ptr else None))

    def select_device(self, x_ptr, B, mass, np_min, np_max, probes_ptr, counts_ptr, stream_ptr=None):
        check(lib().pn_probe_select_device(self._h, _vp(x_ptr), int(B), float(mass), int(np_min), int(np_max),
                                           _vp(probes_ptr), _vp(counts_ptr), _vp(stream_ptr) if stream_ptr else None))
