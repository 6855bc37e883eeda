"""Learned probe predictor for IVFFlatL2Index: an MLP that ranks a query's clusters, trained on
the device on exact neighbours, and an adaptive number of probes per query.

    probe = LearnedProbe(d, nlist)                  # d -> 256 (ReLU) -> nlist, bf16 on MFMA
    probe.fit(train_queries, neighbour_clusters)    # clusters of each query's exact top-L
    probes, counts = probe.select(q_dev, nq, mass=0.9, np_min=1, np_max=64)
    ivf.search_probes_device(q_dev.ptr, nq, probes.ptr, counts.ptr, 64, k, ...)

Training runs entirely in the library's kernels (forward, soft cross-entropy, backward, Adam);
the host only shuffles and slices batches.
"""

import numpy as np

from . import _native as N
from . import hipdev


class LearnedProbe:
    def __init__(self, d, nlist, hidden=256, seed=0):
        self.d, self.nlist, self.hidden = int(d), int(nlist), int(hidden)
        self.net = N.ProbeHandle(self.d, self.hidden, self.nlist, seed)
        self._sel = None

    def fit(self, queries, labels, epochs=3, batch=1024, lr=1e-3, seed=0, log=None):
        """Train on queries (n x d float32) and labels (n x L int32: clusters of each query's exact
        top-L neighbours). Returns the mean loss of every epoch."""
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        n, L = labels.shape
        batch = max(16, batch // 16 * 16)
        steps = n // batch
        if steps == 0:
            raise ValueError("fewer training queries than one batch")
        rng = np.random.default_rng(seed)
        loss_dev = hipdev.DeviceArray((steps,), np.float32)
        history = []
        for ep in range(epochs):
            perm = rng.permutation(n)[: steps * batch]
            xq = hipdev.DeviceArray.from_numpy(queries[perm])
            xl = hipdev.DeviceArray.from_numpy(labels[perm])
            for s in range(steps):
                self.net.train_step_device(xq.ptr + s * batch * self.d * 4, xl.ptr + s * batch * L * 4, L, batch, lr,
                                           loss_dev.ptr + 4 * s)
            losses = loss_dev.numpy()
            history.append(float(losses.mean()))
            if log:
                log(f"epoch {ep}: loss {history[-1]:.4f}")
            del xq, xl
        return history

    def logits(self, queries):
        q = hipdev.DeviceArray.from_numpy(np.ascontiguousarray(queries, dtype=np.float32))
        out = hipdev.DeviceArray((len(queries), self.nlist), np.float32)
        self.net.forward_device(q.ptr, len(queries), out.ptr)
        return out.numpy()

    def select(self, q_ptr, nq, mass, np_min, np_max, stream=None):
        """Device (probes int32[nq, np_max], counts int32[nq]) for device queries q_ptr."""
        if self._sel is None or self._sel[0] < nq or self._sel[1] != np_max:
            self._sel = (nq, np_max, hipdev.DeviceArray((nq, np_max), np.int32), hipdev.DeviceArray((nq,), np.int32))
        _, _, probes, counts = self._sel
        self.net.select_device(q_ptr, nq, mass, np_min, np_max, probes.ptr, counts.ptr,
                               stream.handle if stream is not None else None)
        return probes, counts
