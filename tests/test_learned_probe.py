"""Learned probe MLP: the oracle's gradients are pinned by finite differences (CPU); the device's
forward/backward kernels are checked against the oracle; training lowers the loss; the adaptive
selection honours its bounds; and IVF search over given probes equals the fixed-nprobe search
when the probes are the same clusters."""

import numpy as np
import pytest

from oracle import mlp

D, HID, NCLS, B, L = 32, 48, 64, 64, 5


def _problem(seed=0):
    rng = np.random.default_rng(seed)
    n1, n2 = HID * D, NCLS * HID
    params = np.concatenate([rng.standard_normal(n1) * np.sqrt(2 / D), rng.standard_normal(HID) * 0.1,
                             rng.standard_normal(n2) * np.sqrt(1 / HID), rng.standard_normal(NCLS) * 0.1])
    x = rng.standard_normal((B, D)).astype(np.float32)
    labels = rng.integers(0, NCLS, (B, L)).astype(np.int32)
    labels[0, 2] = labels[0, 1]  # a repeated cluster carries 2/L
    return params.astype(np.float32), x, labels


# ------------------------------------------------------------------ CPU: pin the oracle
def test_oracle_gradient_matches_finite_differences():
    params, x, labels = _problem()
    g = mlp.grads_exact(params, x, labels, D, HID, NCLS)
    rng = np.random.default_rng(1)
    p64 = params.astype(np.float64)
    for i in rng.choice(len(params), 60, replace=False):
        h = 1e-5
        e = np.zeros_like(p64)
        e[i] = h
        fd = (mlp.loss_exact(p64 + e, x, labels, D, HID, NCLS) - mlp.loss_exact(p64 - e, x, labels, D, HID, NCLS)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_oracle_bf16_path_close_to_exact():
    params, x, labels = _problem(2)
    loss, g = mlp.loss_and_grads(params, x, labels, D, HID, NCLS)
    ge = mlp.grads_exact(params, x, labels, D, HID, NCLS)
    assert abs(loss - mlp.loss_exact(params, x, labels, D, HID, NCLS)) < 2e-2
    assert np.linalg.norm(g - ge) / np.linalg.norm(ge) < 3e-2


def test_planted_mistake_in_oracle_is_caught():
    """A dropped ReLU mask in the backward pass must fail the finite-difference check."""
    params, x, labels = _problem(3)
    W1, b1, W2, b2 = mlp.unpack(params.astype(np.float64), D, HID, NCLS)
    X = x.astype(np.float64)
    H = np.maximum(X @ W1.T + b1, 0)
    logits = H @ W2.T + b2
    P = np.exp(logits - logits.max(1, keepdims=True))
    P /= P.sum(1, keepdims=True)
    dL = (P - mlp.targets(labels, NCLS)) / B
    wrong = (dL @ W2).T @ X  # no mask
    right = mlp.grads_exact(params, x, labels, D, HID, NCLS)[:HID * D].reshape(HID, D)
    assert np.abs(wrong - right).max() > 1e-3


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_device_gradients_match_oracle():
    from pynear_amd import _native as N, hipdev
    params, x, labels = _problem(4)
    net = N.ProbeHandle(D, HID, NCLS, seed=0)
    net.set_params(params)
    assert np.array_equal(net.get_params(), params)
    dx = hipdev.DeviceArray.from_numpy(x)
    dl = hipdev.DeviceArray.from_numpy(labels)
    loss = hipdev.DeviceArray((1,), np.float32)
    grads = hipdev.DeviceArray((net.n_params,), np.float32)
    net.grads_device(dx.ptr, dl.ptr, L, B, loss.ptr, grads.ptr)
    g = grads.numpy().astype(np.float64)
    ref_loss, ref_g = mlp.loss_and_grads(params, x, labels, D, HID, NCLS)
    assert abs(float(loss.numpy()[0]) - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
    n1, n2 = HID * D, NCLS * HID
    for name, sl in (("W1", slice(0, n1)), ("b1", slice(n1, n1 + HID)), ("W2", slice(n1 + HID, n1 + HID + n2)),
                     ("b2", slice(n1 + HID + n2, None))):
        err = np.linalg.norm(g[sl] - ref_g[sl]) / max(np.linalg.norm(ref_g[sl]), 1e-12)
        assert err < 1e-2, (name, err)
    # and the device logits equal the oracle forward (bf16 inputs, fp32 sums)
    lg = hipdev.DeviceArray((B, NCLS), np.float32)
    net.forward_device(dx.ptr, B, lg.ptr)
    _, _, ref_logits = mlp.forward(params, x, D, HID, NCLS)
    np.testing.assert_allclose(lg.numpy(), ref_logits, rtol=1e-3, atol=1e-3)


@pytest.mark.gpu
def test_training_lowers_loss_and_selection_bounds():
    from pynear_amd.learned import LearnedProbe
    from pynear_amd import hipdev
    rng = np.random.default_rng(5)
    centres = rng.standard_normal((NCLS, D)).astype(np.float32) * 3
    lab = rng.integers(0, NCLS, 4096)
    q = centres[lab] + rng.standard_normal((4096, D)).astype(np.float32)
    labels = np.repeat(lab[:, None], 4, axis=1).astype(np.int32)
    probe = LearnedProbe(D, NCLS, hidden=HID, seed=1)
    hist = probe.fit(q, labels, epochs=4, batch=256, lr=3e-3)
    assert hist[-1] < 0.5 * hist[0], hist
    acc = float((probe.logits(q[:512]).argmax(1) == lab[:512]).mean())
    assert acc > 0.9
    dq = hipdev.DeviceArray.from_numpy(q[:300])
    for mass, lo, hi in ((0.5, 1, 8), (0.99, 2, 6), (0.999, 1, 64)):
        probes, counts = probe.select(dq.ptr, 300, mass, lo, hi)
        c = counts.numpy()
        p = probes.numpy()
        assert np.all((c >= lo) & (c <= hi))
        logits = probe.logits(q[:300])
        P = np.exp(logits - logits.max(1, keepdims=True))
        P /= P.sum(1, keepdims=True)
        for i in range(300):
            sel = p[i, :c[i]]
            assert len(set(sel.tolist())) == c[i] and np.all((sel >= 0) & (sel < NCLS))
            if c[i] < hi:  # stopped by the mass, not by the cap: the chosen mass reaches the target
                assert P[i, sel].sum() >= mass - 1e-3
            rest = np.setdiff1d(np.arange(NCLS), sel)
            assert P[i, sel].min() >= P[i, rest].max() - 1e-6  # the most probable clusters


@pytest.mark.gpu
@pytest.mark.parametrize("ncls", [1500, 3000, 5000])
def test_selection_at_many_clusters(ncls):
    """The selection for 1,500 and 3,000 clusters (one wave per query, 32 or 64 probabilities per
    lane) and 5,000 (a workgroup per query) on an untrained net: the chosen set is the most
    probable clusters, within [np_min, np_max], in ascending cluster order up to 4,096 clusters,
    reaches the mass unless capped, and is the fewest that do (dropping its least probable
    cluster falls short, ties aside)."""
    from pynear_amd.learned import LearnedProbe
    from pynear_amd import hipdev
    rng = np.random.default_rng(11)
    q = rng.standard_normal((200, D)).astype(np.float32) * 4
    probe = LearnedProbe(D, ncls, hidden=HID, seed=3)
    dq = hipdev.DeviceArray.from_numpy(q)
    logits = probe.logits(q).astype(np.float64)
    P = np.exp(logits - logits.max(1, keepdims=True))
    P /= P.sum(1, keepdims=True)
    for mass, lo, hi in ((0.05, 1, 300), (0.2, 3, 40)):
        probes, counts = probe.select(dq.ptr, len(q), mass, lo, hi)
        c, p = counts.numpy(), probes.numpy()
        assert np.all((c >= lo) & (c <= hi))
        for i in range(len(q)):
            sel = p[i, :c[i]]
            assert len(set(sel.tolist())) == c[i] and np.all((sel >= 0) & (sel < ncls))
            if ncls <= 4096:
                assert np.all(np.diff(sel) > 0)
            rest = np.setdiff1d(np.arange(ncls), sel)
            assert P[i, sel].min() >= P[i, rest].max() - 1e-6
            if c[i] < hi:
                assert P[i, sel].sum() >= mass - 1e-4
                if c[i] > lo:
                    smallest = np.sort(P[i, sel])
                    if smallest[1] - smallest[0] > 1e-6:  # a strict least member (no tie at the cut)
                        assert P[i, sel].sum() - smallest[0] < mass + 1e-4


@pytest.mark.gpu
def test_ivf_search_with_given_probes_equals_fixed_nprobe():
    from pynear_amd import _native as N, hipdev
    rng = np.random.default_rng(6)
    X = rng.standard_normal((20000, 64)).astype(np.float32)
    Q = rng.standard_normal((300, 64)).astype(np.float32)
    h = N.IvfL2Handle()
    labels, cent = h.build(X, 64, 10, 42)
    np_ = 6
    ref_i, ref_d, ref_c = h.search(Q, np_, 10)
    # the np_ nearest centroids in the same order as the coarse stage ((distance, id) ascending)
    keep = np.flatnonzero(np.bincount(labels, minlength=64))
    d2 = ((Q[:, None, :].astype(np.float64) - cent[keep][None].astype(np.float64)) ** 2).sum(-1)
    pr = keep[np.argsort(d2, axis=1, kind="stable")[:, :np_]].astype(np.int32)
    pad = np.full((len(Q), 9), -1, np.int32)  # np_max 9 with unused slots
    pad[:, :np_] = pr
    cnt = np.full(len(Q), np_, np.int32)
    dq, dp, dc = hipdev.DeviceArray.from_numpy(Q), hipdev.DeviceArray.from_numpy(pad), hipdev.DeviceArray.from_numpy(cnt)
    oi = hipdev.DeviceArray((len(Q), 10), np.int32)
    od = hipdev.DeviceArray((len(Q), 10), np.float32)
    oc = hipdev.DeviceArray((len(Q),), np.int32)
    h.search_probes_device(dq.ptr, len(Q), dp.ptr, dc.ptr, 9, 10, oi.ptr, od.ptr, oc.ptr)
    assert np.array_equal(oc.numpy(), ref_c)
    assert np.array_equal(oi.numpy().astype(np.int64), ref_i)
    assert np.array_equal(od.numpy(), ref_d)
