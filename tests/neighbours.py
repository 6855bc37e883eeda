"""Checks shared by the approximate-search tests: equality with an exact top-k up to near ties.

An approximate search (IVF with bf16 / fp8 candidate distances and an fp32 rerank) is correct over
the rows it searched when its result is the exact top-k of those rows, except where rows sit so
close to the k-th distance that the candidate precision cannot order them. ``near_tie_check``
makes that precise: with D_k the exact (fp64) k-th smallest distance over the searched rows,

* every exact top-k row that is missing from the result has distance >= D_k * (1 - tol);
* every returned row outside the exact top-k has distance <= D_k * (1 + tol);
* every returned distance equals the row's fp64 distance within ``rtol_d``;
* the result has min(k, rows) entries, in ascending order of distance.

A search that drops a true neighbour well inside the k-th distance fails the first rule whatever
the recall over many queries.
"""

import numpy as np


def sq_l2_rows(X, q, rows):
    diff = X[rows].astype(np.float64) - q.astype(np.float64)
    return (diff * diff).sum(1)


def near_tie_check(got_idx, got_dist, rows, dd, k, tol, rtol_d=1e-5, atol_d=1e-6):
    """Asserts the rules above for one query; ``rows`` are the searched rows (sorted ids) and
    ``dd`` their fp64 squared distances to the query (``sq_l2_rows``). Returns the largest relative
    distance from D_k of a row swapped at the boundary (0 if none)."""
    rows = np.asarray(rows)
    got_idx = np.asarray(got_idx, np.int64)
    got_dist = np.asarray(got_dist, np.float64)
    kk = min(k, len(rows))
    assert len(got_idx) == kk, (len(got_idx), kk)
    assert np.all(np.diff(got_dist) >= 0), "distances not ascending"
    order = np.lexsort((rows, dd))
    if kk == 0:
        return 0.0
    dk = dd[order[kk - 1]]
    pos = np.searchsorted(rows, got_idx)
    assert np.all(pos < len(rows)) and np.all(rows[np.minimum(pos, len(rows) - 1)] == got_idx), \
        "a returned row is not among the searched rows"
    exact_got = dd[pos]
    np.testing.assert_allclose(got_dist, exact_got, rtol=rtol_d, atol=atol_d)
    top = set(rows[order[:kk]].tolist())
    got = set(got_idx.tolist())
    worst = 0.0
    scale = max(dk, 1e-30)
    for r in top - got:
        dev = (dk - dd[np.searchsorted(rows, r)]) / scale
        assert dev <= tol, f"true neighbour {r} missing at {dev:.3g} below the k-th distance (tol {tol})"
        worst = max(worst, dev)
    for r, e in zip(got_idx.tolist(), exact_got.tolist()):
        if r not in top:
            dev = (e - dk) / scale
            assert dev <= tol, f"row {r} returned at {dev:.3g} above the k-th distance (tol {tol})"
            worst = max(worst, dev)
    return worst


def nearest_lists(Q, cent, lists, nprobe, gap_rel=1e-5):
    """The nprobe nearest of the given centroid ids per query by fp64 distance, ties by id, and a
    mask of the queries whose nprobe-th and next centroid distances are apart by more than
    ``gap_rel`` (for the others an fp32 coarse step may pick either list)."""
    c = cent[lists].astype(np.float64)
    q = Q.astype(np.float64)
    d2 = (q * q).sum(1)[:, None] + (c * c).sum(1)[None] - 2.0 * (q @ c.T)
    order = np.lexsort((np.broadcast_to(lists, d2.shape), d2), axis=1)
    chosen = lists[order[:, :nprobe]]
    if nprobe >= len(lists):
        return chosen, np.ones(len(Q), bool)
    a = np.take_along_axis(d2, order[:, nprobe - 1:nprobe], 1)[:, 0]
    b = np.take_along_axis(d2, order[:, nprobe:nprobe + 1], 1)[:, 0]
    return chosen, (b - a) > gap_rel * np.maximum(b, 1e-30)


def probed_exact(X, labels, chosen, Qs):
    """For each query of ``Qs`` and its chosen lists (cluster ids, ``labels`` gives each row's
    cluster): the sorted rows of those lists and their fp64 squared distances, computed list by
    list as ||x||^2 + ||q||^2 - 2 x.q in fp64 (relative error ~1e-15 at these magnitudes)."""
    order = np.argsort(labels, kind="stable")
    bounds = np.searchsorted(labels[order], np.arange(labels.max() + 2))
    q64 = Qs.astype(np.float64)
    qn = (q64 * q64).sum(1)
    parts = [[] for _ in range(len(Qs))]
    for c in np.unique(chosen):
        rows = order[bounds[c]:bounds[c + 1]]
        if not len(rows):
            continue
        who = np.flatnonzero((chosen == c).any(1))
        x = X[rows].astype(np.float64)
        dd = (x * x).sum(1)[:, None] + qn[who][None] - 2.0 * (x @ q64[who].T)
        for j, qi in enumerate(who):
            parts[qi].append((rows, dd[:, j]))
    out = []
    for p in parts:
        rows = np.concatenate([r for r, _ in p]) if p else np.zeros(0, np.int64)
        dd = np.concatenate([d for _, d in p]) if p else np.zeros(0)
        o = np.argsort(rows, kind="stable")
        out.append((rows[o], dd[o]))
    return out
