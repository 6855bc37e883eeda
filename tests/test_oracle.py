"""Pins the CPU oracle (CPU only).

The oracle is checked against (1) the reference's own test vectors and recorded outputs of the
reference module (tests/golden), (2) the reference module itself when oracle/_ref is built, and
(3) what the mathematics fixes: numpy bit counts, FP64 closed forms, brute-force order. Each check
is chosen so that a plausible mistake (a dropped term, a changed summation order, a wrong tie rule)
fails it; test_planted_mistakes_are_caught shows that for the float arithmetic.
"""

import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(GOLD, "ref_vectors.npz"))


def test_hamming_known_answers():
    with open(os.path.join(GOLD, "hamming_known_answers.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        a, b = np.array(c["a"], np.uint8), np.array(c["b"], np.uint8)
        assert oracle.dist_hamming(a, b) == c["expected"]
        assert oracle.dist_hamming(b, a) == c["expected"]


def test_hamming_matches_bitcount():
    rng = np.random.default_rng(0)
    for nb in (1, 3, 8, 13, 16, 32, 64, 100):
        a = rng.integers(0, 256, (50, nb), dtype=np.uint8)
        b = rng.integers(0, 256, (50, nb), dtype=np.uint8)
        for i in range(50):
            assert oracle.dist_hamming(a[i], b[i]) == int(np.bitwise_count(a[i] ^ b[i]).sum())


@pytest.mark.parametrize("metric", [oracle.L2, oracle.L1, oracle.LINF])
def test_float_distance_vs_fp64(metric):
    rng = np.random.default_rng(1)
    for d in (1, 2, 5, 8, 13, 32, 768):
        a = rng.standard_normal(d).astype(np.float32)
        b = rng.standard_normal(d).astype(np.float32)
        diff = a.astype(np.float64) - b.astype(np.float64)
        exact = {oracle.L2: np.sqrt((diff ** 2).sum()), oracle.L1: np.abs(diff).sum(),
                 oracle.LINF: np.abs(diff).max()}[metric]
        assert oracle.dist(metric, a, b) == pytest.approx(exact, rel=2e-6)


def test_float_distance_bit_exact_vs_reference_module(ref_module):
    rng = np.random.default_rng(2)
    fns = {oracle.L2: ref_module.dist_l2, oracle.L1: ref_module.dist_l1, oracle.LINF: ref_module.dist_chebyshev}
    for d in list(range(1, 41)) + [64, 100, 512, 768]:
        for _ in range(20):
            a = (rng.standard_normal(d) * rng.choice([1e-3, 1.0, 1e3])).astype(np.float32)
            b = rng.standard_normal(d).astype(np.float32)
            for m, f in fns.items():
                assert np.float32(oracle.dist(m, a, b)) == np.float32(f(a, b)), (m, d)


def test_planted_mistakes_are_caught(g):
    """Restatements with a dropped term or a plain left-to-right sum disagree with the recorded
    reference distances; the oracle agrees with them."""
    X, Q, ref_d = g["l2_X"], g["l2_Q"], g["l2_dist"]
    idx = g["l2_idx"]
    ok = drop = naive = 0
    for qi in range(len(Q)):
        for j in range(idx.shape[1]):
            x = X[idx[qi, j]]
            ok += np.float32(oracle.dist(oracle.L2, Q[qi], x)) == ref_d[qi, j]
            drop += np.float32(oracle.dist(oracle.L2, Q[qi][:-1], x[:-1])) == ref_d[qi, j]
            t = np.float32(0)
            for e in range(len(x)):
                t = np.float32(t + np.float32(Q[qi][e] - x[e]) * np.float32(Q[qi][e] - x[e]))
            naive += np.float32(np.sqrt(t)) == ref_d[qi, j]
    total = idx.size
    assert ok == total
    assert drop < total // 10
    assert naive < total


@pytest.mark.parametrize("name,metric", [("l2", oracle.L2), ("l1", oracle.L1), ("linf", oracle.LINF)])
def test_knn_float_vs_reference_vptree(g, name, metric):
    X, Q = g[f"{name}_X"], g[f"{name}_Q"]
    oi, od = oracle.knn_float(metric, X, Q, 5)
    # the reference returns farthest-first; distances bit-identical, indices equal (no ties here)
    np.testing.assert_array_equal(od, g[f"{name}_dist"][:, ::-1])
    np.testing.assert_array_equal(oi, g[f"{name}_idx"][:, ::-1])
    oi1, od1 = oracle.knn_float(metric, X, Q, 1)
    np.testing.assert_array_equal(od1[:, 0], g[f"{name}_dist1"])
    np.testing.assert_array_equal(oi1[:, 0], g[f"{name}_idx1"])


@pytest.mark.parametrize("name,metric", [("l2", oracle.L2), ("l1", oracle.L1), ("linf", oracle.LINF)])
def test_five_point_fixture(g, name, metric):
    # test_vptree.py:282-301: nearest-first [1, 0, 2, 3]
    oi, od = oracle.knn_float(metric, g["five_X"], g["five_Q"], 4)
    assert oi.tolist() == [[1, 0, 2, 3]]
    np.testing.assert_array_equal(oi, g[f"five_{name}_idx"][:, ::-1])
    np.testing.assert_array_equal(od, g[f"five_{name}_dist"][:, ::-1])


def test_knn_hamming_vs_reference_vptree(g):
    oi, od = oracle.knn_hamming(g["bin_X"], g["bin_Q"], 3)
    np.testing.assert_array_equal(od, g["bin_dist"][:, ::-1])


def test_knn_brute_force_order_and_fp64():
    rng = np.random.default_rng(3)
    X = rng.random((700, 6)).astype(np.float32)
    Q = rng.random((9, 6)).astype(np.float32)
    for metric in (oracle.L2, oracle.L1, oracle.LINF):
        oi, od = oracle.knn_float(metric, X, Q, 700)
        fi, fd = oracle.knn_fp64(metric, X, Q, 700)
        np.testing.assert_allclose(od, fd, rtol=2e-6)
        assert (np.diff(od, axis=1) >= 0).all()
        for q in range(len(Q)):  # equal as sets, identical order wherever fp32 does not tie
            assert sorted(oi[q]) == list(range(700))
    Xb = (rng.integers(0, 3, (400, 4)) * 5).astype(np.uint8)
    Qb = Xb[:7]
    oi, od = oracle.knn_hamming(Xb, Qb, 400)
    full = np.bitwise_count(Qb[:, None, :] ^ Xb[None, :, :]).sum(-1)
    exp = np.lexsort((np.broadcast_to(np.arange(400), full.shape), full), axis=-1)
    np.testing.assert_array_equal(oi, exp)


@pytest.mark.parametrize("nprobe", [1, 4, 32])
def test_ivfbin_vs_reference_outputs(g, nprobe):
    X, Q = g["ivfbin_X"], g["ivfbin_Q"]
    cent, labels = oracle.ivfbin_build(X, 32, 20, 42)
    off, ids = oracle.lists_from_labels(labels, len(cent))
    oi, od = oracle.ivfbin_search(cent, off, ids, X, Q, nprobe, 10)
    assert oi == g[f"ivfbin_idx_p{nprobe}"].tolist()
    assert od == g[f"ivfbin_dist_p{nprobe}"].tolist()


def test_ivfbin_probe_all_is_exact_up_to_order(g):
    X, Q = g["ivfbin_X"], g["ivfbin_Q"]
    cent, labels = oracle.ivfbin_build(X, 32, 20, 42)
    off, ids = oracle.lists_from_labels(labels, len(cent))
    _, od = oracle.ivfbin_search(cent, off, ids, X, Q, 32, 10)
    _, ed = oracle.knn_hamming(X, Q, 10)
    np.testing.assert_array_equal(np.array(od), ed)


def test_kmeans_vs_reference_outputs(g):
    lab, cen = oracle.kmeans_l2(g["km_X"], 12, 100, 42)
    np.testing.assert_array_equal(lab, g["km_labels"])
    np.testing.assert_array_equal(cen, g["km_centroids"])


def test_kmeans_invariants():
    rng = np.random.default_rng(4)
    X = rng.random((500, 5)).astype(np.float32)
    lab, cen = oracle.kmeans_l2(X, 7, 100, 1)
    d = ((X[:, None, :].astype(np.float64) - cen[None].astype(np.float64)) ** 2).sum(-1)
    # converged: every point sits at its nearest centroid; centroids are member means
    assert (lab == d.argmin(1)).mean() > 0.99
    for c in range(7):
        if (lab == c).any():
            np.testing.assert_allclose(cen[c], X[lab == c].mean(0), rtol=1e-4, atol=1e-5)


def test_ivf_l2_numpy_restatement_exact_at_full_probe():
    rng = np.random.default_rng(5)
    X = rng.random((600, 12)).astype(np.float32)
    Q = rng.random((15, 12)).astype(np.float32)
    lab, cen = oracle.kmeans_l2(X, 10, 100, 42)
    data = [X[lab == c] for c in range(10)]
    norms = [(d * d).sum(1) for d in data]
    ids = [np.where(lab == c)[0] for c in range(10)]
    ii, dd = oracle.ivf_l2_search(cen, data, norms, ids, Q, 10, 5)
    fi, fd = oracle.knn_fp64(oracle.L2, X, Q, 5)
    assert [sorted(r) for r in ii] == [sorted(r) for r in fi.tolist()]
    np.testing.assert_allclose(np.sqrt(np.array(dd)), fd, rtol=1e-4)


# ---------------------------------------------------------------- binary widths, dist_*, range, MIH
def _true_hamming(X, Q):
    return np.bitwise_count(Q[:, None, :] ^ X[None, :, :]).sum(-1).astype(np.int64)


@pytest.mark.parametrize("nb", [64, 16, 8, 12, 6, 17])
def test_knn_hamming_vs_reference_width_classes(g, nb):
    """The fixed 512/128/64-bit classes and the generic class (int32 / int16 / int8 accumulators,
    none of which wraps here): distances equal, rows equal below the k-th distance."""
    X, Q = g[f"binw{nb}_X"], g[f"binw{nb}_Q"]
    oi, od = oracle.knn_hamming(X, Q, 4)
    ri, rd = g[f"binw{nb}_idx"][:, ::-1], g[f"binw{nb}_dist"][:, ::-1]
    np.testing.assert_array_equal(od, rd)
    true = _true_hamming(X, Q)
    for q in range(len(Q)):
        assert (true[q, ri[q]] == rd[q]).all()
        below = od[q] < od[q].max()
        assert sorted(oi[q][below]) == sorted(ri[q][rd[q] < rd[q].max()])
    np.testing.assert_array_equal(oracle.knn_hamming(X, Q, 1)[1][:, 0], g[f"binw{nb}_dist1"])


def test_generic_width_33_bytes_reference_wraps_int8(g):
    """At 33 bytes the reference's generic distance sums into int8_t (DistanceFunctions.hpp:125-146):
    its recorded distances are the int8 wrap of the true popcount of the rows it returned, and
    differ from the true k-NN, which the oracle returns."""
    X, Q = g["binw33_X"], g["binw33_Q"]
    true = _true_hamming(X, Q)
    ri, rd = g["binw33_idx"], g["binw33_dist"]
    np.testing.assert_array_equal(np.take_along_axis(true, ri, 1).astype(np.int8), rd)
    oi, od = oracle.knn_hamming(X, Q, 4)
    np.testing.assert_array_equal(od, np.sort(true, 1)[:, :4])
    assert (rd < 0).all() and (od > 0).all()


def test_free_distance_functions_vs_reference(g):
    fa, fb = g["dist_fa"], g["dist_fb"]
    for name, m in (("l2", oracle.L2), ("l1", oracle.L1), ("linf", oracle.LINF)):
        got = np.array([oracle.dist(m, fa[j], fb[j]) for j in range(len(fa))], np.float32)
        np.testing.assert_array_equal(got, g[f"dist_{name}"])
    for nb in (64, 32, 16, 8):
        a, b = g[f"dist_h{nb}_a"], g[f"dist_h{nb}_b"]
        assert [oracle.dist_hamming(a[j], b[j]) for j in range(len(a))] == g[f"dist_h{nb}"].tolist()


def _csr(g, key):
    c = g[f"{key}_count"]
    off = np.concatenate([[0], np.cumsum(c)])
    return [(g[f"{key}_idx"][off[i]:off[i + 1]], g[f"{key}_dist"][off[i]:off[i + 1]]) for i in range(len(c))]


@pytest.mark.parametrize("thr", [8, 40, 52])
def test_hamming_range_vs_reference_bktree(g, thr):
    """BKTreeBinaryIndex128.find_threshold: the same (row, distance) pairs, in any order (the
    reference lists them in its tree's breadth-first order)."""
    oi, od = oracle.hamming_range(g["bk_X"], g["bk_Q"], thr)
    for (ri, rd), a, b in zip(_csr(g, f"bk_thr{thr}"), oi, od):
        assert sorted(zip(ri.tolist(), rd.tolist())) == list(zip(a.tolist(), b.tolist()))
    assert any(len(a) > 1 for a in oi)  # the duplicated row comes back twice


def test_hamming_range_brute_force_small():
    rng = np.random.default_rng(11)
    X = rng.integers(0, 256, (300, 5), dtype=np.uint8)
    Q = np.concatenate([X[:3] ^ np.uint8(1), rng.integers(0, 256, (3, 5), dtype=np.uint8)])
    for thr in (-1, 0, 3, 12, 40):
        oi, od = oracle.hamming_range(X, Q, thr)
        for q in range(len(Q)):
            exp = [(r, bin(int.from_bytes(Q[q].tobytes(), "little") ^ int.from_bytes(X[r].tobytes(), "little")).count("1"))
                   for r in range(len(X))]
            exp = [(r, d) for r, d in exp if d <= thr]
            assert list(zip(oi[q].tolist(), od[q].tolist())) == exp


@pytest.mark.parametrize("m,radius", [(4, 8), (2, 9), (8, 16)])
def test_mih_vs_reference(g, m, radius):
    """MIHBinaryIndex.searchKNN(k=5): distances equal; rows below the k-th distance equal in
    order; rows at the k-th distance are candidates of that distance (the reference keeps them in
    hash-set order)."""
    X, Q = g["bk_X"], g["bk_Q"]
    oi, od = oracle.mih_search(X, Q, m, radius, 5)
    true = _true_hamming(X, Q)
    sb = X.shape[1] // m
    for q, ((ri, rd), a, b) in enumerate(zip(_csr(g, f"mih_m{m}_r{radius}"), oi, od)):
        np.testing.assert_array_equal(rd, b)
        if len(b):
            below = b < b.max()
            np.testing.assert_array_equal(ri[below], a[below])
            assert (true[q, ri] == rd).all()
            sub = np.bitwise_count(X[ri] ^ Q[q]).reshape(len(ri), m, sb).sum(-1)
            assert ((sub <= radius // m).any(1)).all()
    assert sum(len(a) for a in oi) >= 10


def test_mih_invariants_and_planted_mistakes(g):
    """Pigeonhole: every row within the radius is a candidate, so when the k nearest rows are
    within it MIH equals the exact k-NN; a sub-string radius of all bits makes every row a
    candidate. Restatements that use the full radius per sub-string, or sub-strings shifted by
    a byte, disagree with the reference's recorded results."""
    X, Q = g["bk_X"], g["bk_Q"]
    ei, ed = oracle.knn_hamming(X, Q, 5)
    for m, radius in ((4, 8), (2, 9), (8, 16), (16, 15)):
        oi, od = oracle.mih_search(X, Q, m, radius, 5)
        for q in range(len(Q)):
            within = ed[q] <= radius
            if within.all():
                np.testing.assert_array_equal(oi[q], ei[q])
            else:
                assert set(ei[q][within].tolist()) <= set(oi[q].tolist())
    ai, ad = oracle.mih_search(X, Q, 4, 4 * 128, 5)
    np.testing.assert_array_equal(np.array(ai), ei)
    # planted mistakes, checked against the m = 8, radius 16 record
    rec = _csr(g, "mih_m8_r16")
    true = _true_hamming(X, Q)
    sub = np.bitwise_count(Q[:, None, :] ^ X[None]).reshape(len(Q), len(X), 8, 2).sum(-1)
    wrong_radius = (sub <= 16).any(-1)
    shifted = np.bitwise_count(np.roll(Q, 1, 1)[:, None, :] ^ np.roll(X, 1, 1)[None])
    shifted = (shifted[:, :, :16].reshape(len(Q), len(X), 8, 2).sum(-1) <= 2).any(-1)
    for cand in (wrong_radius, shifted):
        bad = 0
        for q, (ri, rd) in enumerate(rec):
            rows = np.flatnonzero(cand[q])
            o = rows[np.lexsort((rows, true[q, rows]))][:5]
            bad += not np.array_equal(true[q, o], rd)
        assert bad > 0


def test_ivf_l2_restatement_partial_probe_vs_fp64():
    """At n_probe < nlist the restatement's neighbours are the fp64 exact top-k over the union of
    the n_probe nearest clusters (fp64 centroid distances), up to fp32 near ties."""
    from neighbours import near_tie_check, nearest_lists, sq_l2_rows
    rng = np.random.default_rng(12)
    centres = rng.standard_normal((12, 10)).astype(np.float32)
    X = (centres[rng.integers(0, 12, 1500)] + 0.4 * rng.standard_normal((1500, 10))).astype(np.float32)
    Q = (centres[rng.integers(0, 12, 40)] + 0.4 * rng.standard_normal((40, 10))).astype(np.float32)
    lab, cen = oracle.kmeans_l2(X, 12, 100, 42)
    keep = [c for c in range(12) if (lab == c).any()]
    data = [X[lab == c] for c in keep]
    norms = [(d * d).sum(1) for d in data]
    ids = [np.where(lab == c)[0] for c in keep]
    for n_probe in (1, 2, 5):
        ii, dd = oracle.ivf_l2_search(cen[keep], data, norms, ids, Q, n_probe, 7)
        chosen, clear = nearest_lists(Q, cen[keep], np.arange(len(keep)), n_probe)
        assert clear.all()
        for q in range(len(Q)):
            rows = np.sort(np.concatenate([ids[c] for c in chosen[q]]))
            # the expanded form |q|^2 + |x|^2 - 2 q.x in fp32 costs ~1e-6 of |x|^2 absolute
            near_tie_check(ii[q], dd[q], rows, sq_l2_rows(X, Q[q], rows), 7, 1e-4, rtol_d=1e-3, atol_d=1e-4)
        # a restatement that probed the farthest clusters instead fails the same check
    far_i, far_d = oracle.ivf_l2_search(cen[keep], data, norms, ids, -Q * 50, 1, 7)
    chosen, _ = nearest_lists(Q, cen[keep], np.arange(len(keep)), 1)
    with pytest.raises(AssertionError):
        for q in range(len(Q)):
            rows = np.sort(ids[chosen[q][0]])
            near_tie_check(far_i[q], np.sort(sq_l2_rows(X, Q[q], np.array(far_i[q]))), rows,
                           sq_l2_rows(X, Q[q], rows), 7, 1e-4, rtol_d=1e-3, atol_d=1e-4)
