"""Batched exact L2 search over uneven indices: every sub-search must equal the plain search of its
index bit for bit (indices and distances), before and after autotuning, on every graph replay."""

import numpy as np
import pytest

import oracle
from pynear_amd import VPTreeL2Index
from pynear_amd.batch import BatchSearch


def _make(seed, specs):
    rng = np.random.default_rng(seed)
    out = []
    for n, d, nq, k in specs:
        X = rng.standard_normal((n, d), dtype=np.float32)
        Q = rng.standard_normal((nq, d), dtype=np.float32)
        idx = VPTreeL2Index()
        idx.set(X)
        out.append((idx, X, Q, k))
    return out


# (n, d, queries, k); indices of at most 32 / 64 queries take one / two query groups per scan
# workgroup (its waves on different slices): wide rows (d 128 .. 256) and narrow ones, one or many
# slices, a slice group with empty trailing slices (300 rows)
SPECS = [(1000, 64, 7, 1), (5000, 128, 300, 10), (20000, 256, 40, 32), (3000, 64, 1, 10), (777, 100, 129, 5),
         (64, 16, 20, 32), (40, 32, 5, 32), (12000, 128, 257, 1), (2500, 8, 33, 12), (30000, 256, 3, 1),
         (9000, 192, 64, 10), (7000, 64, 33, 5), (300, 256, 1, 10)]


def _plain(idx, Q, k):
    i, d = idx._h.search(Q, k)
    return i, d


@pytest.mark.gpu
def test_batch_equals_plain_searches():
    subs = _make(0, SPECS)
    b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
    for rep in range(3):  # graph replays
        b.run()
        res = b.results()
        for (idx, X, Q, k), (bi, bd) in zip(subs, res):
            pi, pd = _plain(idx, Q, k)
            assert np.array_equal(bi, pi), (rep, X.shape, Q.shape, k)
            assert np.array_equal(bd, pd)


@pytest.mark.gpu
def test_batch_matches_oracle():
    subs = _make(1, [(3000, 64, 50, 10), (800, 128, 17, 32), (1500, 20, 9, 1)])
    b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
    b.run()
    for (idx, X, Q, k), (bi, bd) in zip(subs, b.results()):
        ri, rd = oracle.knn_float(oracle.L2, X, Q, k)
        assert np.array_equal(bi, ri)
        assert np.array_equal(bd, rd)


@pytest.mark.gpu
@pytest.mark.parametrize("lean,fuse", [(0, 1), (2, 1), (1, 0), (2, 0)])
def test_batch_merge_variants(lean, fuse):
    """The merge with private lists (one launch per list size), the lean merge's rounds path, and
    merge and rerank in separate launches give the plain searches' results (the default, the lean
    merge fused with the rerank, is covered by the other tests)."""
    from pynear_amd import _native as N

    subs = _make(4, SPECS[:8])
    old = (N.tune_get("merge.lean"), N.tune_get("l2.fuse_rerank"))
    try:
        N.tune_set("merge.lean", lean)
        N.tune_set("l2.fuse_rerank", fuse)
        b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
        b.run()
        res = b.results()
    finally:
        N.tune_set("merge.lean", old[0])
        N.tune_set("l2.fuse_rerank", old[1])
    for (idx, X, Q, k), (bi, bd) in zip(subs, res):
        pi, pd = _plain(idx, Q, k)
        assert np.array_equal(bi, pi), (X.shape, Q.shape, k)
        assert np.array_equal(bd, pd)


@pytest.mark.gpu
@pytest.mark.parametrize("slice_rows,order,share,pipe", [(32, 2, 1, 3), (256, 0, 1, 2), (2048, 2, 1, 3),
                                                         (2048, 2, 1, 0), (4096, 2, 1, 1), (4096, 1, 0, 3),
                                                         (8192, 2, 1, 3)])
def test_batch_schedules_keep_results(slice_rows, order, share, pipe):
    """The slice lists hold keys only: a key's low bits carry its row's offset in the slice (5 bits
    for 32-row slices up to 12 for 4,096; longer settings are cut to 4,096), and the slices of a
    query share the bound of their lists (in index order, longest first, or slice by slice so that
    later slices start from finished ones), and the wide and narrow rows go through the
    software-pipelined scans or not (pipe bits 1 and 2). Every schedule returns the plain
    searches' results,
    bit for bit, on indices of many slices (60,000 and 45,000 rows, k = 32 and 10) and on the
    wide-row scan's edge cases."""
    from pynear_amd import _native as N

    # (the wide-row class with more than 64 queries at 192 and 256 columns and k = 32, a partial
    # last tile, an index of one partial tile and one of a full and a partial one)
    subs = _make(5, SPECS[:6] + [(60000, 64, 200, 32), (45000, 256, 300, 10), (20000, 128, 90, 1),
                                 (7001, 192, 150, 32), (9001, 256, 130, 32), (31, 128, 129, 1), (33, 256, 200, 10)])
    names = ("batch.slice_rows", "batch.order", "batch.share", "batch.pipe")
    old = [N.tune_get(n) for n in names]
    try:
        for n, v in zip(names, (slice_rows, order, share, pipe)):
            N.tune_set(n, v)
        b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
        for _ in range(2):
            b.run()
        res = b.results()
    finally:
        for n, v in zip(names, old):
            N.tune_set(n, v)
    for (idx, X, Q, k), (bi, bd) in zip(subs, res):
        pi, pd = _plain(idx, Q, k)
        assert np.array_equal(bi, pi), (slice_rows, order, share, pipe, X.shape, Q.shape, k)
        assert np.array_equal(bd, pd)


@pytest.mark.gpu
def test_batch_fallback_many_flagged_queries():
    """An index of identical rows ties every distance, so none of its 1,500 queries can be proved
    complete from the candidates: all go to the fallback scan (more queries than fallback
    workgroups, so each takes several in turn). Results are the oracle's: rows 0 .. k-1."""
    rng = np.random.default_rng(8)
    Xe = np.tile(rng.standard_normal((1, 64)).astype(np.float32), (3000, 1))
    Qe = rng.standard_normal((1500, 64)).astype(np.float32)
    ie = VPTreeL2Index()
    ie.set(Xe)
    subs = [(ie, Xe, Qe, 10)] + _make(9, [(2000, 128, 100, 5)])
    b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
    b.run()
    for (idx, X, Q, k), (bi, bd) in zip(subs, b.results()):
        ri, rd = oracle.knn_float(oracle.L2, X, Q, k)
        assert np.array_equal(bi, ri)
        assert np.array_equal(bd, rd)


@pytest.mark.gpu
def test_batch_tuned_schedule_keeps_results(tmp_path):
    subs = _make(2, SPECS[:5])
    b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
    b.run()
    before = b.results()
    table = str(tmp_path / "tune.txt")
    b._plan.tune(table, reps=2)
    tiles, cfg = b.info()
    assert tiles > 0 and cfg["slice_rows"] in (128, 256, 512, 1024, 2048, 4096)
    assert open(table).read().split()[1:4] == [str(cfg["slice_rows"]), str(cfg["wg_per_cu"]), str(cfg["order"])]
    b.run()
    for (bi, bd), (ai, ad) in zip(b.results(), before):
        assert np.array_equal(bi, ai) and np.array_equal(bd, ad)
    # a second plan of the same shape takes the stored schedule without measuring
    b2 = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs], tune=True, table_path=table)
    assert b2.info()[1] == cfg


@pytest.mark.gpu
def test_batch_rejects_unsupported():
    subs = _make(3, [(100, 8, 3, 5)])
    idx, X, Q, _ = subs[0]
    with pytest.raises(ValueError):
        BatchSearch([(idx, Q, 33)])
    with pytest.raises(ValueError):
        BatchSearch([(idx, Q[:, :4], 3)])


@pytest.mark.gpu
def test_batch_refuses_to_run_after_an_index_is_set_again():
    """The plan binds each index's device rows: set() on an index after planning (a larger set
    reallocates the rows, an empty one releases them) makes run() and tune() raise instead of
    replaying over stale memory; a new plan over the same indices serves the new rows."""
    subs = _make(5, [(2000, 64, 20, 5), (1500, 128, 10, 10)])
    b = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
    b.run()
    idx0, X0, Q0, k0 = subs[0]
    X_new = np.random.default_rng(6).standard_normal((5000, 64), dtype=np.float32)
    idx0.set(X_new)
    with pytest.raises(ValueError, match="set"):
        b.run()
    with pytest.raises(ValueError, match="set"):
        b._plan.tune(None, reps=1)
    b2 = BatchSearch([(idx, Q, k) for idx, X, Q, k in subs])
    b2.run()
    bi, bd = b2.results()[0]
    ri, rd = oracle.knn_float(oracle.L2, X_new, Q0, k0)
    assert np.array_equal(bi, ri) and np.array_equal(bd, rd)
