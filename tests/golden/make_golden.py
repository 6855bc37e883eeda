"""Regenerate the golden fixtures in this folder (run in a checkout that has the reference).

    make -C oracle ref && python tests/golden/make_golden.py

Writes
  * hamming_known_answers.json -- the Hamming known-answer vectors of the reference's C++ tests
    (pynear/tests/VPTreeTests.cpp:51-98), as data;
  * ref_vectors.npz -- inputs and outputs of the reference's own module (built from its sources
    into oracle/_ref) on small seeded cases: VPTree k-NN / 1-NN for L2, L1, Chebyshev and Hamming,
    the 5-point fixture of pynear/tests/test_vptree.py:282-301, IVFFlatBinaryIndex search,
    kmeans_l2, binary VP-trees at every width class (including the generic class at odd widths),
    the free dist_* functions, BKTreeBinaryIndex128.find_threshold and MIHBinaryIndex.searchKNN.
The fixtures are data only; nothing of the reference's sources is stored.
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def known_answers():
    ff = [255] * 32
    cases = [
        ([255] * 8, [255] * 8, 0),
        ([1] + ff[1:], [2] + ff[1:], 2),
        (ff, [0] + ff[1:], 8),
        ([255, 255, 255, 255, 255, 253] + ff[6:], [0] + ff[1:], 9),
        ([255, 255, 255, 255, 255, 253, 255, 9] + ff[8:], [0, 255, 255, 255, 255, 255, 255, 0] + ff[8:], 11),
        (ff, [0] * 32, 256),
    ]
    return [{"a": a, "b": b, "expected": e} for a, b, e in cases]


def main():
    import oracle

    R = oracle.load_ref()
    if R is None:
        raise SystemExit("reference module not built: make -C oracle ref")
    with open(os.path.join(HERE, "hamming_known_answers.json"), "w") as f:
        json.dump({"source": "pynear/tests/VPTreeTests.cpp:51-98", "cases": known_answers()}, f, indent=1)

    out = {}
    rng = np.random.default_rng(2024)
    # float VP-trees (farthest-first lists as returned by the reference)
    for name, cls in (("l2", R.VPTreeL2Index), ("l1", R.VPTreeL1Index), ("linf", R.VPTreeChebyshevIndex)):
        X = rng.random((2021, 8)).astype(np.float32)
        Q = rng.random((23, 8)).astype(np.float32)
        t = cls()
        t.set(X)
        i, d = t.searchKNN(Q, 5)
        i1, d1 = t.search1NN(Q)
        out.update({f"{name}_X": X, f"{name}_Q": Q, f"{name}_idx": np.array(i, np.int64),
                    f"{name}_dist": np.array(d, np.float32), f"{name}_idx1": np.array(i1, np.int64),
                    f"{name}_dist1": np.array(d1, np.float32)})
    # 5-point fixture (test_vptree.py:282-301)
    X5 = np.array([[-2.5, 0], [-2.58, 0], [0, 0], [2.5, 0], [2.6, 0]], np.float32)
    Q5 = np.array([[-2.55, 0]], np.float32)
    for name, cls in (("l2", R.VPTreeL2Index), ("l1", R.VPTreeL1Index), ("linf", R.VPTreeChebyshevIndex)):
        t = cls()
        t.set(X5)
        i, d = t.searchKNN(Q5, 4)
        out[f"five_{name}_idx"] = np.array(i, np.int64)
        out[f"five_{name}_dist"] = np.array(d, np.float32)
    out["five_X"], out["five_Q"] = X5, Q5
    # binary VP-tree, 256-bit
    Xb = rng.integers(0, 256, (2021, 32), dtype=np.uint8)
    Qb = rng.integers(0, 256, (8, 32), dtype=np.uint8)
    t = R.VPTreeBinaryIndex256()
    t.set(Xb)
    i, d = t.searchKNN(Qb, 3)
    out.update({"bin_X": Xb, "bin_Q": Qb, "bin_idx": np.array(i, np.int64), "bin_dist": np.array(d, np.int64)})
    # IVFFlatBinaryIndex (build + search are deterministic given the seed)
    Xi = rng.integers(0, 256, (1500, 64), dtype=np.uint8)
    Qi = np.concatenate([Xi[rng.integers(0, 1500, 20)] ^ np.uint8(5), rng.integers(0, 256, (12, 64), dtype=np.uint8)])
    for nprobe in (1, 4, 32):
        t = R.IVFFlatBinaryIndex(nlist=32, nprobe=nprobe, max_iter=20, seed=42)
        t.set(Xi)
        i, d = t.searchKNN(Qi, 10)
        out[f"ivfbin_idx_p{nprobe}"] = np.array(i, np.int64)
        out[f"ivfbin_dist_p{nprobe}"] = np.array(d, np.int64)
    out["ivfbin_X"], out["ivfbin_Q"] = Xi, Qi
    # kmeans_l2
    Xk = rng.random((800, 16)).astype(np.float32)
    lab, cen = R.kmeans_l2(Xk, 12, 100, 42)
    out.update({"km_X": Xk, "km_labels": lab, "km_centroids": cen})
    # binary VP-trees at every width class the reference dispatches on: the 512/128/64-bit classes
    # and the generic class at 12 and 6 bytes (int32 / int16 accumulators) and at 17 and 33 bytes,
    # where its int8_t accumulator (DistanceFunctions.hpp:125-146) wraps above 127
    for nb, cls in ((64, R.VPTreeBinaryIndex512), (16, R.VPTreeBinaryIndex128), (8, R.VPTreeBinaryIndex64),
                    (12, R.VPTreeBinaryIndex), (6, R.VPTreeBinaryIndex), (17, R.VPTreeBinaryIndex),
                    (33, R.VPTreeBinaryIndex)):
        Xw = rng.integers(0, 256, (1500, nb), dtype=np.uint8)
        Qw = np.concatenate([Xw[rng.integers(0, 1500, 5)] ^ np.uint8(3), rng.integers(0, 256, (4, nb), dtype=np.uint8)])
        t = cls()
        t.set(Xw)
        i, d = t.searchKNN(Qw, 4)
        i1, d1 = t.search1NN(Qw)
        out.update({f"binw{nb}_X": Xw, f"binw{nb}_Q": Qw, f"binw{nb}_idx": np.array(i, np.int64),
                    f"binw{nb}_dist": np.array(d, np.int64), f"binw{nb}_idx1": np.array(i1, np.int64),
                    f"binw{nb}_dist1": np.array(d1, np.int64)})
    # the free distance functions
    fa = rng.standard_normal((7, 13)).astype(np.float32)
    fb = rng.standard_normal((7, 13)).astype(np.float32)
    out.update({"dist_fa": fa, "dist_fb": fb,
                "dist_l2": np.array([R.dist_l2(fa[j], fb[j]) for j in range(7)], np.float32),
                "dist_l1": np.array([R.dist_l1(fa[j], fb[j]) for j in range(7)], np.float32),
                "dist_linf": np.array([R.dist_chebyshev(fa[j], fb[j]) for j in range(7)], np.float32)})
    for nb, fn in ((64, R.dist_hamming_512), (32, R.dist_hamming_256), (16, R.dist_hamming_128), (8, R.dist_hamming_64)):
        ha = rng.integers(0, 256, (5, nb), dtype=np.uint8)
        hb = rng.integers(0, 256, (5, nb), dtype=np.uint8)
        out.update({f"dist_h{nb}_a": ha, f"dist_h{nb}_b": hb,
                    f"dist_h{nb}": np.array([fn(list(ha[j]), list(hb[j])) for j in range(5)], np.int64)})
    # BK-tree range search (find_threshold; results in its tree order) and MIH top-k over the
    # candidates of its sub-string tables
    Xt = rng.integers(0, 256, (3000, 16), dtype=np.uint8)
    Xt[100] = Xt[7]  # an exact duplicate
    src = np.concatenate([[7], rng.integers(0, 3000, 13)])
    Qt = Xt[src].copy()
    for j, nf in enumerate((0, 1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 30)):  # distinct bits flipped
        bits = rng.permutation(128)[:nf]
        Qt[j, bits // 8] ^= (np.uint8(1) << (bits % 8).astype(np.uint8))
    Qt = np.concatenate([Qt, rng.integers(0, 256, (4, 16), dtype=np.uint8)])
    for thr in (8, 40, 52):
        t = R.BKTreeBinaryIndex128()
        t.set(Xt)
        ti, td, _ = t.find_threshold(Qt, thr)
        out[f"bk_thr{thr}_count"] = np.array([len(r) for r in ti], np.int64)
        out[f"bk_thr{thr}_idx"] = np.array([v for r in ti for v in r], np.int64)
        out[f"bk_thr{thr}_dist"] = np.array([v for r in td for v in r], np.int64)
    out["bk_X"], out["bk_Q"] = Xt, Qt
    for m_, radius in ((4, 8), (2, 9), (8, 16)):
        mi = R.MIHBinaryIndex(m_)
        mi.set(Xt)
        ii, dd = mi.searchKNN(Qt, 5, radius)
        out[f"mih_m{m_}_r{radius}_count"] = np.array([len(r) for r in ii], np.int64)
        out[f"mih_m{m_}_r{radius}_idx"] = np.array([v for r in ii for v in r], np.int64)
        out[f"mih_m{m_}_r{radius}_dist"] = np.array([v for r in dd for v in r], np.int64)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
