"""CPU tests that pin the oracle (oracle/) independently of its own restatement.

The reference ships no golden vectors or fixtures (its drivers only print a
relative error against a dense cuBLAS product, hmglib_test.cu:124-128), so each
function of the oracle is pinned by what the mathematics fixes: published
known-answer vectors, brute force on small inputs, closed forms and invariants.
"""
import collections

import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------- generator
def test_splitmix64_known_answer():
    # first outputs of Vigna's splitmix64.c for seed 1234567
    assert O.splitmix64(1234567, 5) == [6457827717110365317, 3203168211198807973,
                                        9817491932198370423, 4593380528125082431,
                                        16408922859458223821]


def test_uniform_range_and_moments():
    u = O.uniform(2001, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5e-3
    assert abs(u.var() - 1.0 / 12.0) < 5e-3
    # 53-bit granularity
    assert np.all(u * 2.0 ** 53 == np.floor(u * 2.0 ** 53))
    z = O.normal(2002, 200001)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1.0) < 1e-2


def test_product_generator_matches_oracle():
    from hmglib_amd.hmatrix import rng_normal, rng_uniform
    np.testing.assert_array_equal(rng_uniform(1001, 10001), O.uniform(1001, 10001))
    np.testing.assert_array_equal(rng_normal(3001, 999), O.normal(3001, 999))


# ---------------------------------------------------------------- Morton
def _interleave(q, dim, bits):
    out = np.zeros(q.shape[-1], dtype=np.uint64)
    for d in range(dim):
        for b in range(bits):
            bit = (q[d] >> np.uint64(b)) & np.uint64(1)
            out |= bit << np.uint64(dim * b + d)
    return out


def test_stretch_exhaustive_3d_and_sampled_2d():
    rng = np.random.default_rng(0)
    vals3 = list(range(0, 1 << 20, 4099)) + [0, 1, (1 << 20) - 1, 0xAAAAA, 0x55555]
    for v in vals3:
        assert O.stretch(v, 3, 20) == sum(((v >> b) & 1) << (3 * b) for b in range(20))
    for v in list(rng.integers(0, 1 << 32, 2000)) + [0, (1 << 32) - 1, 0xAAAAAAAA]:
        v = int(v)
        assert O.stretch(v, 2, 32) == sum(((v >> b) & 1) << (2 * b) for b in range(32))
    # undefined (dim,bits) pairs give 0 (morton.cu:114)
    assert O.stretch(12345, 3, 21) == 0
    assert O.stretch(7, 10, 4) == sum(((7 >> b) & 1) << (10 * b) for b in range(4))


@pytest.mark.parametrize("dim,bits", [(2, 32), (3, 20)])
def test_morton_quantisation_closed_form(dim, bits):
    n = 5000
    X = O.uniform(7 + dim, dim * n).reshape(dim, n)
    codes = O.morton(X, bits)
    mn, mx = X.min(axis=1), X.max(axis=1)
    q = np.empty((dim, n), dtype=np.uint64)
    for d in range(dim):
        t = (X[d] - mn[d]) / (mx[d] - mn[d])
        qq = np.floor(t * float(1 << bits)).astype(np.uint64)
        q[d] = np.minimum(qq, np.uint64((1 << bits) - 1))
    np.testing.assert_array_equal(codes, _interleave(q, dim, bits))
    # the maximum point maps to the last cell, the minimum to cell 0
    assert np.all(q.max(axis=1) == (1 << bits) - 1) and np.all(q.min(axis=1) == 0)


def test_sort_is_stable():
    rng = np.random.default_rng(1)
    codes = rng.integers(0, 50, 20000).astype(np.uint64)
    s, order = O.sort_codes(codes)
    ref = np.argsort(codes, kind="stable")
    np.testing.assert_array_equal(order, ref.astype(np.uint64))
    np.testing.assert_array_equal(s, codes[ref])


def _split_brute(codes, first, last):
    c = [int(v) for v in codes]
    if c[first] == c[last]:
        return (first + last) >> 1
    common = 64 - (c[first] ^ c[last]).bit_length()
    best = first
    for i in range(first, last):
        if 64 - (c[first] ^ c[i]).bit_length() > common:
            best = i
    return best


def test_find_split_brute_force():
    rng = np.random.default_rng(2)
    for trial in range(300):
        n = int(rng.integers(2, 60))
        codes = np.sort(rng.integers(0, 1 << int(rng.integers(3, 62)), n).astype(np.uint64))
        a = int(rng.integers(0, n - 1))
        b = int(rng.integers(a + 1, n))
        assert O.find_split(codes, a, b) == _split_brute(codes, a, b)


# ---------------------------------------------------------------- tree
def _admissible(b1, b2, eta):
    """tree.cu:121-160 with the diameter accumulated as fma(t, t, acc)."""
    def diam(lo, hi):
        acc = 0.0
        for t in hi - lo:
            acc = _fma(float(t), float(t), acc)
        return np.sqrt(acc)
    (lo1, hi1), (lo2, hi2) = b1, b2
    dist = 0.0
    for d in range(lo1.size):
        a = max(0.0, lo1[d] - hi2[d])
        b = max(0.0, lo2[d] - hi1[d])
        dist = dist + (a * a + b * b)
    return max(diam(lo1, hi1), diam(lo2, hi2)) <= eta * np.sqrt(dist)


def _fma(a, b, c):
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _tree_brute(X1s, X2s, codes1, codes2, eta, max_level, c_leaf, roots=(0, 0)):
    """Level-by-level traversal with exact boxes computed directly from the points.
    Nodes whose cluster levels differ split only the side with the lower level
    into two children, whatever their admissibility (tree.cu:258-264, 395-444);
    empty clusters (a one-point cluster split at its middle) are dropped."""
    def box(X, l, u):
        return X[:, l:u + 1].min(axis=1), X[:, l:u + 1].max(axis=1)

    cur = [(0, X1s.shape[1] - 1, 0, X2s.shape[1] - 1, roots[0], roots[1])]
    leaves = []
    for level in range(max_level):
        nxt = []
        for (l1, u1, l2, u2, v1, v2) in cur:
            if l1 > u1 or l2 > u2:
                continue
            if v1 != v2:
                if v1 < v2:
                    s1 = _split_brute(codes1, l1, u1)
                    nxt += [(l1, s1, l2, u2, v1 + 1, v2), (s1 + 1, u1, l2, u2, v1 + 1, v2)]
                else:
                    s2 = _split_brute(codes2, l2, u2)
                    nxt += [(l1, u1, l2, s2, v1, v2 + 1), (l1, u1, s2 + 1, u2, v1, v2 + 1)]
                continue
            adm = _admissible(box(X1s, l1, u1), box(X2s, l2, u2), eta)
            if adm or (u1 - l1 + 1) <= c_leaf or (u2 - l2 + 1) <= c_leaf or level + 1 >= max_level:
                leaves.append((l1, u1, l2, u2, O.WT_ACA if adm else O.WT_DENSE, v1, v2))
            else:
                s1 = _split_brute(codes1, l1, u1)
                s2 = _split_brute(codes2, l2, u2)
                nxt += [(l1, s1, l2, s2, v1 + 1, v2 + 1), (l1, s1, s2 + 1, u2, v1 + 1, v2 + 1),
                        (s1 + 1, u1, l2, s2, v1 + 1, v2 + 1), (s1 + 1, u1, s2 + 1, u2, v1 + 1, v2 + 1)]
        cur = nxt
        if not cur:
            break
    dense = [b for b in leaves if b[4] == O.WT_DENSE]
    return dense + [b for b in leaves if b[4] == O.WT_ACA]


@pytest.mark.parametrize("dim,bits,n1,n2,eta,c_leaf,seed", [
    (2, 32, 300, 300, 1.5, 8, 1), (2, 32, 400, 250, 0.7, 8, 2), (3, 20, 350, 350, 1.0, 16, 3),
    (3, 20, 200, 500, 2.0, 8, 4)])
def test_tree_matches_brute_force(dim, bits, n1, n2, eta, c_leaf, seed):
    X1 = O.uniform(seed, dim * n1).reshape(dim, n1)
    X2 = X1 if n1 == n2 else O.uniform(seed + 100, dim * n2).reshape(dim, n2)
    hm = O.HMatrix(X1, X2, bits, eta, 50, c_leaf)
    diag = hm.diagnostics()
    assert diag["box_differs"] == 0 and diag["undefined_table"] == 0
    ref = _tree_brute(hm.sorted_coords(0), hm.sorted_coords(1), hm.codes(0), hm.codes(1), eta, 50,
                      c_leaf)
    got = hm.blocks()
    got_t = [(int(b["set1_l"]), int(b["set1_u"]), int(b["set2_l"]), int(b["set2_u"]),
              int(b["work_type"]), int(b["level_1"]), int(b["level_2"])) for b in got]
    assert got_t == ref
    assert hm.ndense == sum(1 for b in ref if b[4] == O.WT_DENSE)


@pytest.mark.parametrize("dim,bits,n1,n2,eta,c_leaf,roots,seed", [
    (2, 32, 300, 260, 1.2, 16, (0, 2), 5), (2, 32, 300, 260, 1.2, 16, (3, 1), 6),
    (2, 32, 300, 260, 1.2, 16, (1, 0), 7), (2, 32, 300, 260, 1.2, 24, (0, 2), 10),
    (3, 20, 320, 280, 1.2, 16, (1, 0), 5)])
def test_tree_unequal_root_levels_matches_brute_force(dim, bits, n1, n2, eta, c_leaf, roots, seed):
    """Root levels that differ: the traversal first splits the lower-level side in
    two until the levels meet (tree.cu:258-264, 395-444), then proceeds as usual;
    block list, work types and cluster levels against the brute-force traversal."""
    X1 = O.uniform(seed, dim * n1).reshape(dim, n1)
    X2 = O.uniform(seed + 100, dim * n2).reshape(dim, n2)
    hm = O.HMatrix(X1, X2, bits, eta, 50, c_leaf, *roots)
    diag = hm.diagnostics()
    assert diag["box_differs"] == 0 and diag["undefined_table"] == 0
    ref = _tree_brute(hm.sorted_coords(0), hm.sorted_coords(1), hm.codes(0), hm.codes(1), eta, 50,
                      c_leaf, roots)
    got = hm.blocks()
    got_t = [(int(b["set1_l"]), int(b["set1_u"]), int(b["set2_l"]), int(b["set2_u"]),
              int(b["work_type"]), int(b["level_1"]), int(b["level_2"])) for b in got]
    assert got_t == ref
    # every leaf was reached after the levels met
    assert all(b[5] == b[6] for b in ref) and hm.naca > 0


def test_tree_invariants_tiling_and_admissibility():
    n = 4096
    X = O.uniform(1001, 2 * n).reshape(2, n)
    hm = O.HMatrix(X, X, 32, 1.5, 50, 64)
    b = hm.blocks()
    # blocks tile [0,N1) x [0,N2) exactly once
    cover = np.zeros((n, n), dtype=np.int8)
    for w in b:
        cover[w["set1_l"]:w["set1_u"] + 1, w["set2_l"]:w["set2_u"] + 1] += 1
    assert cover.min() == 1 and cover.max() == 1
    Xs = hm.sorted_coords(0)
    for w in b:
        l1, u1, l2, u2 = (int(w[f]) for f in ("set1_l", "set1_u", "set2_l", "set2_u"))
        b1 = (Xs[:, l1:u1 + 1].min(1), Xs[:, l1:u1 + 1].max(1))
        b2 = (Xs[:, l2:u2 + 1].min(1), Xs[:, l2:u2 + 1].max(1))
        np.testing.assert_array_equal(w["min1"][:2], b1[0])
        np.testing.assert_array_equal(w["max2"][:2], b2[1])
        if w["work_type"] == O.WT_ACA:
            assert _admissible(b1, b2, 1.5)
        else:
            assert (u1 - l1 + 1 <= 64 or u2 - l2 + 1 <= 64) and not _admissible(b1, b2, 1.5)
    # DENSE first, then ACA
    wt = b["work_type"]
    assert np.all(wt[:hm.ndense] == O.WT_DENSE) and np.all(wt[hm.ndense:] == O.WT_ACA)


@pytest.mark.parametrize("dim,bits,n,c_leaf,eta", [(2, 32, 4096, 64, 1.5), (3, 20, 6000, 128, 1.0),
                                                   (2, 32, 3000, 8, 1.5)])
def test_same_sets_block_list_is_transpose_closed(dim, bits, n, c_leaf, eta):
    """With the same points on both sides the traversal is symmetric: (C,D) is a
    leaf of a work type iff (D,C) is, and off-diagonal DENSE leaves have disjoint
    clusters.  The engine's symmetric near field (each transposed DENSE pair
    evaluated once) relies on exactly this; it also checks the kernel matrix is
    bitwise symmetric on such a pair."""
    X = O.uniform(7 + dim, dim * n).reshape(dim, n)
    hm = O.HMatrix(X, X, bits, eta, 50, c_leaf)
    b = hm.blocks()
    key = {(int(w["set1_l"]), int(w["set1_u"]), int(w["set2_l"]), int(w["set2_u"])): int(w["work_type"])
           for w in b}
    offdiag = 0
    for (l1, u1, l2, u2), wt in key.items():
        assert key.get((l2, u2, l1, u1)) == wt
        if (l1, u1) != (l2, u2):
            assert u1 < l2 or u2 < l1
            offdiag += wt == O.WT_DENSE
    assert offdiag > 0
    ok = O.kernel(O.K_EXPONENTIAL, dim, 0.0, 0.1)
    for (l1, u1, l2, u2), wt in list(key.items())[:50]:
        if wt == O.WT_DENSE and l1 != l2:
            for i, j in ((l1, l2), (u1, u2), (l1, u2)):
                assert hm.kernel_entry(ok, i, j) == hm.kernel_entry(ok, j, i)


def test_single_block_when_sets_are_small():
    X = O.uniform(5, 2 * 40).reshape(2, 40)
    hm = O.HMatrix(X, X, 32, 1.5, 50, 64)
    b = hm.blocks()
    assert len(b) == 1 and b[0]["work_type"] == O.WT_DENSE
    assert (b[0]["set1_l"], b[0]["set1_u"], b[0]["set2_l"], b[0]["set2_u"]) == (0, 39, 0, 39)


def _quirk_points(n):
    """One point at the origin and n-1 points in [0.6,1]^2: the first split of the
    Morton order isolates the origin in a size-1 cluster."""
    rest = 0.6 + 0.4 * O.uniform(77, 2 * (n - 1)).reshape(2, n - 1)
    X = np.concatenate([np.zeros((2, 1)), rest], axis=1)
    X[:, 1] = 1.0  # pin the extent to [0,1]^2
    return X


def _literal_level_boxes(X, ranges):
    """Boxes of one level's clusters as the reference's lookup table forms them
    (tree.cu:852-1147, the rule written out in the hmglib survey's appendix A.4),
    restated independently with numpy: keys +t+1 at each distinct lower bound,
    then -(t+1) at each distinct upper bound (a one-point cluster keeps the minus),
    an inclusive prefix sum, +t+1 written back at each upper bound, maximal runs of
    equal keys with key 0 dropped, the k-th run's min/max as table entry k, and
    table entry t for the clusters starting at the t-th lower bound."""
    n = X.shape[1]
    L = sorted({l for l, u in ranges})
    U = sorted({u for l, u in ranges})
    assert len(L) == len(U)
    key = np.zeros(n, dtype=np.int64)
    for t in range(len(L)):
        key[L[t]] = t + 1
        key[U[t]] = -(t + 1)
    key = np.cumsum(key)
    for t in range(len(U)):
        key[U[t]] = t + 1
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    ends = np.r_[starts[1:], n]
    table = [(X[:, a:b].min(axis=1), X[:, a:b].max(axis=1)) for a, b in zip(starts, ends) if key[a] != 0]
    assert len(table) >= len(L)   # no entry read past the last run
    where = {l: t for t, l in enumerate(L)}
    return {(l, u): table[where[l]] for l, u in ranges}


def _tree_brute_literal(X1s, X2s, codes1, codes2, eta, max_level, c_leaf):
    """The brute-force traversal with every level's boxes from _literal_level_boxes
    (equal root levels); returns the leaves with their boxes, DENSE first."""
    cur = [(0, X1s.shape[1] - 1, 0, X2s.shape[1] - 1)]
    leaves = []
    for level in range(max_level):
        cur = [c for c in cur if c[0] <= c[1] and c[2] <= c[3]]
        if not cur:
            break
        b1 = _literal_level_boxes(X1s, [(c[0], c[1]) for c in cur])
        b2 = _literal_level_boxes(X2s, [(c[2], c[3]) for c in cur])
        nxt = []
        for (l1, u1, l2, u2) in cur:
            bx1, bx2 = b1[(l1, u1)], b2[(l2, u2)]
            adm = _admissible(bx1, bx2, eta)
            if adm or (u1 - l1 + 1) <= c_leaf or (u2 - l2 + 1) <= c_leaf or level + 1 >= max_level:
                leaves.append((l1, u1, l2, u2, O.WT_ACA if adm else O.WT_DENSE, level, bx1, bx2))
            else:
                s1 = _split_brute(codes1, l1, u1)
                s2 = _split_brute(codes2, l2, u2)
                nxt += [(l1, s1, l2, s2), (l1, s1, s2 + 1, u2), (s1 + 1, u1, l2, s2), (s1 + 1, u1, s2 + 1, u2)]
        cur = nxt
    return [b for b in leaves if b[4] == O.WT_DENSE] + [b for b in leaves if b[4] == O.WT_ACA]


@pytest.mark.parametrize("n,c_leaf,eta,seed", [(400, 2, 1.2, 91), (600, 1, 1.0, 92), (500, 3, 1.5, 93)])
def test_quirk_literal_boxes_match_brute_force(n, c_leaf, eta, seed):
    """Clustered points and tiny leaves make one-point clusters on many levels, so
    the lookup table shifts boxes (the reference's size-1 quirk).  The whole block
    list, boxes included (bit for bit), against the brute-force traversal that
    forms every level's boxes with the literal key/scan rule."""
    X = O.uniform(seed, 2 * n).reshape(2, n) ** 3
    hm = O.HMatrix(X, X, 32, eta, 50, c_leaf)
    d = hm.diagnostics()
    assert d["singleton_levels"] >= 2 and d["box_differs"] > 0 and d["undefined_table"] == 0
    ref = _tree_brute_literal(hm.sorted_coords(0), hm.sorted_coords(1), hm.codes(0), hm.codes(1), eta, 50,
                              c_leaf)
    got = hm.blocks()
    assert len(got) == len(ref)
    for w, r in zip(got, ref):
        assert (int(w["set1_l"]), int(w["set1_u"]), int(w["set2_l"]), int(w["set2_u"]), int(w["work_type"]),
                int(w["level_1"])) == r[:6]
        np.testing.assert_array_equal(w["min1"][:2], r[6][0])
        np.testing.assert_array_equal(w["max1"][:2], r[6][1])
        np.testing.assert_array_equal(w["min2"][:2], r[7][0])
        np.testing.assert_array_equal(w["max2"][:2], r[7][1])


@pytest.mark.parametrize("n,c_leaf,eta,seed", [(400, 2, 1.2, 91), (600, 1, 1.0, 92), (500, 3, 1.5, 93)])
def test_exact_boxes_mode_matches_brute_force(n, c_leaf, eta, seed):
    """The exact-box mode (a cluster's box is the min/max of its own points, no
    lookup table) on the clouds where the literal mode shifts boxes: the whole
    block list, boxes included, against the brute-force traversal with exact boxes;
    every ACA block admissible and every DENSE block small or inadmissible under
    exact boxes; and the block list differs from the literal mode's."""
    X = O.uniform(seed, 2 * n).reshape(2, n) ** 3
    hm = O.HMatrix(X, X, 32, eta, 50, c_leaf, exact_boxes=True)
    lit = O.HMatrix(X, X, 32, eta, 50, c_leaf)
    assert lit.diagnostics()["box_differs"] > 0
    ref = _tree_brute(hm.sorted_coords(0), hm.sorted_coords(1), hm.codes(0), hm.codes(1), eta, 50, c_leaf)
    got = hm.blocks()
    got_t = [(int(b["set1_l"]), int(b["set1_u"]), int(b["set2_l"]), int(b["set2_u"]),
              int(b["work_type"]), int(b["level_1"]), int(b["level_2"])) for b in got]
    assert got_t == ref
    Xs = hm.sorted_coords(0)
    for w in got:
        l1, u1, l2, u2 = (int(w[f]) for f in ("set1_l", "set1_u", "set2_l", "set2_u"))
        b1 = (Xs[:, l1:u1 + 1].min(1), Xs[:, l1:u1 + 1].max(1))
        b2 = (Xs[:, l2:u2 + 1].min(1), Xs[:, l2:u2 + 1].max(1))
        np.testing.assert_array_equal(w["min1"][:2], b1[0])
        np.testing.assert_array_equal(w["max1"][:2], b1[1])
        np.testing.assert_array_equal(w["min2"][:2], b2[0])
        np.testing.assert_array_equal(w["max2"][:2], b2[1])
        if w["work_type"] == O.WT_ACA:
            assert _admissible(b1, b2, eta)
    lb = lit.blocks()
    assert len(lb) != len(got) or not np.array_equal(lb["work_type"], got["work_type"]) or \
        not np.array_equal(lb["min1"], got["min1"])


def test_quirk_hand_worked_example():
    """tree.cu:1000-1024 with a singleton cluster t=0 = [0,0] and t=1 = [1,N-1]:
    keys[0] = +1 then -1 (same thread); keys[1] = +2; keys[N-1] = -2; inclusive
    scan gives [-1, 1, ..., 1, -1]; fix-up keys[0] = 1, keys[N-1] = 2; runs are
    [0, N-2] (key 1) and [N-1] (key 2).  So the singleton's box is the box of
    points 0..N-2 and the big cluster's box is the single point N-1."""
    n = 64
    X = _quirk_points(n)
    hm = O.HMatrix(X, X, 32, 1.0, 2, 0)
    Xs = hm.sorted_coords(0)
    assert np.all(Xs[:, 0] == 0.0)
    b = hm.blocks()
    assert len(b) == 4  # level-1 nodes become leaves at max_level 2
    box_run0 = (Xs[:, :n - 1].min(1), Xs[:, :n - 1].max(1))
    box_run1 = (Xs[:, n - 1], Xs[:, n - 1])
    for w in b:
        for s, (lo_f, hi_f, l_f) in enumerate((("min1", "max1", "set1_l"), ("min2", "max2", "set2_l"))):
            expect = box_run0 if w[l_f] == 0 else box_run1
            np.testing.assert_array_equal(w[lo_f][:2], expect[0])
            np.testing.assert_array_equal(w[hi_f][:2], expect[1])
    d = hm.diagnostics()
    assert d["singleton_levels"] >= 2 and d["box_differs"] > 0


# ---------------------------------------------------------------- batch sizing
@pytest.mark.parametrize("max_aca", [65536, 1 << 30, 3000, 200, 10])
def test_batch_sizes_are_greedy_maximal_partitions(max_aca):
    """Batch sizing of setup (hmglib.cu:211-219, linear_algebra.cu:4059-4288).
    The dense cap is 1.05 x the dense entries, so the dense walk takes every DENSE
    block in one batch and then records its terminating 0.  The ACA walk cuts the
    ACA blocks, in block order, into batches whose sums of m1 stay within the cap,
    each as long as possible (the next block would cross the cap), ending with a
    recorded 0; a first block larger than the cap gives [0] (the reference then
    builds no ACA block at all, hmglib.cu:232)."""
    n = 4096
    X = O.uniform(1001, 2 * n).reshape(2, n)
    hm = O.HMatrix(X, X, 32, 1.5, 50, 64)
    b = hm.blocks()
    m1 = (b["set1_u"] - b["set1_l"] + 1).astype(np.int64)
    m2 = (b["set2_u"] - b["set2_l"] + 1).astype(np.int64)
    cap, dws, aws = hm.work_sizes(max_aca)
    e_near = int((m1 * m2)[:hm.ndense].sum())
    assert e_near == hm.work()[0]
    assert cap == int(e_near * 1.05)
    assert list(dws) == [hm.ndense, 0]
    assert aws[-1] == 0 and np.all(aws[:-1] > 0)
    m1a = m1[hm.ndense:]
    if m1a[0] > max_aca:
        assert list(aws) == [0]
        return
    assert aws.sum() == hm.naca
    at = 0
    for w in aws[:-1]:
        assert m1a[at:at + w].sum() <= max_aca
        if at + w < hm.naca:
            assert m1a[at:at + w + 1].sum() > max_aca
        at += w


# ---------------------------------------------------------------- kernels
def test_kernel_entries_closed_form():
    n = 50
    X1 = O.uniform(3, 3 * n).reshape(3, n)
    X2 = O.uniform(4, 3 * n).reshape(3, n)
    hm = O.HMatrix(X1, X2, 20, 1.0, 50, 64)
    A, B = hm.sorted_coords(0), hm.sorted_coords(1)
    from scipy.special import gamma, yn
    for kind, ls in [(O.K_GAUSSIAN_REF, 1.0), (O.K_GAUSSIAN, 0.1), (O.K_EXPONENTIAL, 0.1),
                     (O.K_MATERN_REF, 1.0)]:
        k = O.kernel(kind, 3, 0.25, ls)
        for i, j in [(0, 0), (3, 7), (49, 2), (10, 10)]:
            r = np.sqrt(np.sum((A[:, i] - B[:, j]) ** 2))
            if kind == O.K_GAUSSIAN_REF:
                v = np.exp(-r * r)
            elif kind == O.K_GAUSSIAN:
                v = np.exp(-r * r / (2 * ls * ls))
            elif kind == O.K_EXPONENTIAL:
                v = np.exp(-r / ls)
            else:
                v = yn(1, r + 1e-15) * r / (2.0 ** 1.5 * gamma(2.5))
            if i == j:
                v += 0.25
            assert hm.kernel_entry(k, i, j) == pytest.approx(v, rel=1e-13)


# ---------------------------------------------------------------- ACA
def test_aca_rank_one_is_exact():
    a = np.array([1.0, -2.0, 3.5, 0.25])
    b = np.array([0.5, 4.0, -1.0])
    A = np.outer(a, b)
    U, V, rp, cp, r = O.aca_matrix(A, 3)
    assert r == 1  # residual rows are zero, so the row search runs out of rows
    np.testing.assert_allclose(U @ V.T, A, rtol=0, atol=1e-15)
    assert rp[0] == 0 and cp[0] == 1  # |4.0| is the max of row 0


def test_aca_pivot_rules():
    # row 0 is zero and must be skipped; row 1 has a tie |3| at columns 0 and 2
    A = np.array([[0.0, 0.0, 0.0], [3.0, 1.0, -3.0], [1.0, 2.0, 0.5]])
    _, _, rp, cp, r = O.aca_matrix(A, 1, tie_last=True)
    assert rp[0] == 1 and cp[0] == 2  # batched path: last argmax
    _, _, rp, cp, r = O.aca_matrix(A, 1, tie_last=False)
    assert rp[0] == 1 and cp[0] == 0  # unbatched path: first argmax
    # scaled by the pivot read before scaling: V[cp] == 1
    U, V, rp, cp, r = O.aca_matrix(A, 1)
    assert V[cp[0], 0] == 1.0
    np.testing.assert_array_equal(U[:, 0], A[:, cp[0]])


def test_aca_full_rank_reproduces_matrix():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((9, 6))
    U, V, rp, cp, r = O.aca_matrix(A, 100)
    assert r == 6
    np.testing.assert_allclose(U @ V.T, A, atol=1e-12)
    assert list(rp) == list(range(6))  # sequential rows


def test_aca_rank_deficient_stops_when_rows_run_out():
    rng = np.random.default_rng(6)
    A = rng.standard_normal((5, 2)) @ rng.standard_normal((2, 7))
    U, V, rp, cp, r = O.aca_matrix(A, 5)
    assert r == 2
    np.testing.assert_allclose(U @ V.T, A, atol=1e-12)


def _smooth_block(m1=60, m2=50, ls=0.3):
    """Kernel block between two separated point clusters: numerically low rank."""
    rng = np.random.default_rng(11)
    P = rng.random((m1, 2))
    Q = rng.random((m2, 2)) + np.array([2.0, 0.5])
    return np.exp(-np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1)) / ls)


def _first_stop(U, V, c, every):
    """First rank step r (r % every == 0) where ||u_r|| ||v_r|| <= c ||U V^T||_F over
    the first r+1 terms, the norm taken of the explicit product."""
    for r in range(U.shape[1]):
        if r % every:
            continue
        fro = np.linalg.norm(U[:, :r + 1] @ V[:, :r + 1].T)
        if np.linalg.norm(U[:, r]) * np.linalg.norm(V[:, r]) <= c * fro:
            return r + 1
    return U.shape[1]


def test_aca_tolerance_stop_rank_matches_the_explicit_product():
    """The tolerance stop only truncates the fixed-rank run (pivots are unchanged),
    at the first checked step where the criterion holds for the explicit product
    U V^T; mode 1 is the reference's rule (every 5th step, eps (1-eta)/(1+eps)),
    mode 2 the plain relative rule (every step)."""
    A = _smooth_block()
    U, V, rp, cp, r = O.aca_matrix(A, 40)
    assert r >= 30  # the rows' residuals reach the 1e-13 floor near rank 40
    U, V = U[:, :r], V[:, :r]
    for eps in (1e-3, 1e-6, 1e-9):
        Ut, Vt, rpt, cpt, rt = O.aca_matrix(A, 40, tol_mode=2, eps=eps)
        assert rt == _first_stop(U, V, eps, 1) < r
        np.testing.assert_array_equal(Ut[:, :rt], U[:, :rt])
        np.testing.assert_array_equal(cpt[:rt], cp[:rt])
        # the classic ACA heuristic: the truncation error is of the order of eps
        assert np.linalg.norm(A - Ut[:, :rt] @ Vt[:, :rt].T) <= 10 * eps * np.linalg.norm(A)
    eta = 0.5
    for eps in (1e-4, 1e-8):
        _, _, _, _, rt = O.aca_matrix(A, 40, tol_mode=1, eps=eps, eta=eta)
        assert rt == _first_stop(U, V, eps * (1 - eta) / (1 + eps), 5)
        assert (rt - 1) % 5 == 0


def test_aca_tolerance_limits():
    """eps ~ 0 never stops early; the reference rule with eta >= 1 (every
    configuration here) never stops, since eps (1 - eta) <= 0; smaller eps never
    gives a smaller rank."""
    A = _smooth_block()
    _, _, _, _, r = O.aca_matrix(A, 30)
    assert O.aca_matrix(A, 30, tol_mode=2, eps=1e-300)[4] == r
    assert O.aca_matrix(A, 30, tol_mode=1, eps=1e-2, eta=1.0)[4] == r
    assert O.aca_matrix(A, 30, tol_mode=1, eps=1e-2, eta=1.5)[4] == r
    ranks = [O.aca_matrix(A, 30, tol_mode=2, eps=e)[4] for e in (1e-2, 1e-4, 1e-6, 1e-8, 1e-10)]
    assert ranks == sorted(ranks) and ranks[0] < ranks[-1]


def test_hmvp_full_rank_is_exact_against_numpy_dense():
    n = 600
    X = O.uniform(8, 2 * n).reshape(2, n)
    hm = O.HMatrix(X, X, 32, 1.5, 50, 16)
    k = O.kernel(O.K_GAUSSIAN, 2, 0.0, 0.2)
    hm.aca(k, 10000)  # k_b = min(m1, m2): every ACA block is reproduced exactly
    x = O.normal(9, n)
    Xs = hm.sorted_coords(0)
    D = np.exp(-((Xs[:, :, None] - Xs[:, None, :]) ** 2).sum(0) / (2 * 0.2 ** 2))
    order = hm.order(0).astype(np.int64)
    y_ref = np.empty(n)
    y_ref[order] = D @ x[order]
    np.testing.assert_allclose(hm.dense_mvp(k, x), y_ref, rtol=1e-12, atol=1e-12)
    y = hm.hmvp(k, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < 1e-10


def test_c1_configuration_error():
    """Config 1 (N=4096, U[0,1]^2, exp(-r^2/0.02), leaf 64, eta 1.5, k 32): the
    fixed-rank H-matrix error against the dense FP64 product."""
    n = 4096
    X = O.uniform(1001, 2 * n).reshape(2, n)
    hm = O.HMatrix(X, X, 32, 1.5, 50, 64)
    k = O.kernel(O.K_GAUSSIAN, 2, 0.0, 0.1)
    hm.aca(k, 32)
    x = O.uniform(1003, n) * 2 - 1
    err = np.linalg.norm(hm.hmvp(k, x) - hm.dense_mvp(k, x)) / np.linalg.norm(hm.dense_mvp(k, x))
    assert err < 1e-4
    rows = np.array([0, 17, 4095])
    np.testing.assert_allclose(hm.dense_rows(k, rows, x), hm.dense_mvp(k, x)[rows], rtol=1e-13)


def test_gaussian_rhs_is_the_integral():
    """b(x) = int_[0,1]^2 exp(-|x-y|^2) dy (hmglib.cu:403-418), by a tensor Gauss rule."""
    n = 30
    X = O.uniform(10, 2 * n).reshape(2, n)
    hm = O.HMatrix(X, X, 32, 1.5, 50, 64)
    b = hm.gaussian_rhs()
    g, w = np.polynomial.legendre.leggauss(40)
    g = 0.5 * (g + 1.0)
    w = 0.5 * w
    for i in range(n):
        f1 = (w * np.exp(-(X[0, i] - g) ** 2)).sum()
        f2 = (w * np.exp(-(X[1, i] - g) ** 2)).sum()
        assert b[i] == pytest.approx(f1 * f2, rel=1e-13)
