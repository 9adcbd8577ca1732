"""Pins of the CPU oracle (runs without a GPU).

The reference ships no tests or golden vectors, so every pin here comes from
mathematics or from an independent method: hand-derived Morton codes, numpy
sorting, an independent recursive trie for the Karras tree, a brute-force
Morton-cell census for the octree, two independent exact kNN methods, and
LAPACK (numpy.linalg.eigh) for the Jacobi eigensolver.  Each test is chosen so
that a plausible mistake (swapped axis, off-by-one, wrong root, dropped tie
rule, sign error) fails it.
"""
import numpy as np
import pytest

from conftest import rng


# ---------------------------------------------------------------- Morton
@pytest.mark.parametrize("bits", [30, 63])
def test_interleave_known_answers(O, bits):
    m = 2 ** (bits // 3) - 1
    assert O.interleave(1, 0, 0, bits) == 1
    assert O.interleave(0, 1, 0, bits) == 2
    assert O.interleave(0, 0, 1, bits) == 4
    assert O.interleave(2, 0, 0, bits) == 8
    assert O.interleave(0, 0, 2, bits) == 32
    assert O.interleave(m, m, m, bits) == 2**bits - 1
    assert O.interleave(m, 0, 0, bits) == int("001" * (bits // 3), 2)


def test_interleave_digit_order(O):
    # x=3 (011), y=5 (101), z=6 (110): digit i = (z_i y_i x_i)
    # i=0: z0=0 y0=1 x0=1 -> 011 ; i=1: z1=1 y1=0 x1=1 -> 101 ; i=2: z2=1 y2=1 x2=0 -> 110
    assert O.interleave(3, 5, 6, 63) == 0b110_101_011


@pytest.mark.parametrize("bits", [30, 63])
def test_quantize_known_answers(O, bits):
    s = 2 ** (bits // 3) - 1
    x = np.array([0.0, 1.0, 0.5, 0.25], np.float32)
    z = np.zeros(4, np.float32)
    codes = O.encode(x, z, z, 0.0, 1.0, bits)
    # x occupies bits 3i: codes are interleave(raw, 0, 0)
    raws = [0, s, int(np.floor(0.5 * s)), int(np.floor(0.25 * s))]
    assert [int(c) for c in codes] == [O.interleave(r, 0, 0, bits) for r in raws]
    # the max corner maps to all ones
    one = np.ones(1, np.float32)
    assert int(O.encode(one, one, one, 0.0, 1.0, bits)[0]) == 2**bits - 1


def test_zero_range_maps_to_zero(O):
    x = np.full(5, 3.0, np.float32)
    assert (O.encode(x, x, x, 3.0, 0.0, 63) == 0).all()


@pytest.mark.parametrize("bits", [30, 63])
def test_decode_encode_round_trip_on_grid(O, bits):
    # decode then encode is the identity on codes when the cube is [0, 1]: the
    # decoded coordinate raw/s re-quantises to raw (no intermediate rounding
    # can push s*(raw/s) below raw for s < 2^24).
    g = rng(7)
    s = 2 ** (bits // 3) - 1
    raw = g.integers(0, s + 1, size=(20000, 3), dtype=np.int64)
    codes = np.array([O.interleave(int(a), int(b), int(c), bits) for a, b, c in raw[:2000]], np.uint64)
    pts = O.decode(codes, 0.0, 1.0, bits)
    back = O.encode(pts[:, 0], pts[:, 1], pts[:, 2], 0.0, 1.0, bits)
    assert (back == codes).all()


def test_decode_is_monotone_per_axis(O):
    codes = np.array([O.interleave(r, 0, 0, 63) for r in range(0, 2**21, 997)], np.uint64)
    x = O.decode(codes, -3.25, 117.5, 63)[:, 0]
    assert (np.diff(x) >= 0).all()


def test_bbox(O):
    g = rng(1)
    x, y, z = (g.normal(size=1000).astype(np.float32) for _ in range(3))
    b, bad = O.bbox(x, y, z)
    allv = np.concatenate([x, y, z])
    assert bad == 0 and b[0] == allv.min() and b[1] == allv.max() and b[2] == np.float32(allv.max() - allv.min())
    x[5] = np.nan
    y[7] = np.inf
    b2, bad2 = O.bbox(x, y, z)
    assert bad2 == 2


# ---------------------------------------------------------------- sort / unique
def test_sort_pairs_matches_numpy_stable(O):
    g = rng(2)
    keys = g.integers(0, 2**20, size=50000, dtype=np.uint64)  # many duplicates
    ko, perm = O.sort_pairs(keys)
    ref = np.argsort(keys, kind="stable")
    assert (perm == ref).all() and (ko == keys[ref]).all()


def test_unique_representatives(O):
    g = rng(3)
    keys = g.integers(0, 5000, size=40000, dtype=np.uint64)
    ko, perm = O.sort_pairs(keys)
    uq, rep, uoi = O.unique(ko, perm)
    ref_u, ref_first, ref_inv = np.unique(keys, return_index=True, return_inverse=True)
    assert (uq == ref_u).all()
    assert (rep == ref_first).all()  # lowest input index of each code
    assert (uoi == ref_inv).all()


# ---------------------------------------------------------------- radix tree
def _sorted_unique_codes(O, x, y, z, bits):
    b, _ = O.bbox(x, y, z)
    c = O.encode(x, y, z, b[0], b[2], bits)
    s, p = O.sort_pairs(c)
    uq, _, _ = O.unique(s, p)
    return uq, b


def _tree_equal(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


@pytest.mark.parametrize("seed,n,bits", [(11, 17, 63), (12, 300, 30), (13, 2000, 63), (14, 5000, 30)])
def test_karras_textbook_equals_trie(O, seed, n, bits):
    g = rng(seed)
    x, y, z = (g.random(n).astype(np.float32) for _ in range(3))
    uq, _ = _sorted_unique_codes(O, x, y, z, bits)
    U = len(uq)
    kt = O.karras(uq, bits, 0)
    tr = O.trie(uq, bits)
    for key in kt:
        assert np.array_equal(kt[key][: U - 1], tr[key][: U - 1]), key
    assert np.array_equal(kt["parent"], tr["parent"])


def test_karras_sphere100k_equals_trie_and_reference_root(O, sx):
    x, y, z = sx.gen_sphere(1, 100_000, 0.002)
    uq, _ = _sorted_unique_codes(O, x, y, z, 30)
    U = len(uq)
    kt, kr, tr = O.karras(uq, 30, 0), O.karras(uq, 30, 1), O.trie(uq, 30)
    for key in kt:
        assert np.array_equal(kt[key][: U - 1], tr[key][: U - 1]), key
        assert np.array_equal(kt[key][: U - 1], kr[key][: U - 1]), key  # benchmark input: roots agree


def test_reference_root_differs_in_degenerate_case(O):
    # the last code alone forms the root's right subtree: the reference root
    # (l = N-1) leaves it out, the textbook root spans all leaves.
    codes = np.array([0b000, 0b001, 0b010, 0b011, 0b100 << 60], np.uint64)
    kt, kr = O.karras(codes, 63, 0), O.karras(codes, 63, 1)
    assert kt["has_leaf_right"][0] == 1 and kt["left_child"][0] == 3
    assert not _tree_equal({k: v[:4] for k, v in kt.items()}, {k: v[:4] for k, v in kr.items()})


def test_reference_root_hand_derived(O):
    """constructTree (src/RadixTree.cu:99-194) stepped by hand with the root's
    range end l = N-1, N = n_unique - 1 internal nodes (:128-131)."""
    # the root's right subtree is exactly the last two leaves: the reference
    # root spans [0, 2], so its right child is leaf 2 and leaf 3's node 2 has
    # no parent; the textbook root spans [0, 3]
    c = np.array([1, 2, 1 << 62, (1 << 62) + 1], np.uint64)
    ref, txt = O.karras(c, 63, 1), O.karras(c, 63, 0)
    assert list(ref["prefix_n"][:3]) == [0, 61, 62] and list(txt["prefix_n"][:3]) == [0, 61, 62]
    assert list(ref["left_child"][:3]) == [1, 0, 2] and list(txt["left_child"][:3]) == [1, 0, 2]
    assert list(ref["has_leaf_left"][:3]) == [0, 1, 1] and list(txt["has_leaf_left"][:3]) == [0, 1, 1]
    assert list(ref["has_leaf_right"][:3]) == [1, 1, 1]
    assert list(txt["has_leaf_right"][:3]) == [0, 1, 1]
    assert list(ref["parent"][:3]) == [-1, 0, -1] and list(txt["parent"][:3]) == [-1, 0, 0]
    # the last code alone is the right subtree: the reference root covers
    # [0, 2] with delta 61, splits after leaf 0, and node 2 (range [0, 2]
    # seen from its other end) claims node 1 as well
    c = np.array([1, 2, 3, 1 << 62], np.uint64)
    ref, txt = O.karras(c, 63, 1), O.karras(c, 63, 0)
    assert list(ref["prefix_n"][:3]) == [61, 62, 61]
    assert list(ref["left_child"][:3]) == [0, 1, 0]
    assert list(ref["has_leaf_left"][:3]) == [1, 1, 1]
    assert list(ref["has_leaf_right"][:3]) == [0, 1, 0]
    assert ref["parent"][0] == -1 and ref["parent"][1] in (0, 2)
    assert list(txt["prefix_n"][:3]) == [0, 62, 61]
    assert list(txt["left_child"][:3]) == [2, 1, 0]
    assert list(txt["has_leaf_left"][:3]) == [0, 1, 1] and list(txt["has_leaf_right"][:3]) == [1, 1, 0]
    assert list(txt["parent"][:3]) == [-1, 2, 0]


def test_radix_tree_invariants(O, sx):
    x, y, z = sx.gen_scene(21, 20000)
    uq, _ = _sorted_unique_codes(O, x, y, z, 63)
    U = len(uq)
    t = O.karras(uq, 63, 0)
    lc, pn, par = t["left_child"], t["prefix_n"].astype(int), t["parent"]
    # children share the node's prefix; the split bit separates them; parent round trip
    for i in range(U - 1):
        g = lc[i]
        p = pn[i]
        shift = 63 - p
        assert (int(uq[g]) >> shift) == (int(uq[g + 1]) >> shift)
        assert ((int(uq[g]) >> (shift - 1)) & 1) == 0 and ((int(uq[g + 1]) >> (shift - 1)) & 1) == 1
        if not t["has_leaf_left"][i]:
            assert par[g] == i and pn[g] > p
        if not t["has_leaf_right"][i]:
            assert par[g + 1] == i and pn[g + 1] > p


# ---------------------------------------------------------------- octree
def _octree_census_check(O, uq, bits, b, rt, nodes):
    pn = rt["prefix_n"]
    root_level = int(pn[0]) // 3
    cells = O.octree_cells(uq, bits, root_level)
    assert len(nodes) == len(cells)
    census = {(int(L), int(p)) for L, p, _, _ in cells}
    # walk from the root and check every slot against the census
    seen = set()
    stack = [(0, root_level, int(uq[0]) >> (bits - 3 * root_level) if root_level else 0)]
    while stack:
        idx, L, p = stack.pop()
        assert (L, p) in census
        assert (L, p) not in seen
        seen.add((L, p))
        nd = nodes[idx]
        assert nd["cell_size"] == np.float32(np.float32(b[2]) / np.float32(2 ** (L - root_level))) or idx == 0
        code0 = (p << (bits - 3 * L)) if L else 0
        corner = O.decode(np.array([code0], np.uint64), b[0], b[2], bits)[0]
        assert np.array_equal(nd["corner"], corner)
        for s in range(8):
            child = (L + 1, p * 8 + s)
            is_node = bool(nd["child_node_mask"] >> s & 1)
            is_leaf = bool(nd["child_leaf_mask"] >> s & 1)
            assert not (is_node and is_leaf)
            if is_node:
                stack.append((int(nd["children"][s]), L + 1, p * 8 + s))
            elif is_leaf:
                leaf = int(nd["children"][s])
                assert (int(uq[leaf]) >> (bits - 3 * (L + 1))) == p * 8 + s
                assert child not in census
            else:
                assert child not in census
    assert seen == census
    assert nodes[0]["cell_size"] == np.float32(b[2])


@pytest.mark.parametrize("kind,seed,n,bits", [("sphere", 1, 100_000, 30), ("scene", 31, 50_000, 63),
                                               ("degenerate", 32, 30_000, 63)])
def test_octree_matches_cell_census(O, sx, kind, seed, n, bits):
    if kind == "sphere":
        x, y, z = sx.gen_sphere(seed, n, 0.002)
    elif kind == "scene":
        x, y, z = sx.gen_scene(seed, n)
    else:
        x, y, z = sx.gen_degenerate(seed, n)[:3]
    uq, b = _sorted_unique_codes(O, x, y, z, bits)
    rt = O.karras(uq, bits, 0)
    nodes = O.octree(uq, bits, b[0], b[2], rt)
    _octree_census_check(O, uq, bits, b, rt, nodes)


# ---------------------------------------------------------------- kNN
def test_knn_brute_equals_grid_sphere(O, sx):
    x, y, z = sx.gen_sphere(1, 100_000, 0.002)
    uq, b = _sorted_unique_codes(O, x, y, z, 30)
    pts = O.decode(uq, b[0], b[2], 30)
    q = np.arange(0, len(pts), 7)
    g = O.knn_grid(pts, 8, q)
    br = np.concatenate([O.knn_brute(pts, 8, int(s), 1) for s in q[:300]])
    assert np.array_equal(g[:300], br)
    # rows: self first, ascending (d2, index)
    d2 = np.sum((pts[g] - pts[q][:, None, :]) ** 2, axis=2)
    assert (g[:, 0] == q).all()
    assert (np.diff(d2, axis=1) >= -1e-12).all()


def test_knn_ties_broken_by_index(O):
    # lattice: many exactly equal distances
    a = np.arange(6, dtype=np.float32)
    P = np.stack(np.meshgrid(a, a, a, indexing="ij"), -1).reshape(-1, 3)
    k = 8
    res = O.knn_brute(P, k)
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    for q in range(len(P)):
        order = np.lexsort((np.arange(len(P)), d2[q]))[:k]
        assert np.array_equal(res[q], order)
    assert np.array_equal(O.knn_grid(P, k, np.arange(len(P))), res)


def test_knn_pads_when_fewer_points_than_k(O):
    P = np.array([[0, 0, 0], [1, 0, 0], [3, 0, 0]], np.float32)
    res = O.knn_brute(P, 8)
    assert list(res[1]) == [1, 0, 2, 1, 1, 1, 1, 1]
    assert np.array_equal(O.knn_grid(P, 8, np.arange(3)), res)


def test_reference_knn_emulation_recall(O, sx):
    x, y, z = sx.gen_scene(41, 30000)
    uq, b = _sorted_unique_codes(O, x, y, z, 63)
    rt = O.karras(uq, 63, 1)
    nodes = O.octree(uq, 63, b[0], b[2], rt)
    pts = O.decode(uq, b[0], b[2], 63)
    ref = O.knn_reference(nodes, pts, 16, 0, 3000)
    ex = O.knn_brute(pts, 16, 0, 3000)
    recall = np.mean([len(set(r) & set(e)) / 16 for r, e in zip(ref, ex)])
    assert recall > 0.5  # diagnostic: the reference search is approximate


def _otnodes(O, n):
    nodes = np.zeros(n, O.OTNODE_DTYPE)
    nodes["cell_size"] = 1.0
    nodes["corner"] = -100.0
    return nodes


def _three_squares(lo, count, skip=()):
    """Integer points (a, b, c) with distinct squared norms >= lo, ascending."""
    out, t = [], lo
    while len(out) < count:
        hit = next(((a, b, c) for a in range(13) for b in range(a + 1) for c in range(b + 1)
                    if a * a + b * b + c * c == t), None)
        if hit is not None and t not in skip:
            out.append(hit)
        t += 1
    return out


def test_reference_result_queue_hand_derived(O):
    """The k-slot result queue of knnSearchKernel (src/Octree.hpp:136-193):
    when full, an item replaces the maximum if that maximum's priority is >=
    its own, else it is dropped (src/dequeue.hpp:60-68); an equal priority
    replaces.  Root (leaves 0-6, d2 0 1 2 4 5 6 9, then node 1) and node 1
    (leaves 7-11, d2 10 3 9 11 12) visited in that order; hand-derived queue:
    {0..6} + 7 -> full; 8 (3) evicts 7 (10); 9 (9) evicts 6 (9, equal);
    10 and 11 dropped.  Rows ascending: 0 1 2 8 3 4 5 9 - the exact search
    (ties by index) ends with 6 instead."""
    pts = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 0, 0), (2, 1, 0), (2, 1, 1), (3, 0, 0),
                    (3, 1, 0), (1, 1, 1), (2, 2, 1), (3, 1, 1), (2, 2, 2)], np.float32)
    nodes = _otnodes(O, 2)
    nodes["children"][0, :7] = np.arange(7)
    nodes["children"][0, 7] = 1
    nodes["child_leaf_mask"][0] = 0x7F
    nodes["child_node_mask"][0] = 0x80
    nodes["children"][1, :5] = np.arange(7, 12)
    nodes["child_leaf_mask"][1] = 0x1F
    nodes["corner"][1] = (0.5, -100.0, -100.0)  # box distance 0.5
    got = O.knn_reference(nodes, pts, 8, 0, 1)[0]
    assert list(got) == [0, 1, 2, 8, 3, 4, 5, 9]
    assert list(O.knn_brute(pts, 8, 0, 1)[0]) == [0, 1, 2, 8, 3, 4, 5, 6]


def test_reference_search_queue_overflow_hand_derived(O):
    """The 32-slot search queue (SEARCH_Q_SIZE, src/Octree.hpp:17) drops
    nodes when full.  Root: child nodes C1-C7 (box distance 1 + i/64) and
    leaf 0 (the query, at the origin); C_i: grandchildren g = 8(i-1) + s (box
    distance 2 + g/64), each holding one point 1 + g.  Best first, C1-C3
    queue 24 grandchildren (28 items with C4-C7); C4 adds g24-g28, then the
    queue is full and g29-g31 exceed its maximum (g28): dropped; C5, C6, C7
    each free one slot for g32, g40, g48 and drop the rest.  The points of
    g29 (d2 1) and g30 (d2 3) are never seen; g48's (d2 2) evicts the
    largest result.  Rows: 0, 49 (g48), 1 (g0, d2 50), 2-6 (g1-g5)."""
    n_gc = 56
    nodes = _otnodes(O, 8 + n_gc)
    nodes["children"][0, :7] = np.arange(1, 8)
    nodes["children"][0, 7] = 0
    nodes["child_node_mask"][0] = 0x7F
    nodes["child_leaf_mask"][0] = 0x80
    for i in range(1, 8):
        nodes["corner"][i] = (1.0 + i / 64.0, -100.0, -100.0)
        nodes["children"][i] = 8 + 8 * (i - 1) + np.arange(8)
        nodes["child_node_mask"][i] = 0xFF
    for g in range(n_gc):
        nodes["corner"][8 + g] = (2.0 + g / 64.0, -100.0, -100.0)
        nodes["children"][8 + g, 0] = 1 + g
        nodes["child_leaf_mask"][8 + g] = 1
    special = {29: (1, 0, 0), 30: (1, 1, 1), 48: (1, 1, 0), 0: (5, 4, 3)}  # d2 1, 3, 2, 50
    far = iter(_three_squares(100, n_gc))
    pts = np.zeros((1 + n_gc, 3), np.float32)
    for g in range(n_gc):
        pts[1 + g] = special.get(g) or next(far)
    got = O.knn_reference(nodes, pts, 8, 0, 1)[0]
    assert list(got) == [0, 49, 1, 2, 3, 4, 5, 6]
    assert list(O.knn_brute(pts, 8, 0, 1)[0]) == [0, 30, 49, 31, 1, 2, 3, 4]


# ---------------------------------------------------------------- eigen / PCA
def test_jacobi_matches_lapack(O):
    g = rng(5)
    for _ in range(500):
        M = g.normal(size=(3, 3))
        A = M @ M.T * 10 ** g.uniform(-6, 6)
        if g.random() < 0.2:  # near-repeated eigenvalues
            Q, _ = np.linalg.qr(g.normal(size=(3, 3)))
            A = Q @ np.diag([1.0, 1.0 + 1e-9, 5.0]) @ Q.T
        w, V = O.jacobi3(A)
        wr, Vr = np.linalg.eigh(A)
        assert np.allclose(w, wr, rtol=1e-10, atol=1e-12 * abs(wr).max())
        assert np.allclose(A @ V, V * w, atol=1e-10 * abs(wr).max())
        gap = wr[1] - wr[0]
        if gap > 1e-6 * abs(wr).max():
            assert abs(abs(V[:, 0] @ Vr[:, 0]) - 1) < 1e-9


def test_pca_plane_normal_closed_form(O):
    g = rng(6)
    u, v = g.random(64), g.random(64)
    n = np.array([1.0, 2.0, 2.0]) / 3.0
    e1 = np.array([2.0, -1.0, 0]) / np.sqrt(5)
    e2 = np.cross(n, e1)
    P = (u[:, None] * e1 + v[:, None] * e2 + 5.0).astype(np.float32)
    nn = np.arange(16, dtype=np.int32)[None, :].repeat(4, 0) + np.arange(4, dtype=np.int32)[:, None] * 16
    normals, evals = O.pca_f64(P, nn)
    assert (O.angular_error_deg(normals, np.repeat(n[None], 4, 0)) < 1e-3).all()
    assert (evals[:, 0] < 1e-12 * evals[:, 2] + 1e-12).all()


def test_reference_f32_normals_close_to_f64(O, sx):
    x, y, z = sx.gen_sphere(1, 100_000, 0.002)
    uq, b = _sorted_unique_codes(O, x, y, z, 30)
    pts = O.decode(uq, b[0], b[2], 30)
    nn = O.knn_brute(pts, 8, 0, 5000)
    ref32 = O.normals_ref_f32(pts, nn)
    n64, _ = O.pca_f64(pts, nn)
    err = O.angular_error_deg(ref32, n64)
    assert np.median(err) < 0.5
    # orientation rule: (0 - p) . n >= 0
    assert (np.sum(-pts[:5000] * ref32, axis=1) >= 0).all()


# ---------------------------------------------------------------- MLP oracle
def test_mlp_oracle_gradient_matches_finite_differences(O):
    g = rng(8)
    P = (g.normal(size=4801) * 0.2).astype(np.float32)
    f = g.normal(size=(40, 8)).astype(np.float32)
    t = g.normal(size=40).astype(np.float32)
    mu, sg = np.zeros(8, np.float32), np.ones(8, np.float32)
    grad, _ = O.mlp_grad(P, f, t, mu, sg, emulate=False)

    def loss(PP):
        return np.mean((O.mlp_forward(PP, f, mu, sg, emulate=False) - t) ** 2)

    # one index per parameter block: W1, b1, W2, b2, w3, b3
    for j in (3, 517, 1000, 4200, 4650, 4750, 4800):
        e = 3e-5
        a, b = P.copy(), P.copy()
        a[j] += e
        b[j] -= e
        num = (loss(a) - loss(b)) / (np.float64(a[j]) - np.float64(b[j]))
        assert abs(num - grad[j]) <= 1e-3 * abs(grad[j]) + 1e-7, (j, num, grad[j])


def test_mlp_oracle_bf16_emulation_is_close(O):
    g = rng(9)
    P = (g.normal(size=4801) * 0.2).astype(np.float32)
    f = g.normal(size=(200, 8)).astype(np.float32)
    mu, sg = np.zeros(8, np.float32), np.ones(8, np.float32)
    y64 = O.mlp_forward(P, f, mu, sg, emulate=False)
    y16 = O.mlp_forward(P, f, mu, sg, emulate=True)
    assert np.max(np.abs(y64 - y16)) < 0.05 * (1 + np.abs(y64).max())
    assert not np.array_equal(y64, y16)
