"""GPU parity of every stage against the CPU oracle, through the C ABI.

Bit-exact: cube bounds, Morton codes, sort permutation, unique codes and
representatives, radix tree arrays, octree nodes (byte for byte), decoded
points, kNN index rows.  Normals: angular error against FP64 PCA on the same
neighbour rows, median <= 0.1 deg and p99 <= 1 deg (the metric's precision gate),
plus the orientation rule.
"""
import numpy as np
import pytest

from conftest import ROOT, rng

pytestmark = pytest.mark.gpu


def _stages(sx, x, y, z, bits, root_mode=0):
    D = sx.DeviceArray
    dx, dy, dz = D.from_numpy(x), D.from_numpy(y), D.from_numpy(z)
    n = len(x)
    bounds, bad = sx.bbox(dx, dy, dz)
    codes = sx.morton_encode(dx, dy, dz, bounds, bits)
    ko, perm = sx.sort_pairs(codes, bits)
    uq, rep, uos, cnt = sx.unique(ko, perm)
    U, _ = sx.check_counts(cnt, n, bad=bad)
    rt = sx.radix_tree(uq, cnt, bits, root_mode)
    nodes, ocnt = sx.octree(uq, cnt, bits, bounds, rt)
    _, n_oct = sx.check_counts(cnt, n, ocnt, len(nodes))
    pts = sx.morton_decode(uq, bounds, bits, cnt)
    out = {
        "bounds": bounds.numpy()[:3],
        "codes": codes.numpy(),
        "sorted": ko.numpy(),
        "perm": perm.numpy(),
        "unique": uq.numpy()[:U],
        "rep": rep.numpy()[:U],
        "unique_of_sorted": uos.numpy(),
        "points": pts.numpy()[:U],
        "oct_nodes": nodes.numpy()[:n_oct],
    }
    for k, v in rt.items():
        out[k] = v.numpy()[:U]
    return out


def _oracle(O, x, y, z, bits, root_mode=0):
    b, bad = O.bbox(x, y, z)
    codes = O.encode(x, y, z, b[0], b[2], bits)
    sc, perm = O.sort_pairs(codes)
    uq, rep, uoi = O.unique(sc, perm)
    rt = O.karras(uq, bits, root_mode)
    nodes = O.octree(uq, bits, b[0], b[2], rt)
    pts = O.decode(uq, b[0], b[2], bits)
    return dict(bounds=b, codes=codes, sorted=sc, perm=perm, unique=uq, rep=rep, unique_of_input=uoi,
                points=pts, oct_nodes=nodes, **rt)


def _assert_stages_equal(g, o):
    U = len(o["unique"])
    assert np.array_equal(g["bounds"], o["bounds"])
    assert np.array_equal(g["codes"], o["codes"])
    assert np.array_equal(g["sorted"], o["sorted"])
    assert np.array_equal(g["perm"], o["perm"])
    assert np.array_equal(g["unique"], o["unique"])
    assert np.array_equal(g["rep"], o["rep"])
    assert np.array_equal(g["points"], o["points"])
    for k in ("prefix_n", "left_child", "has_leaf_left", "has_leaf_right"):
        assert np.array_equal(g[k][: U - 1], o[k][: U - 1]), k
    assert np.array_equal(g["parent"][:U], o["parent"][:U])
    assert len(g["oct_nodes"]) == len(o["oct_nodes"])
    assert g["oct_nodes"].tobytes() == o["oct_nodes"].tobytes()


def _workload(sx, kind, seed, n):
    if kind == "sphere":
        return sx.gen_sphere(seed, n, 0.002)
    if kind == "scene":
        return sx.gen_scene(seed, n)
    if kind == "gmm":
        return sx.gen_gmm_scene(seed, n)
    return sx.gen_degenerate(seed, n)[:3]


def test_stages_match_golden_fixture(gpu):
    f = np.load(ROOT / "tests" / "golden" / "small_scene.npz")
    g = _stages(gpu, f["x"], f["y"], f["z"], 63)
    assert np.array_equal(g["bounds"], f["bounds"])
    assert np.array_equal(g["codes"], f["codes"])
    assert np.array_equal(g["perm"], f["perm"])
    assert np.array_equal(g["unique"], f["unique"])
    assert np.array_equal(g["rep"], f["rep"])
    assert np.array_equal(g["points"], f["points"])
    U = len(f["unique"])
    for k in ("prefix_n", "left_child", "has_leaf_left", "has_leaf_right"):
        assert np.array_equal(g[k][: U - 1], f[k][: U - 1]), k
    assert np.array_equal(g["parent"], f["parent"])
    assert g["oct_nodes"].view(np.uint8).reshape(-1, 56).tobytes() == f["oct_nodes"].tobytes()


@pytest.mark.parametrize("kind,seed,n,bits", [
    ("sphere", 1, 100_000, 30),
    ("scene", 2, 1_000_000, 63),
    ("scene", 9, 777_777, 30),
    ("gmm", 3, 500_000, 63),
    ("degenerate", 5, 300_000, 63),
])
def test_stages_bit_exact(gpu, O, kind, seed, n, bits):
    x, y, z = _workload(gpu, kind, seed, n)
    _assert_stages_equal(_stages(gpu, x, y, z, bits), _oracle(O, x, y, z, bits))


def test_reference_root_mode_bit_exact(gpu, O):
    x, y, z = gpu.gen_sphere(1, 100_000, 0.002)
    g = _stages(gpu, x, y, z, 30, root_mode=1)
    o = _oracle(O, x, y, z, 30, root_mode=1)
    _assert_stages_equal(g, o)


def test_reference_root_degenerate_codes(gpu, O):
    codes = np.array([0b000, 0b001, 0b010, 0b011, 0b100 << 60], np.uint64)
    dc = gpu.DeviceArray.from_numpy(codes)
    cnt = gpu.DeviceArray.from_numpy(np.array([5], np.int32))
    for mode in (0, 1):
        rt = gpu.radix_tree(dc, cnt, 63, mode)
        o = O.karras(codes, 63, mode)
        for k in ("prefix_n", "left_child", "has_leaf_left", "has_leaf_right"):
            assert np.array_equal(rt[k].numpy()[:4], o[k][:4]), (mode, k)
        # Under the reference root this input makes two nodes claim the same
        # children (the reference's constructTree races on parent[] there), so
        # only children claimed once have a defined parent.
        claims = np.zeros(len(codes), int)
        for i in range(len(codes) - 1):
            g = o["left_child"][i]
            claims[g] += 0 if o["has_leaf_left"][i] else 1
            claims[g + 1] += 0 if o["has_leaf_right"][i] else 1
        once = claims <= 1
        assert np.array_equal(rt["parent"].numpy()[once], o["parent"][once]), mode
        if mode == 0:
            assert once.all()


def test_sort_pairs_arbitrary_keys(gpu):
    g = rng(3)
    for n, bits in ((1, 64), (2047, 64), (2049, 17), (300_001, 40), (1 << 20, 64)):
        keys = g.integers(0, 2**min(bits, 63), size=n, dtype=np.uint64)
        keys[: n // 3] = keys[n // 2]  # duplicates exercise stability
        ko, vo = gpu.sort_pairs(gpu.DeviceArray.from_numpy(keys), bits)
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(vo.numpy(), ref) and np.array_equal(ko.numpy(), keys[ref])


@pytest.mark.parametrize("bits", [63, 64, 52])
def test_sort_pairs_bucket_edges(gpu, bits):
    """Keys whose top bits (the four digits the sort places first) form
    buckets of every size class and at every class edge: ranked in the local
    pass (up to 64 keys), sorted by one wave (65-1024), sorted by the full
    passes (larger); buckets across 2048-key tile edges, large ones next to
    small ones, duplicate keys inside buckets of every class; explicit
    values; a shuffled input."""
    g = rng(11 + bits)
    low_bits = 8 * ((bits + 7) // 8 - 4)  # bits below the four top digits
    sizes = [1, 2, 3, 63, 64, 65, 255, 256, 257, 511, 1023, 1024, 1025, 2047, 2048, 2049, 5000, 70_000, 17, 64, 1]
    sizes += list(g.integers(1, 300, size=3000)) + list(g.integers(900, 1200, size=40))
    top = np.repeat(np.arange(len(sizes), dtype=np.uint64) * np.uint64(7919) + np.uint64(3), sizes)
    top_bits = bits - low_bits
    top = top % np.uint64(1 << top_bits)
    if bits == 64:
        top |= np.uint64(1 << (top_bits - 1))  # keys with bit 63 set
    low = g.integers(0, 1 << low_bits, size=len(top), dtype=np.uint64)
    low[g.random(len(low)) < 0.3] = np.uint64(5)  # duplicate keys within buckets
    keys = (top << np.uint64(low_bits)) | low
    perm = g.permutation(len(keys))
    keys = keys[perm]
    vals = g.integers(0, 2**32, size=len(keys), dtype=np.uint64).astype(np.uint32)
    ko, vo = gpu.sort_pairs(gpu.DeviceArray.from_numpy(keys), bits, gpu.DeviceArray.from_numpy(vals), check=True)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(ko.numpy(), keys[ref])
    assert np.array_equal(vo.numpy(), vals[ref])


@pytest.mark.parametrize("n", [2048, 2049, 2111, 2177, 4097, 6000])
def test_sort_pairs_hybrid_small_inputs(gpu, n):
    """The hybrid schedule from its smallest input (one 2,048-key local tile)
    up: clustered 63-bit keys (buckets of every class inside a few tiles),
    duplicates, and the last tile partly filled."""
    g = rng(n)
    top = g.integers(0, 40, size=n, dtype=np.uint64) * np.uint64(7)  # few buckets: large ones
    low = g.integers(0, 1 << 32, size=n, dtype=np.uint64)
    low[g.random(n) < 0.25] = np.uint64(9)
    keys = (top << np.uint64(32)) | low
    keys[: n // 8] = g.integers(0, 2**63, size=n // 8, dtype=np.uint64)  # and some singletons
    ko, vo = gpu.sort_pairs(gpu.DeviceArray.from_numpy(keys), 63, check=True)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(vo.numpy(), ref) and np.array_equal(ko.numpy(), keys[ref])


def test_index_object_matches_oracle(gpu, O):
    x, y, z = gpu.gen_scene(2, 1_000_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    f = ix.fetch()
    o = _oracle(O, x, y, z, 63)
    U = f["n_unique"]
    assert U == len(o["unique"]) and f["n_oct"] == len(o["oct_nodes"])
    assert f["min_coord"] == o["bounds"][0] and f["max_coord"] == o["bounds"][1]
    assert np.array_equal(f["codes"], o["unique"])
    assert np.array_equal(f["sorted_perm"], o["perm"])
    assert np.array_equal(f["rep"], o["rep"])
    assert np.array_equal(f["unique_of_input"], o["unique_of_input"])
    assert np.array_equal(f["points"], o["points"])
    assert f["oct_nodes"].tobytes() == o["oct_nodes"].tobytes()
    for k in ("prefix_n", "left_child", "has_leaf_left", "has_leaf_right"):
        assert np.array_equal(f[k][: U - 1], o[k][: U - 1]), k


# ---------------------------------------------------------------- edge cases
@pytest.mark.parametrize("pts", [
    [[0.5, 0.5, 0.5]],
    [[0, 0, 0], [1, 1, 1]],
    [[0, 0, 0], [1, 0, 0], [0, 1, 0]],
    [[2, 2, 2]] * 1000,
    [[0, 0, 0]] * 10 + [[1, 2, 3]] * 7,
])
def test_degenerate_inputs(gpu, O, pts):
    P = np.array(pts, np.float32)
    x, y, z = P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy()
    ix = gpu.Index(len(P), 63).build(x, y, z)
    f = ix.fetch()
    o = _oracle(O, x, y, z, 63)
    U = f["n_unique"]
    assert U == len(o["unique"])
    assert np.array_equal(f["codes"], o["unique"]) and np.array_equal(f["rep"], o["rep"])
    assert f["n_oct"] == len(o["oct_nodes"]) and f["oct_nodes"].tobytes() == o["oct_nodes"].tobytes()
    for k in (8, 16):
        nn = ix.knn(k).numpy()
        assert np.array_equal(nn, O.knn_brute(o["points"], k))
        nrm = ix.normals(ix.knn(k), k).numpy()
        assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-5)


def test_non_finite_input_reported(gpu):
    x, y, z = gpu.gen_scene(2, 10_000)
    x[17] = np.nan
    ix = gpu.Index(len(x), 63).build(x, y, z)
    with pytest.raises(gpu.SxError) as e:
        ix.view()
    assert e.value.code == 5  # SX_EDATA


def test_octree_capacity_overflow_reported(gpu):
    x, y, z = gpu.gen_scene(2, 50_000)
    ix = gpu.Index(len(x), 63, oct_factor=0.05).build(x, y, z)
    with pytest.raises(gpu.SxError) as e:
        ix.view()
    assert e.value.code == 4  # SX_ERANGE


def test_query_range_validated(gpu):
    x, y, z = gpu.gen_sphere(1, 1000)
    ix = gpu.Index(1000, 30).build(x, y, z)
    U = ix.n_unique
    with pytest.raises(gpu.SxError):
        ix.knn(8, q_start=U - 5, n_q=10)
    with pytest.raises(gpu.SxError):
        ix.knn(12)


# ---------------------------------------------------------------- kNN
@pytest.mark.parametrize("k", [8, 16, 24, 32])
def test_knn_sphere_all_queries_vs_brute(gpu, O, k):
    x, y, z = gpu.gen_sphere(1, 100_000, 0.002)
    ix = gpu.Index(len(x), 30).build(x, y, z)
    f = ix.fetch()
    got = ix.knn(k).numpy()
    assert np.array_equal(got, O.knn_brute(f["points"], k))


@pytest.mark.parametrize("kind,seed,n,k", [("scene", 2, 1_000_000, 16), ("gmm", 3, 1_000_000, 32),
                                            ("degenerate", 5, 500_000, 24)])
def test_knn_sampled_vs_grid(gpu, O, kind, seed, n, k):
    x, y, z = _workload(gpu, kind, seed, n)
    ix = gpu.Index(n, 63).build(x, y, z)
    f = ix.fetch()
    q = np.sort(rng(seed).choice(f["n_unique"], 20_000, replace=False))
    got = ix.knn(k).numpy()[q]
    assert np.array_equal(got, O.knn_grid(f["points"], k, q))


def test_knn_slice_and_bucket_invariance(gpu):
    x, y, z = gpu.gen_scene(4, 300_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    full = ix.knn(16).numpy()
    U = ix.n_unique
    part = ix.knn(16, q_start=1234, n_q=U // 3).numpy()
    assert np.array_equal(part, full[1234:1234 + U // 3])
    for b in (1, 2, 4, 32, 64, 256):
        ix.set_bucket(b)
        assert np.array_equal(ix.knn(16).numpy(), full), b
    ix.set_bucket(16)


def test_knn_radius_seed_is_exact(gpu):
    x, y, z = gpu.gen_scene(6, 400_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    k = 16
    plain = ix.knn(k).numpy()
    pts = ix.fetch()["points"]
    kth = np.sum((pts[plain[:, -1]] - pts) ** 2, axis=1).astype(np.float32)
    g = rng(6)
    for scale in (0.25, 1.0, 1.0001, 4.0, None):
        r2 = kth * np.float32(scale) if scale is not None else (kth * g.uniform(0, 3, len(kth))).astype(np.float32)
        ix.knn_reruns(reset=True)
        got = ix.knn(k, mode=gpu.SX_KNN_RADIUS, radius2=gpu.DeviceArray.from_numpy(r2.astype(np.float32))).numpy()
        assert np.array_equal(got, plain), scale
        reruns = ix.knn_reruns()
        print(f"[radius] scale {scale}: {reruns / len(kth):.3%} searched again")
        if scale == 0.25:
            assert reruns > 0.9 * len(kth)
        # a radius at (or just above) the k-th distance proves almost every
        # tile lane exact; a 4x radius only widens the tile boxes (more LDS
        # overflows, still exact results)
        if scale in (1.0, 1.0001):
            assert reruns < 0.05 * len(kth)


def test_knn_equidistant_shell_exact(gpu, O):
    """A query at the centre of a shell has every subtree within its k-th
    distance, so its deferred search overflows the 16-lane group's stack and
    is finished by the one-lane fallback; every row must still be exact."""
    g = rng(11)
    v = g.normal(size=(30_000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    p = np.vstack([v, np.zeros((1, 3))]).astype(np.float32)
    x, y, z = (np.ascontiguousarray(p[:, i]) for i in range(3))
    ix = gpu.Index(len(x), 63).build(x, y, z)
    ix.set_bucket(1)
    ref = O.knn_brute(ix.fetch()["points"], 16)
    assert np.array_equal(ix.knn(16).numpy(), ref)
    # a tiny radius defers every query to the group kernel, with the bound of
    # its 17-point seed list
    r2 = gpu.DeviceArray.from_numpy(np.full(ix.n_unique, 1e-12, np.float32))
    ix.knn_reruns(reset=True)
    got = ix.knn(16, mode=gpu.SX_KNN_RADIUS, radius2=r2).numpy()
    assert ix.knn_reruns() > 0.9 * ix.n_unique
    assert np.array_equal(got, ref)
    ix.set_bucket(16)


@pytest.mark.parametrize("kind,seed,n,k", [("scene", 2, 1_000_000, 16), ("gmm", 3, 600_000, 32),
                                            ("degenerate", 5, 400_000, 24), ("sphere", 1, 100_000, 8)])
def test_knn_algorithms_agree(gpu, kind, seed, n, k):
    """Tiles (+ queue) and one-lane-per-query give identical rows, plain and seeded."""
    x, y, z = _workload(gpu, kind, seed, n)
    ix = gpu.Index(n, 63).build(x, y, z)
    tiles = ix.knn(k).numpy()
    deferred = ix.knn_reruns()
    ix.set_algorithms(knn=gpu.SX_KNN_ALGO_LANE)
    lane = ix.knn(k).numpy()
    assert np.array_equal(tiles, lane)
    pts = ix.fetch()["points"]
    kth = np.sum((pts[lane[:, -1]] - pts) ** 2, axis=1).astype(np.float32)
    r2 = gpu.DeviceArray.from_numpy((kth * rng(seed).uniform(0.3, 3.0, len(kth))).astype(np.float32))
    seeded_lane = ix.knn(k, mode=gpu.SX_KNN_RADIUS, radius2=r2).numpy()
    ix.set_algorithms(knn=gpu.SX_KNN_ALGO_TILES)
    seeded_tiles = ix.knn(k, mode=gpu.SX_KNN_RADIUS, radius2=r2).numpy()
    assert np.array_equal(seeded_lane, lane) and np.array_equal(seeded_tiles, lane)
    print(f"[knn] {kind}: plain tiles deferred {deferred / len(lane):.3%} to one lane per query")


@pytest.mark.parametrize("algo,n", [("SX_SORT_THREE_KERNEL", 2_000_000), ("SX_SORT_ONESWEEP", 2_000_000),
                                    ("SX_SORT_ONESWEEP", 2_100)])
def test_sort_algorithms_agree(gpu, algo, n):
    x, y, z = gpu.gen_scene(2, n)
    a = gpu.Index(len(x), 63).build(x, y, z).fetch()  # hybrid sort (default)
    ixb = gpu.Index(len(x), 63)
    ixb.set_algorithms(sort=getattr(gpu, algo))
    b = ixb.build(x, y, z).fetch()
    for key in ("codes", "sorted_perm", "rep", "points"):
        assert np.array_equal(a[key], b[key]), key
    assert a["oct_nodes"].tobytes() == b["oct_nodes"].tobytes()


# ---------------------------------------------------------------- normals
def _normal_gate(O, pts, nn, normals, q_start=0):
    n64, _ = O.pca_f64(pts, nn)
    err = O.angular_error_deg(normals, n64)
    return float(np.median(err)), float(np.percentile(err, 99)), err


@pytest.mark.parametrize("kind,seed,n,k,bits", [("sphere", 1, 100_000, 8, 30), ("scene", 2, 1_000_000, 16, 63),
                                                 ("gmm", 3, 1_000_000, 32, 63)])
def test_normals_vs_fp64_pca(gpu, O, kind, seed, n, k, bits):
    x, y, z = _workload(gpu, kind, seed, n)
    ix = gpu.Index(n, bits).build(x, y, z)
    nn_d = ix.knn(k)
    nrm, flags = ix.normals(nn_d, k, flags=True)
    nrm = nrm.numpy()
    nn = nn_d.numpy()
    pts = ix.fetch()["points"]
    med, p99, err = _normal_gate(O, pts, nn, nrm)
    print(f"[normals] {kind} k={k}: median {med:.2e} deg, p99 {p99:.2e} deg, max {err.max():.2e}, "
          f"fp64 fallback {flags.numpy()[:len(nrm)].mean():.4f}")
    assert med <= 0.1 and p99 <= 1.0
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    dots = np.sum((0.0 - pts.astype(np.float64)) * nrm, axis=1)
    assert (dots >= -1e-6 * np.linalg.norm(pts, axis=1)).all()
    # the same normals through the fused entry point
    fused = ix.estimate_normals(k).numpy()
    assert np.array_equal(fused, nrm)


def test_normals_fp64_mode(gpu, O):
    x, y, z = gpu.gen_scene(2, 200_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    nn_d = ix.knn(16)
    n64g = ix.normals(nn_d, 16, precision=gpu.SX_PREC_FP64).numpy()
    med, p99, err = _normal_gate(O, ix.fetch()["points"], nn_d.numpy(), n64g)
    assert p99 < 1e-3


def test_host_copy_waits_for_nonblocking_stream(gpu):
    """A host copy without a stream follows work launched on a library stream
    (those streams are non-blocking, so the null stream alone would not order
    the copy): the slow FP64 normals of 1M points, read at once, are whole."""
    x, y, z = gpu.gen_scene(2, 1_000_000)
    st = gpu.Stream()
    ix = gpu.Index(len(x), 63).build(x, y, z)
    nn = ix.knn(16)
    for _ in range(3):
        n64 = ix.normals(nn, 16, precision=gpu.SX_PREC_FP64, stream=st.handle)
        early = n64.numpy()
        st.sync()
        assert np.array_equal(early, n64.numpy(st.handle))
        assert np.all(np.abs(np.linalg.norm(early, axis=1) - 1.0) < 1e-3)


def test_normals_degenerate_classes(gpu, O):
    x, y, z, cls, d = gpu.gen_degenerate(5, 600_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    f = ix.fetch()
    k = 24
    nn_d = ix.knn(k)
    nrm, flags = ix.normals(nn_d, k, flags=True)
    nrm, flags = nrm.numpy(), flags.numpy()
    ucls = cls[f["rep"]]
    udir = d[f["rep"]]
    n64, ev = O.pca_f64(f["points"], nn_d.numpy())
    err = O.angular_error_deg(nrm, n64)
    for c, name in enumerate(("lines", "corners", "slabs")):
        m = ucls == c
        print(f"[degenerate] {name}: median {np.median(err[m]):.2e} p99 {np.percentile(err[m], 99):.2e} "
              f"fp64 {flags[m].mean():.3f}")
    # lines: the normal is only defined up to a rotation about the line; it
    # must be perpendicular to the line direction
    lines = ucls == 0

    def off_perp(n):
        return np.degrees(np.arcsin(np.clip(np.abs(np.sum(n[lines] * udir[lines], axis=1)), 0, 1)))

    perp_gpu, perp_f64 = off_perp(nrm), off_perp(n64)
    print(f"[degenerate] lines: angle to the line's normal plane p99 gpu {np.percentile(perp_gpu, 99):.3f} "
          f"fp64 {np.percentile(perp_f64, 99):.3f} deg")
    assert np.percentile(perp_gpu, 99) <= np.percentile(perp_f64, 99) + 0.1
    for c in (1, 2):
        m = ucls == c
        assert np.median(err[m]) <= 0.1 and np.percentile(err[m], 99) <= 1.0
