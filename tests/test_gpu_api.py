"""GPU tests of the query and build entry points around the hot path, through
the C ABI: external-query kNN (the reference's Octree::knnSearch<k> query API,
src/Octree.hpp:214-242) against brute force, rows of positions past the
unique count, the sort's look-back give-up report, rows in input order
(src/main.cu:119-120 returns Morton order only) and the index built from
given codes (the path of the host-data Octree constructor,
src/Octree.cu:179-192)."""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu


def _cloud(sx, kind, n, seed):
    if kind == "sphere":
        return sx.gen_sphere(seed, n, 0.002)
    return sx.gen_scene(seed, n)


def _queries(g, pts, n_in, n_out, n_on):
    """Points inside the cube, outside it (near and far) and exactly on index points."""
    lo, hi = pts.min(0), pts.max(0)
    ext = hi - lo
    inside = lo + g.random((n_in, 3)) * ext
    near = lo - 0.5 * ext + g.random((n_out // 2, 3)) * 2.0 * ext
    far = (g.normal(size=(n_out - n_out // 2, 3)) * 50.0 + 1.0) * ext.max()
    on = pts[g.choice(len(pts), n_on, replace=False)]
    return np.vstack([inside, near, far, on]).astype(np.float32)


@pytest.mark.parametrize("root_mode", [0, 1])
@pytest.mark.parametrize("k", [8, 16, 24, 32])
def test_knn_points_vs_brute(gpu, O, k, root_mode):
    x, y, z = gpu.gen_scene(21, 200_000)
    ix = gpu.Index(len(x), 63, root_mode=root_mode).build(x, y, z)
    pts = ix.fetch()["points"]
    q = _queries(rng(100 + k), pts, 3000, 1000, 1000)
    got = ix.knn_points(q, k).numpy()
    assert np.array_equal(got, O.knn_brute_points(pts, q, k))
    # a query on an index point finds that point first (distance 0, lowest index of equal points)
    on = q[-1000:]
    assert np.array_equal(pts[got[-1000:, 0]], on)


def test_knn_points_sphere_30bit_and_bucket_invariance(gpu, O):
    x, y, z = gpu.gen_sphere(1, 100_000, 0.002)
    ix = gpu.Index(len(x), 30).build(x, y, z)
    pts = ix.fetch()["points"]
    q = _queries(rng(7), pts, 2000, 500, 500)
    ref = O.knn_brute_points(pts, q, 16)
    for b in (1, 16, 64):
        ix.set_bucket(b)
        assert np.array_equal(ix.knn_points(q, 16).numpy(), ref), b
    ix.set_bucket(16)


def test_knn_points_fewer_points_than_k(gpu, O):
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 3], [1, 1, 1]], np.float32)
    ix = gpu.Index(len(P), 63).build(*(np.ascontiguousarray(P[:, i]) for i in range(3)))
    pts = ix.fetch()["points"]
    q = np.array([[0.1, 0.1, 0.1], [5, 5, 5], [-3, 0, 0]], np.float32)
    got = ix.knn_points(q, 8).numpy()
    assert np.array_equal(got, O.knn_brute_points(pts, q, 8))
    assert (got[:, 5:] == got[:, :1]).all()  # padded with the first entry


def test_rows_past_unique_count(gpu, O):
    """Before a view has read the counts, positions up to the input count are
    accepted; those at or past the unique count get rows of -1 / NaN."""
    x, y, z = gpu.gen_scene(2, 50_000)
    x = np.concatenate([x, x[:5000]])  # 5,000 duplicate inputs
    y = np.concatenate([y, y[:5000]])
    z = np.concatenate([z, z[:5000]])
    N = len(x)
    ix = gpu.Index(N, 63).build(x, y, z)
    nn = ix.knn(16, q_start=0, n_q=N)  # no view yet: checked against N
    nrm = ix.normals(nn, 16, q_start=0, n_q=N).numpy()
    got = nn.numpy()
    U = ix.n_unique  # reads the view
    assert U <= 50_000
    assert (got[U:] == -1).all()
    assert np.isnan(nrm[U:]).all()
    pts = ix.fetch()["points"]
    q = np.sort(rng(3).choice(U, 5000, replace=False))
    assert np.array_equal(got[q], O.knn_grid(pts, 16, q))
    assert np.isfinite(nrm[:U]).all()
    with pytest.raises(gpu.SxError):  # now checked against U
        ix.knn(16, q_start=0, n_q=N)


def test_sort_lookback_giveup_is_reported(gpu, O):
    """A look-back that gives up (forced by a spin limit of 0: the first wait
    gives up, so every kernel still exits) is reported instead of SX_OK: by
    sx_sort_check for the stage call, by the view for an index (whose later
    stages then run on no points)."""
    g = rng(17)
    n = 4_000_000
    keys = g.integers(0, 2**63, size=n, dtype=np.uint64)
    dk = gpu.DeviceArray.from_numpy(keys)
    ref = np.argsort(keys, kind="stable")
    x, y, z = gpu.gen_scene(2, 2_000_000)
    raised = 0
    try:
        gpu.set_sort_spin_limit(0)
        for attempt in range(3):
            try:
                ko, vo = gpu.sort_pairs(dk, 63, check=True)
                assert np.array_equal(vo.numpy(), ref), "a sort reported OK must be correct"
            except gpu.SxError as e:
                assert e.code == 7  # SX_EINTERNAL
                raised += 1
            ix = gpu.Index(len(x), 63).build(x, y, z)
            try:
                ix.view()
            except gpu.SxError as e:
                assert e.code == 7
                raised += 1
                rows = ix.knn(16, q_start=0, n_q=1000).numpy()  # an empty index: rows of -1
                assert (rows == -1).all()
    finally:
        gpu.set_sort_spin_limit(gpu.DEFAULT_SORT_SPIN_LIMIT)
    print(f"[sort] give-ups reported: {raised} of 6 sorts")
    assert raised > 0
    ko, vo = gpu.sort_pairs(dk, 63, check=True)
    assert np.array_equal(vo.numpy(), ref)


@pytest.mark.parametrize("kind,n,bits,k", [("scene", 300_000, 63, 16), ("sphere", 100_000, 30, 8)])
def test_rows_in_input_order(gpu, O, kind, n, bits, k):
    x, y, z = _cloud(gpu, kind, n, 5)
    x = np.concatenate([x, x[::7]])  # duplicates map to their code's row
    y = np.concatenate([y, y[::7]])
    z = np.concatenate([z, z[::7]])
    ix = gpu.Index(len(x), bits).build(x, y, z)
    nn = ix.knn(k)
    nrm = ix.normals(nn, k)
    nrm_in = ix.to_input_order(nrm).numpy()
    nn_in = ix.to_input_order(nn).numpy()
    b, _ = O.bbox(x, y, z)
    codes = O.encode(x, y, z, b[0], b[2], bits)
    sc, perm = O.sort_pairs(codes)
    uq, rep, uoi = O.unique(sc, perm)
    assert nrm_in.shape == (len(x), 3) and nn_in.shape == (len(x), k)
    assert np.array_equal(nrm_in, nrm.numpy()[uoi])
    assert np.array_equal(nn_in, nn.numpy()[uoi])
    # each input point's own decoded position is its row's first neighbour
    pts = O.decode(uq, b[0], b[2], bits)
    assert np.array_equal(pts[nn_in[:, 0]], pts[uoi])


def test_index_from_codes_equals_index_from_points(gpu, O):
    x, y, z = gpu.gen_scene(2, 500_000)
    a = gpu.Index(len(x), 63).build(x, y, z)
    fa = a.fetch()
    b = gpu.Index(fa["n_unique"], 63).build_codes(fa["codes"], fa["bounds"][0], fa["bounds"][2])
    fb = b.fetch()
    assert fb["n_unique"] == fa["n_unique"] and fb["n_input"] == fa["n_unique"]
    assert np.array_equal(fb["points"], fa["points"])
    assert fb["oct_nodes"].tobytes() == fa["oct_nodes"].tobytes()
    assert np.array_equal(fb["rep"], np.arange(fa["n_unique"], dtype=np.uint32))
    for key in ("prefix_n", "left_child", "has_leaf_left", "has_leaf_right", "parent"):
        assert np.array_equal(fb[key], fa[key]), key
    nn_a, nn_b = a.knn(16), b.knn(16)
    assert np.array_equal(nn_a.numpy(), nn_b.numpy())
    assert np.array_equal(a.normals(nn_a, 16).numpy(), b.normals(nn_b, 16).numpy())


def test_index_from_codes_rejects_unsorted(gpu):
    codes = np.array([5, 9, 9, 12, 3], np.uint64)
    ix = gpu.Index(8, 63).build_codes(codes, 0.0, 1.0)
    with pytest.raises(gpu.SxError) as e:
        ix.view()
    assert e.value.code == 5  # SX_EDATA
    assert (ix.knn(8, q_start=0, n_q=5).numpy() == -1).all()  # nothing was built on them
    wide = np.array([1, 2, 1 << 40], np.uint64)
    with pytest.raises(gpu.SxError):
        gpu.Index(8, 30).build_codes(wide, 0.0, 1.0).view()
