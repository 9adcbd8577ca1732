"""Learned search radius on the GPU: BF16-MFMA forward and backward kernels
against the CPU oracle (with the same bf16 rounding points), training, and
exactness of the radius-seeded kNN."""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L(gpu):
    from shapeextraction_amd import learned

    return learned


def _random_model(L, seed, n_params=4801):
    g = rng(seed)
    m = L.RadiusModel(16, max_batch=4096)
    P = (g.normal(size=n_params) * 0.25).astype(np.float32)
    m.set_params(P)
    mu = g.normal(size=8).astype(np.float32)
    sg = g.uniform(0.5, 2.0, 8).astype(np.float32)
    m.set_normalizer(mu, sg, 0.3)
    return m, P, mu, sg


@pytest.mark.parametrize("n", [1, 15, 16, 64, 100, 1000, 4096])
def test_mlp_forward_matches_oracle(gpu, O, L, n):
    m, P, mu, sg = _random_model(L, 1)
    g = rng(2)
    f = g.normal(size=(n, 8)).astype(np.float32) * sg + mu
    y = m.forward(gpu.DeviceArray.from_numpy(f)).numpy()
    ref = O.mlp_forward(P, f, mu, sg, emulate=True)
    err = np.abs(y - ref)
    assert np.median(err) < 1e-3 * (1 + np.abs(ref).max())
    assert err.max() < 2e-2 * (1 + np.abs(ref).max())
    r2 = m.predict_radius2(gpu.DeviceArray.from_numpy(f)).numpy()
    assert np.allclose(r2, np.exp(y + 0.3), rtol=1e-5)


@pytest.mark.parametrize("n", [64, 333, 4096])
def test_mlp_gradient_matches_oracle(gpu, O, L, n):
    m, P, mu, sg = _random_model(L, 3)
    g = rng(4)
    f = (g.normal(size=(n, 8)) * sg + mu).astype(np.float32)
    t = g.normal(size=n).astype(np.float32)
    grad, loss = m.backward(gpu.DeviceArray.from_numpy(f), gpu.DeviceArray.from_numpy(t))
    ref, ref_loss = O.mlp_grad(P, f, t, mu, sg, emulate=True)
    rel = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    cos = grad @ ref / (np.linalg.norm(grad) * np.linalg.norm(ref))
    print(f"[mlp grad] n={n}: relative error {rel:.2e}, cosine {cos:.8f}, loss {loss:.5f} vs {ref_loss:.5f}")
    assert rel < 1e-2 and cos > 0.9999
    assert abs(loss - ref_loss) < 1e-3 * ref_loss
    # per block: every parameter group gets its own gradient right
    for a, b in ((0, 512), (512, 576), (576, 4672), (4672, 4736), (4736, 4800), (4800, 4801)):
        nb = np.linalg.norm(ref[a:b])
        assert np.linalg.norm(grad[a:b] - ref[a:b]) <= 2e-2 * nb + 1e-6, (a, b)


def test_mlp_training_fits_a_smooth_target(gpu, L):
    g = rng(5)
    X = g.normal(size=(65536, 8)).astype(np.float32)
    T = (0.5 * X[:, 0] - 0.3 * X[:, 1] * X[:, 2] + np.sin(X[:, 3])).astype(np.float32)
    m = L.RadiusModel(16, max_batch=4096)
    m.init(0)
    dX, dT = gpu.DeviceArray.from_numpy(X), gpu.DeviceArray.from_numpy(T)
    y0 = m.forward(dX).numpy()
    for s in range(600):
        off = (s * 4096) % (65536 - 4096)
        m.adam_step(dX.offset(off * 8), dT.offset(off), 4096, 3e-3)
    y1 = m.forward(dX).numpy()
    mse0, mse1 = np.mean((y0 - T) ** 2), np.mean((y1 - T) ** 2)
    print(f"[mlp train] mse {mse0:.4f} -> {mse1:.4f}")
    assert mse1 < 0.1 * mse0 and mse1 < 0.05


def test_model_file_round_trip(gpu, L, tmp_path):
    m, P, mu, sg = _random_model(L, 6)
    m.save(tmp_path / "m.bin")
    m2 = L.RadiusModel.load(tmp_path / "m.bin")
    assert np.array_equal(m2.params(), P) and np.array_equal(m2.mu, mu) and m2.margin == np.float32(0.3)


def test_features_are_finite_and_consistent(gpu, L):
    x, y, z = gpu.gen_scene(2, 300_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    f = L.features(ix).numpy()
    assert np.isfinite(f).all()
    U = ix.n_unique
    # counts are monotone with coarser cells; depth >= 0; slots within [0, 8]
    assert (f[:, 1] >= f[:, 0]).all() and (f[:, 2] >= f[:, 1]).all() and (f[:, 3] >= f[:, 2]).all()
    assert (f[:, 0] >= 1).all()  # the leaf cell holds the point and at least one other
    assert (f[:, 5] >= 0).all() and (f[:, 6] + f[:, 7] <= 8).all() and (f[:, 6] >= 1).all()
    part = L.features(ix, 1000, 5000).numpy()
    assert np.array_equal(part, f[1000:6000])
    assert len(f) == U


def _features_by_definition(f, bits):
    """The eight features from their definition, on the fetched (oracle-pinned)
    tree and octree: point counts of the Morton cells at depths D0..D0-3
    holding each point (D0 = prefix length of its radix parent / 3, clamped at
    the root's depth), log2 edge, depth below the root, and the leaf / child
    slots of the octree node holding the point."""
    codes = f["codes"]
    U = len(codes)
    M = U - 1
    pn = f["prefix_n"].astype(np.int64)
    lc = f["left_child"].astype(np.int64)[:M]
    hll = f["has_leaf_left"][:M] == 1
    hlr = f["has_leaf_right"][:M] == 1
    lp = np.empty(U, np.int64)
    i = np.arange(M)
    lp[lc[hll]] = i[hll]
    lp[lc[hlr] + 1] = i[hlr]
    d0 = pn[lp] // 3
    root = pn[0] // 3
    out = np.zeros((U, 8), np.float64)
    for s in range(4):
        L = np.maximum(d0 - s, root)
        cnt = np.zeros(U)
        for lv in np.unique(L):
            sel = L == lv
            pv = codes >> np.uint64(bits - 3 * int(lv))
            cnt[sel] = np.searchsorted(pv, pv[sel], "right") - np.searchsorted(pv, pv[sel], "left")
        out[:, s] = np.log2(cnt)
    out[:, 4] = np.log2(np.float64(f["bounds"][2])) - d0
    out[:, 5] = d0 - root
    oct = f["oct_nodes"]
    holder = np.full(U, -1, np.int64)
    for t in range(8):
        has = (oct["child_leaf_mask"] >> t) & 1 == 1
        holder[oct["children"][has, t]] = np.nonzero(has)[0]
    pop = np.vectorize(lambda v: bin(int(v)).count("1"))
    out[:, 6] = pop(oct["child_leaf_mask"][holder])
    out[:, 7] = pop(oct["child_node_mask"][holder])
    return out


@pytest.mark.parametrize("root_mode", [0, 1])
def test_features_match_definition(gpu, L, root_mode):
    """Features against their definition.  With the textbook root the kernel
    reads cell counts from the octree's parent links; with the reference root
    it climbs the traversal tree (f6, f7 then come from the reference octree,
    whose root differs, so only f0..f5 are compared there)."""
    x, y, z = gpu.gen_scene(2, 200_000)
    ix = gpu.Index(len(x), 63, root_mode=root_mode).build(x, y, z)
    got = L.features(ix).numpy().astype(np.float64)
    want = _features_by_definition(ix.fetch() if root_mode == 0 else _textbook_fetch(gpu, x, y, z), 63)
    assert np.allclose(got[:, :5], want[:, :5], rtol=0, atol=1e-5)
    assert np.array_equal(got[:, 5], want[:, 5])
    if root_mode == 0:
        assert np.array_equal(got[:, 6:], want[:, 6:])


def _textbook_fetch(gpu, x, y, z):
    return gpu.Index(len(x), 63).build(x, y, z).fetch()


def test_radius_predict_fused_equals_two_steps(gpu, L):
    """sx_radius_predict (features + MLP in one kernel) returns the same bits
    as sx_radius_features followed by sx_mlp_predict_radius2, also on a slice
    whose length is not a multiple of the 64-query wave."""
    model = L.RadiusModel.load(L.default_model_path(16))
    x, y, z = gpu.gen_scene(7, 500_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    two = model.predict_radius2(L.features(ix)).numpy()
    fused = model.predict_index_radius2(ix).numpy()
    assert np.array_equal(fused, two)
    part = model.predict_index_radius2(ix, 777, 12_345).numpy()
    assert np.array_equal(part, two[777:777 + 12_345])


def test_learned_radius_knn_is_exact(gpu, L):
    """Train a small model quickly, then check that the seeded search returns
    exactly the plain result and report what it saves."""
    model = L.train(16, steps=300, n_clouds=6, log=print)
    x, y, z = gpu.gen_scene(2, 1_000_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    plain = ix.knn(16).numpy()
    r2 = model.predict_radius2(L.features(ix))
    ix.knn_reruns(reset=True)
    got = ix.knn(16, mode=gpu.SX_KNN_RADIUS, radius2=r2).numpy()
    reruns = ix.knn_reruns()
    assert np.array_equal(got, plain)
    print(f"[learned] reruns {reruns / len(plain):.3%}")
    assert reruns < 0.5 * len(plain)
