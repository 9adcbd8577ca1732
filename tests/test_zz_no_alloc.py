"""sx_estimate_normals_noalloc (no device allocation: the neighbour rows live
in the library's fixed per-GPU array) against sx_estimate_normals (rows in a
buffer the library allocates), through the C ABI, on two calls in a row, the
second smaller; and its size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _estimate(gpu, ix, q0, nq, k, noalloc, mode=0, r2=None):
    out = gpu.DeviceArray((nq, 3), np.float32)
    out.fill_(0xFF)  # NaN: any row left unwritten shows
    r2p = r2.ptr if r2 is not None else None
    if noalloc:
        gpu._lib.call("sx_estimate_normals_noalloc", ix.handle, q0, nq, k, mode, r2p, 0, out.ptr, None)
    else:
        gpu._lib.call("sx_estimate_normals", ix.handle, q0, nq, k, mode, r2p, 0, None, out.ptr, None)
    return out.numpy()


def test_noalloc_equals_estimate_normals(gpu):
    x, y, z = gpu.gen_scene(12, 600_000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    U = ix.n_unique
    for q0, nq, k in ((0, U, 16), (1000, 50_000, 32)):  # the second call smaller
        ref = _estimate(gpu, ix, q0, nq, k, noalloc=False)
        got = _estimate(gpu, ix, q0, nq, k, noalloc=True)
        assert np.isfinite(got).all()
        assert np.array_equal(got, ref), (q0, nq, k)
    # the seeded search through the same array
    from shapeextraction_amd import learned

    model = learned.RadiusModel.load(learned.default_model_path(16))
    r2 = model.predict_index_radius2(ix, 0, 70_000)
    ref = _estimate(gpu, ix, 0, 70_000, 16, noalloc=False)
    assert np.array_equal(_estimate(gpu, ix, 0, 70_000, 16, noalloc=True, mode=1, r2=r2), ref)


def test_noalloc_rejects_oversized_range(gpu):
    x, y, z = gpu.gen_sphere(1, 10_000, 0.002)
    ix = gpu.Index(len(x), 30).build(x, y, z)
    out = gpu.DeviceArray((16,), np.float32)
    lib = gpu._lib.load()
    assert lib.sx_estimate_normals_noalloc(ix.handle, 0, (1 << 24) // 16 + 1, 16, 0, None, 0, out.ptr, None) == 1
    assert lib.sx_estimate_normals_noalloc(ix.handle, 0, (1 << 24) // 32 + 1, 32, 0, None, 0, out.ptr, None) == 1
