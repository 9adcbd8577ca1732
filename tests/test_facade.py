"""The drop-in C++ API (shapeextraction_amd/cpp): host parts on the CPU, the
example driver (the reference's main.cu flow) on the GPU."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CPP = ROOT / "shapeextraction_amd" / "cpp"


def test_facade_host_selftest(tmp_path):
    exe = CPP / "facade_selftest"
    assert exe.exists(), "build the facade: make -C shapeextraction_amd/cpp"
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "facade selftest ok" in out.stdout


def _read_ascii_ply(path):
    lines = path.read_text().splitlines()
    n = int(next(l for l in lines if l.startswith("element vertex")).split()[-1])
    start = lines.index("end_header") + 1
    return np.array([[float(v) for v in l.split()] for l in lines[start:start + n]], np.float64)


@pytest.mark.gpu
def test_example_driver_matches_c_abi(gpu, tmp_path):
    out = tmp_path / "cloud.ply"
    r = subprocess.run([str(CPP / "shapes"), str(out), "synthetic:20000", "16"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    ply = _read_ascii_ply(out)
    x, y, z = gpu.gen_scene(2, 20000)
    ix = gpu.Index(len(x), 63).build(x, y, z)
    pts = ix.fetch()["points"]
    normals = ix.estimate_normals(16).numpy()
    assert ply.shape == (len(pts), 6)
    # ASCII output has 6 significant digits
    assert np.allclose(ply[:, :3], pts, rtol=1e-5, atol=1e-5)
    assert np.allclose(ply[:, 3:], normals, rtol=1e-5, atol=1e-5)


def _read_binary_ply(path):
    data = path.read_bytes()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode().splitlines()
    assert "format binary_little_endian 1.0" in header
    n = int(next(l for l in header if l.startswith("element vertex")).split()[-1])
    cols = sum(1 for l in header if l.startswith("property float"))
    return np.frombuffer(data[end:], np.float32, count=n * cols).reshape(n, cols)


def _parse_fixture(path):
    """The fixture's points as the loaders define them (src/DataIO.cpp:33-120)."""
    if path.suffix == ".bin":
        return np.fromfile(path, np.float32).reshape(-1, 4)[:, :3]
    rows = []
    for line in path.read_text().splitlines():
        if path.suffix == ".obj":
            if line.startswith("v "):
                rows.append([float(v) for v in line.split()[1:4]])
        else:
            rows.append([float(v) for v in line.split()[:3]])
    return np.array(rows, np.float32)


def test_loader_fixtures_are_reproducible(tmp_path):
    """The committed loader fixtures are exactly what their script writes."""
    import runpy
    import shutil

    gold = ROOT / "tests" / "golden"
    script = tmp_path / "make_loader_fixtures.py"
    shutil.copy(gold / "make_loader_fixtures.py", script)
    runpy.run_path(str(script), run_name="__main__")
    for name in ("loader_kitti.bin", "loader_scene.obj", "loader_semantic3d.txt"):
        assert (tmp_path / name).read_bytes() == (gold / name).read_bytes(), name
    pts = _parse_fixture(gold / "loader_scene.obj")
    assert pts.shape == (3000, 3)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["loader_kitti.bin", "loader_scene.obj", "loader_semantic3d.txt"])
def test_loaders_through_drop_in_pipeline(gpu, O, tmp_path, name):
    """DataIO::loadFile -> RadixTree -> Octree -> estimateNormals<16> ->
    binary PLY, against the oracle on the same points."""
    src = ROOT / "tests" / "golden" / name
    out = tmp_path / "out.bin.ply"
    r = subprocess.run([str(CPP / "shapes"), str(out), str(src), "16"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    ply = _read_binary_ply(out)
    p = _parse_fixture(src)
    x, y, z = (np.ascontiguousarray(p[:, i]) for i in range(3))
    b, _ = O.bbox(x, y, z)
    codes = O.encode(x, y, z, b[0], b[2], 63)
    sc, perm = O.sort_pairs(codes)
    uq, rep, _ = O.unique(sc, perm)
    pts = O.decode(uq, b[0], b[2], 63)
    assert ply.shape == (len(pts), 6)
    assert np.array_equal(ply[:, :3], pts)  # decoded points, bit for bit
    nn = O.knn_brute(pts, 16)
    n64, _ = O.pca_f64(pts, nn)
    err = O.angular_error_deg(ply[:, 3:], n64)
    assert np.median(err) <= 0.1 and np.percentile(err, 99) <= 1.0
    # the same normals as the C ABI path on the same points
    ix = gpu.Index(len(x), 63).build(x, y, z)
    assert np.array_equal(ix.knn(16).numpy(), nn)
    assert np.array_equal(ply[:, 3:], ix.estimate_normals(16).numpy())


@pytest.mark.gpu
def test_facade_gpu_check(gpu, O, tmp_path):
    """Host-data Octree round trip, Octree::knnSearch<k> on arbitrary queries
    and estimateNormalsInputOrder<16>, through the C++ drop-in API."""
    r = subprocess.run([str(CPP / "facade_gpu_check"), str(tmp_path), "200000"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "facade gpu check ok" in r.stdout
    pts = np.fromfile(tmp_path / "points.f32", np.float32).reshape(-1, 3)
    q = np.fromfile(tmp_path / "queries.f32", np.float32).reshape(-1, 3)
    for k in (8, 16, 24, 32):
        rows = np.fromfile(tmp_path / f"knn{k}.i32", np.int32).reshape(-1, k)
        assert np.array_equal(rows, O.knn_brute_points(pts, q, k)), k
    x, y, z = (np.fromfile(tmp_path / f"cloud_{a}.f32", np.float32) for a in "xyz")
    b, _ = O.bbox(x, y, z)
    sc, perm = O.sort_pairs(O.encode(x, y, z, b[0], b[2], 63))
    uq, rep, uoi = O.unique(sc, perm)
    assert len(uq) == len(pts)
    n_in = np.fromfile(tmp_path / "normals_input.f32", np.float32).reshape(-1, 3)
    n_u = np.fromfile(tmp_path / "normals_unique.f32", np.float32).reshape(-1, 3)
    assert n_in.shape == (len(x), 3)
    assert np.array_equal(n_in, n_u[uoi])
