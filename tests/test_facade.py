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
