"""Writes the small point-cloud files the loader tests read through the
drop-in DataIO (src/DataIO.cpp:13-120 of the reference):

  loader_kitti.bin      KITTI velodyne records: float32 x, y, z, reflectance
  loader_scene.obj      OBJ vertex lines ("v x y z"), with normals, faces and
                        comments the loader must skip
  loader_semantic3d.txt Semantic3D rows "x y z r g b" (the reference reads
                        intensity-free files with colour)

Each holds 3,000 points of a seeded scene (two planes, a sphere, a cylinder
and scattered points).  Coordinates are multiples of 1/1024 below 128, so the
decimal text is exact and every parser reads the same float32 values.
Run from the repository root: python tests/golden/make_loader_fixtures.py
"""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
N = 3000


def scene(seed: int) -> np.ndarray:
    g = np.random.default_rng(seed)
    parts = []
    n = N // 5
    uv = g.random((n, 2))
    parts.append(np.c_[10 + 60 * uv[:, 0], 20 + 50 * uv[:, 1], np.full(n, 5.0)])          # floor
    uv = g.random((n, 2))
    parts.append(np.c_[np.full(n, 80.0), 10 + 70 * uv[:, 0], 5 + 40 * uv[:, 1]])          # wall
    v = g.normal(size=(n, 3))
    parts.append(np.array([40.0, 40.0, 30.0]) + 12.0 * v / np.linalg.norm(v, axis=1, keepdims=True))
    t = g.random(n) * 2 * np.pi
    h = g.random(n) * 30
    parts.append(np.c_[60 + 6 * np.cos(t), 70 + 6 * np.sin(t), 10 + h])                   # cylinder
    parts.append(g.random((N - 4 * n, 3)) * 100.0)                                          # clutter
    p = np.vstack(parts)
    return np.round(p * 1024.0) / 1024.0  # exact in float32 and in decimal


def fmt(v: float) -> str:
    return repr(float(v))


def main() -> None:
    p = scene(31)
    kitti = np.c_[p, np.full(len(p), 255.0)].astype(np.float32)
    (HERE / "loader_kitti.bin").write_bytes(kitti.tobytes())
    q = scene(32)
    with open(HERE / "loader_scene.obj", "w") as f:
        f.write("# loader fixture: vertices of a seeded scene\n")
        for i, (x, y, z) in enumerate(q):
            f.write(f"v {fmt(x)} {fmt(y)} {fmt(z)}\n")
            if i % 3 == 2:
                f.write("vn 0 0 1\n")
        f.write("f 1 2 3\nf 4 5 6\n")
    r = scene(33)
    g = np.random.default_rng(34)
    rgb = g.integers(0, 256, size=(len(r), 3))
    with open(HERE / "loader_semantic3d.txt", "w") as f:
        for (x, y, z), (cr, cg, cb) in zip(r, rgb):
            f.write(f"{fmt(x)} {fmt(y)} {fmt(z)} {cr} {cg} {cb}\n")


if __name__ == "__main__":
    main()
