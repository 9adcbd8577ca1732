"""Writes the committed fixtures under tests/golden/.

The reference holds no golden vectors, test data or fixtures of its own, so
these fixtures are produced by this repository's CPU oracle (itself pinned by
tests/test_oracle.py) and generators:

  checksums.json   FNV-64 of generated workload arrays (generator drift)
  small_scene.npz  a 3000-point scene (seed 77) with every oracle stage output:
                   cube bounds, codes, sort permutation, unique codes and
                   representatives, radix tree (textbook root), octree nodes,
                   decoded points, exact kNN rows (k=8, 16) and FP64 normals.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import shapeextraction_amd as sx  # noqa: E402
from oracle import oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent


def main():
    sums = {
        "sphere100k_x": sx.fnv64(sx.gen_sphere(1, 100_000, 0.002)[0]),
        "scene200k_z": sx.fnv64(sx.gen_scene(2, 200_000)[2]),
        "gmm200k_y": sx.fnv64(sx.gen_gmm_scene(3, 200_000)[1]),
        "batch_sizes": sx.fnv64(sx.gen_batch_sizes(4, 512)),
        "degenerate30k_x": sx.fnv64(sx.gen_degenerate(5, 30_000)[0]),
    }
    (HERE / "checksums.json").write_text(json.dumps({k: str(v) for k, v in sums.items()}, indent=1) + "\n")

    x, y, z = sx.gen_scene(77, 3000)
    b, _ = O.bbox(x, y, z)
    codes = O.encode(x, y, z, b[0], b[2], 63)
    sc, perm = O.sort_pairs(codes)
    uq, rep, uoi = O.unique(sc, perm)
    rt = O.karras(uq, 63, 0)
    nodes = O.octree(uq, 63, b[0], b[2], rt)
    pts = O.decode(uq, b[0], b[2], 63)
    nn8 = O.knn_brute(pts, 8)
    nn16 = O.knn_brute(pts, 16)
    n64, _ = O.pca_f64(pts, nn16)
    np.savez_compressed(
        HERE / "small_scene.npz",
        x=x, y=y, z=z, bounds=b, codes=codes, perm=perm, unique=uq, rep=rep, unique_of_input=uoi,
        prefix_n=rt["prefix_n"], left_child=rt["left_child"], has_leaf_left=rt["has_leaf_left"],
        has_leaf_right=rt["has_leaf_right"], parent=rt["parent"], oct_nodes=nodes.view(np.uint8).reshape(-1, 56),
        points=pts, nn8=nn8, nn16=nn16, normals_f64=n64,
    )
    print("wrote", HERE / "checksums.json", HERE / "small_scene.npz")


if __name__ == "__main__":
    main()
