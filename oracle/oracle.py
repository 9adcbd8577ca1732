"""CPU ORACLE wrapper — test infrastructure only.

numpy front end of ``liboracle.so`` (oracle.cpp), the CPU restatement of the
reference's normal-estimation path.  Only tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of bench.py may import this module, and only as
the checker or the reported CPU baseline; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

OTNODE_DTYPE = np.dtype(
    [
        ("children", "<i4", (8,)),
        ("corner", "<f4", (3,)),
        ("cell_size", "<f4"),
        ("child_node_mask", "<i4"),
        ("child_leaf_mask", "<i4"),
    ]
)

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, I, I64, F = C.c_void_p, C.c_int, C.c_int64, C.c_float
        sig = {
            "orc_threads": (I, []),
            "orc_bbox": (I64, [P, P, P, I64, P]),
            "orc_encode": (None, [P, P, P, I64, F, F, I, P]),
            "orc_decode": (None, [P, I64, F, F, I, P]),
            "orc_interleave": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32, I]),
            "orc_sort_pairs": (None, [P, I64, P, P]),
            "orc_unique": (I64, [P, P, I64, P, P, P]),
            "orc_karras": (None, [P, I64, I, I, P, P, P, P, P]),
            "orc_trie": (None, [P, I64, I, P, P, P, P, P]),
            "orc_octree": (I64, [P, I64, I, F, F, P, P, P, P, P, P, I64]),
            "orc_octree_cells": (I64, [P, I64, I, I, P]),
            "orc_knn_brute": (None, [P, I64, I64, I64, I, P]),
            "orc_knn_brute_points": (None, [P, I64, P, I64, I, P]),
            "orc_knn_grid": (None, [P, I64, P, I64, I, P]),
            "orc_knn_reference": (None, [P, P, I64, I64, I64, I, P]),
            "orc_pca_f64": (None, [P, P, I64, I, P, P]),
            "orc_jacobi3": (None, [P, P, P]),
            "orc_normals_ref_f32": (None, [P, P, I64, I64, I, F, F, F, P]),
            "orc_pipeline": (I64, [P, P, P, I64, I, I, I, P, P]),
            "orc_mlp_num_params": (I, []),
            "orc_mlp_forward": (None, [P, P, I64, P, P, I, P]),
            "orc_mlp_grad": (C.c_double, [P, P, P, I64, P, P, I, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def threads() -> int:
    return int(lib().orc_threads())


def bbox(x, y, z):
    x, y, z = _c(x, np.float32), _c(y, np.float32), _c(z, np.float32)
    out = np.empty(3, np.float32)
    bad = lib().orc_bbox(_p(x), _p(y), _p(z), len(x), _p(out))
    return out, int(bad)


def encode(x, y, z, mn: float, rng: float, code_bits: int) -> np.ndarray:
    x, y, z = _c(x, np.float32), _c(y, np.float32), _c(z, np.float32)
    out = np.empty(len(x), np.uint64)
    lib().orc_encode(_p(x), _p(y), _p(z), len(x), float(mn), float(rng), code_bits, _p(out))
    return out


def decode(codes, mn: float, rng: float, code_bits: int) -> np.ndarray:
    codes = _c(codes, np.uint64)
    out = np.empty((len(codes), 3), np.float32)
    lib().orc_decode(_p(codes), len(codes), float(mn), float(rng), code_bits, _p(out))
    return out


def interleave(x: int, y: int, z: int, code_bits: int) -> int:
    return int(lib().orc_interleave(x, y, z, code_bits))


def sort_pairs(keys):
    keys = _c(keys, np.uint64)
    ko = np.empty_like(keys)
    perm = np.empty(len(keys), np.uint32)
    lib().orc_sort_pairs(_p(keys), len(keys), _p(ko), _p(perm))
    return ko, perm


def unique(sorted_keys, perm):
    s = _c(sorted_keys, np.uint64)
    p = _c(perm, np.uint32)
    uq = np.empty_like(s)
    rep = np.empty(len(s), np.uint32)
    uoi = np.empty(len(s), np.int32)
    U = lib().orc_unique(_p(s), _p(p), len(s), _p(uq), _p(rep), _p(uoi))
    return uq[:U].copy(), rep[:U].copy(), uoi


def _tree_out(U):
    return (np.zeros(U, np.uint8), np.zeros(U, np.int32), np.zeros(U, np.uint8), np.zeros(U, np.uint8),
            np.zeros(U, np.int32))


def karras(codes, code_bits: int, root_mode: int = 0) -> dict:
    codes = _c(codes, np.uint64)
    U = len(codes)
    pn, lc, hll, hlr, par = _tree_out(U)
    lib().orc_karras(_p(codes), U, code_bits, root_mode, _p(pn), _p(lc), _p(hll), _p(hlr), _p(par))
    return {"prefix_n": pn, "left_child": lc, "has_leaf_left": hll, "has_leaf_right": hlr, "parent": par}


def trie(codes, code_bits: int) -> dict:
    codes = _c(codes, np.uint64)
    U = len(codes)
    pn, lc, hll, hlr, par = _tree_out(U)
    lib().orc_trie(_p(codes), U, code_bits, _p(pn), _p(lc), _p(hll), _p(hlr), _p(par))
    return {"prefix_n": pn, "left_child": lc, "has_leaf_left": hll, "has_leaf_right": hlr, "parent": par}


def octree(codes, code_bits: int, mn: float, rng: float, rt: dict) -> np.ndarray:
    codes = _c(codes, np.uint64)
    U = len(codes)
    arrs = [_c(rt[k], dt) for k, dt in (("prefix_n", np.uint8), ("left_child", np.int32),
                                         ("has_leaf_left", np.uint8), ("has_leaf_right", np.uint8),
                                         ("parent", np.int32))]
    args = [_p(codes), U, code_bits, float(mn), float(rng)] + [_p(a) for a in arrs]
    O = lib().orc_octree(*args, None, 0)
    nodes = np.zeros(max(O, 1), OTNODE_DTYPE)
    lib().orc_octree(*args, _p(nodes), O)
    return nodes[:O]


def octree_cells(codes, code_bits: int, root_level: int) -> np.ndarray:
    codes = _c(codes, np.uint64)
    n = lib().orc_octree_cells(_p(codes), len(codes), code_bits, root_level, None)
    cells = np.zeros((max(n, 1), 4), np.int64)
    lib().orc_octree_cells(_p(codes), len(codes), code_bits, root_level, _p(cells))
    return cells[:n]


def knn_brute(points, k: int, q_start: int = 0, n_q: int | None = None) -> np.ndarray:
    pts = _c(points, np.float32).reshape(-1, 3)
    U = len(pts)
    n_q = U - q_start if n_q is None else n_q
    out = np.empty((n_q, k), np.int32)
    lib().orc_knn_brute(_p(pts), U, q_start, n_q, k, _p(out))
    return out


def knn_brute_points(points, queries, k: int) -> np.ndarray:
    """Exact kNN of arbitrary query points (n x 3) among points, by brute force."""
    pts = _c(points, np.float32).reshape(-1, 3)
    q = _c(queries, np.float32).reshape(-1, 3)
    out = np.empty((len(q), k), np.int32)
    lib().orc_knn_brute_points(_p(pts), len(pts), _p(q), len(q), k, _p(out))
    return out


def knn_grid(points, k: int, queries) -> np.ndarray:
    pts = _c(points, np.float32).reshape(-1, 3)
    q = _c(queries, np.int64)
    out = np.empty((len(q), k), np.int32)
    lib().orc_knn_grid(_p(pts), len(pts), _p(q), len(q), k, _p(out))
    return out


def knn_reference(nodes, points, k: int, q_start: int = 0, n_q: int | None = None) -> np.ndarray:
    pts = _c(points, np.float32).reshape(-1, 3)
    nodes = np.ascontiguousarray(nodes, dtype=OTNODE_DTYPE)
    n_q = len(pts) - q_start if n_q is None else n_q
    out = np.empty((n_q, k), np.int32)
    lib().orc_knn_reference(_p(nodes), _p(pts), len(pts), q_start, n_q, k, _p(out))
    return out


def pca_f64(points, nn):
    pts = _c(points, np.float32).reshape(-1, 3)
    nn = _c(nn, np.int32)
    nq, k = nn.shape
    normals = np.empty((nq, 3), np.float64)
    evals = np.empty((nq, 3), np.float64)
    lib().orc_pca_f64(_p(pts), _p(nn), nq, k, _p(normals), _p(evals))
    return normals, evals


def jacobi3(A) -> tuple[np.ndarray, np.ndarray]:
    A = _c(A, np.float64).reshape(3, 3)
    w = np.empty(3, np.float64)
    V = np.empty((3, 3), np.float64)
    lib().orc_jacobi3(_p(A), _p(w), _p(V))
    return w, V


def normals_ref_f32(points, nn, q_start: int = 0, viewpoint=(0.0, 0.0, 0.0)) -> np.ndarray:
    pts = _c(points, np.float32).reshape(-1, 3)
    nn = _c(nn, np.int32)
    nq, k = nn.shape
    out = np.empty((nq, 3), np.float32)
    lib().orc_normals_ref_f32(_p(pts), _p(nn), q_start, nq, k, *map(float, viewpoint), _p(out))
    return out


def pipeline(x, y, z, code_bits: int, k: int, knn_mode: int = 0):
    """The reference pipeline on the host; returns (U, normals, stage_ms)."""
    x, y, z = _c(x, np.float32), _c(y, np.float32), _c(z, np.float32)
    n = len(x)
    normals = np.empty((n, 3), np.float32)
    t = np.zeros(8, np.float64)
    U = lib().orc_pipeline(_p(x), _p(y), _p(z), n, code_bits, k, knn_mode, _p(normals), _p(t))
    return int(U), normals[:U], t


def angular_error_deg(n_test, n_ref) -> np.ndarray:
    """Angle between normals up to sign, degrees."""
    a = np.asarray(n_test, np.float64)
    b = np.asarray(n_ref, np.float64)
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    b = b / np.linalg.norm(b, axis=1, keepdims=True)
    c = np.clip(np.abs(np.sum(a * b, axis=1)), 0.0, 1.0)
    return np.degrees(np.arccos(c))


def mlp_forward(params, feats, mu, sigma, emulate: bool = True) -> np.ndarray:
    P = _c(params, np.float32)
    f = _c(feats, np.float32).reshape(-1, 8)
    mu, sigma = _c(mu, np.float32), _c(sigma, np.float32)
    y = np.empty(len(f), np.float64)
    lib().orc_mlp_forward(_p(P), _p(f), len(f), _p(mu), _p(sigma), int(emulate), _p(y))
    return y


def mlp_grad(params, feats, targets, mu, sigma, emulate: bool = True):
    P = _c(params, np.float32)
    f = _c(feats, np.float32).reshape(-1, 8)
    t = _c(targets, np.float32)
    mu, sigma = _c(mu, np.float32), _c(sigma, np.float32)
    g = np.empty(int(lib().orc_mlp_num_params()), np.float64)
    loss = lib().orc_mlp_grad(_p(P), _p(f), _p(t), len(f), _p(mu), _p(sigma), int(emulate), _p(g))
    return g, float(loss)
