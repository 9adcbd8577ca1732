"""Python host layer over the C ABI: stage calls, the index object, workloads.

The names follow the reference's pipeline (RadixTree / Octree /
NormalEstimation, src/main.cu:80-107); the C++ drop-in facade under
``shapeextraction_amd/cpp`` is the reference-compatible API, this module is
the harness used by tests and the benchmark.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (SX_KNN_ALGO_LANE, SX_KNN_ALGO_TILES, SX_KNN_PLAIN, SX_KNN_RADIUS, SX_PREC_FP32, SX_PREC_FP64,
                   SX_ROOT_REFERENCE, SX_ROOT_TEXTBOOK, SX_SORT_ONESWEEP, SX_SORT_THREE_KERNEL, SX_SORT_HYBRID, call)
from .device import DeviceArray, fetch

OTNODE_DTYPE = np.dtype(
    [
        ("children", "<i4", (8,)),
        ("corner", "<f4", (3,)),
        ("cell_size", "<f4"),
        ("child_node_mask", "<i4"),
        ("child_leaf_mask", "<i4"),
    ]
)
assert OTNODE_DTYPE.itemsize == 56


def _dev(a, dtype=None) -> DeviceArray:
    if isinstance(a, DeviceArray):
        return a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return DeviceArray.from_numpy(arr)


def _ws(name: str, *args_before_ws, stream=None) -> DeviceArray:
    """Two-phase workspace query: returns a device buffer of the asked size."""
    size = C.c_size_t(0)
    call(name, *args_before_ws, None, C.byref(size), stream)
    return DeviceArray((max(int(size.value), 16),), np.uint8)


# ---------------------------------------------------------------------------
# Synthetic workloads (generated by the library's host code).
# ---------------------------------------------------------------------------
def _xyz(n):
    return np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.float32)


def gen_sphere(seed: int, n: int, sigma: float = 0.002):
    x, y, z = _xyz(n)
    call("sx_gen_sphere", seed, n, sigma, x.ctypes.data, y.ctypes.data, z.ctypes.data)
    return x, y, z


def gen_scene(seed: int, n: int, n_prims: int = 200, box: float = 100.0, noise: float = 0.01,
              outliers: float = 0.01):
    x, y, z = _xyz(n)
    call("sx_gen_scene", seed, n, n_prims, box, noise, outliers, x.ctypes.data, y.ctypes.data, z.ctypes.data)
    return x, y, z


def gen_gmm_scene(seed: int, n: int, n_prims: int = 200, n_components: int = 64, box: float = 100.0,
                  noise: float = 0.01, max_ratio: float = 1e4):
    x, y, z = _xyz(n)
    f = C.c_float(0)
    call("sx_gen_gmm_scene", seed, n, n_prims, n_components, box, noise, max_ratio, x.ctypes.data,
         y.ctypes.data, z.ctypes.data, C.byref(f))
    return x, y, z


def gen_batch_sizes(seed: int, n_clouds: int, lo: int = 1000, hi: int = 200000) -> np.ndarray:
    s = np.empty(n_clouds, np.int64)
    call("sx_gen_batch_sizes", seed, n_clouds, lo, hi, s.ctypes.data)
    return s


def gen_degenerate(seed: int, n: int, structs_per_class: int = 64, offset: float = 1e4, box: float = 100.0):
    x, y, z = _xyz(n)
    cls = np.empty(n, np.uint8)
    d = np.empty((n, 3), np.float32)
    call("sx_gen_degenerate", seed, n, structs_per_class, offset, box, x.ctypes.data, y.ctypes.data,
         z.ctypes.data, cls.ctypes.data, d.ctypes.data)
    return x, y, z, cls, d


def fnv64(arr: np.ndarray) -> int:
    arr = np.ascontiguousarray(arr)
    return int(_lib.load().sx_fnv64(arr.ctypes.data, arr.nbytes))


@dataclass(frozen=True)
class Workload:
    name: str
    n: int
    k: int
    code_bits: int
    seed: int
    kind: str

    def generate(self):
        if self.kind == "sphere":
            return gen_sphere(self.seed, self.n, 0.002)
        if self.kind == "scene":
            return gen_scene(self.seed, self.n)
        if self.kind == "gmm":
            return gen_gmm_scene(self.seed, self.n)
        if self.kind == "degenerate":
            return gen_degenerate(self.seed, self.n)[:3]
        raise ValueError(self.kind)


WORKLOADS = {
    "sphere100k": Workload("sphere100k", 100_000, 8, 30, 1, "sphere"),
    "scene8m": Workload("scene8m", 8_388_608, 16, 63, 2, "scene"),
    "gmm16m": Workload("gmm16m", 16_777_216, 32, 63, 3, "gmm"),
    "degenerate4m": Workload("degenerate4m", 4_194_304, 24, 63, 5, "degenerate"),
}


# ---------------------------------------------------------------------------
# Stage calls (one per reference kernel); inputs/outputs are DeviceArrays.
# ---------------------------------------------------------------------------
def bbox(x: DeviceArray, y: DeviceArray, z: DeviceArray, stream=None):
    n = len(x)
    bounds = DeviceArray((4,), np.float32)
    bad = DeviceArray.zeros((1,), np.int32)
    ws = _ws("sx_bbox", x.ptr, y.ptr, z.ptr, n, bounds.ptr, bad.ptr, stream=stream)
    call("sx_bbox", x.ptr, y.ptr, z.ptr, n, bounds.ptr, bad.ptr, ws.ptr, C.byref(C.c_size_t(ws.nbytes)), stream)
    return bounds, bad


def morton_encode(x, y, z, bounds: DeviceArray, code_bits: int, stream=None) -> DeviceArray:
    codes = DeviceArray((len(x),), np.uint64)
    call("sx_morton_encode", x.ptr, y.ptr, z.ptr, len(x), bounds.ptr, code_bits, codes.ptr, stream)
    return codes


def morton_decode(codes: DeviceArray, bounds: DeviceArray, code_bits: int, count: DeviceArray | None = None,
                  stream=None) -> DeviceArray:
    pts = DeviceArray((len(codes), 3), np.float32)
    call("sx_morton_decode", codes.ptr, count.ptr if count else None, len(codes), bounds.ptr, code_bits, pts.ptr,
         stream)
    return pts


def sort_pairs(keys: DeviceArray, key_bits: int, vals: DeviceArray | None = None, stream=None,
               check: bool = False):
    """Stable sort of (key, value) pairs.  check=True also waits for the sort
    and raises SxError(SX_EINTERNAL) if one of its tiles gave up waiting."""
    n = len(keys)
    ko = DeviceArray((n,), np.uint64)
    vo = DeviceArray((n,), np.uint32)
    vp = vals.ptr if vals is not None else None
    ws = _ws("sx_sort_pairs", keys.ptr, vp, n, key_bits, ko.ptr, vo.ptr, stream=stream)
    call("sx_sort_pairs", keys.ptr, vp, n, key_bits, ko.ptr, vo.ptr, ws.ptr, C.byref(C.c_size_t(ws.nbytes)), stream)
    if check:
        call("sx_sort_check", ws.ptr, n, stream)
    return ko, vo


def set_sort_spin_limit(spins: int) -> None:
    """Spins a sort tile waits for an earlier one before giving up (test hook)."""
    call("sx_sort_set_spin_limit", int(spins))


DEFAULT_SORT_SPIN_LIMIT = 1 << 22


def unique(sorted_keys: DeviceArray, vals: DeviceArray | None, stream=None):
    n = len(sorted_keys)
    uq = DeviceArray((n,), np.uint64)
    rep = DeviceArray((n,), np.uint32)
    uos = DeviceArray((n,), np.int32)
    cnt = DeviceArray.zeros((1,), np.int32)
    vp = vals.ptr if vals is not None else None
    ws = _ws("sx_unique", sorted_keys.ptr, vp, n, uq.ptr, rep.ptr, uos.ptr, cnt.ptr, stream=stream)
    call("sx_unique", sorted_keys.ptr, vp, n, uq.ptr, rep.ptr, uos.ptr, cnt.ptr, ws.ptr,
         C.byref(C.c_size_t(ws.nbytes)), stream)
    return uq, rep, uos, cnt


def radix_tree(codes: DeviceArray, count: DeviceArray, code_bits: int, root_mode: int = SX_ROOT_TEXTBOOK,
               stream=None) -> dict:
    n = len(codes)
    out = {
        "prefix_n": DeviceArray((n,), np.uint8),
        "left_child": DeviceArray((n,), np.int32),
        "has_leaf_left": DeviceArray((n,), np.uint8),
        "has_leaf_right": DeviceArray((n,), np.uint8),
        "parent": DeviceArray((n,), np.int32),
        "node_end": DeviceArray((n,), np.int32),
        "leaf_parent": DeviceArray((n,), np.int32),
    }
    call("sx_build_radix_tree", codes.ptr, count.ptr, n, code_bits, root_mode, out["prefix_n"].ptr,
         out["left_child"].ptr, out["has_leaf_left"].ptr, out["has_leaf_right"].ptr, out["parent"].ptr,
         out["node_end"].ptr, out["leaf_parent"].ptr, stream)
    return out


def octree(codes: DeviceArray, count: DeviceArray, code_bits: int, bounds: DeviceArray, rt: dict,
           capacity: int | None = None, stream=None):
    n = len(codes)
    cap = capacity if capacity is not None else max(16, 3 * n)
    nodes = DeviceArray((cap,), OTNODE_DTYPE)
    ocnt = DeviceArray.zeros((1,), np.int32)
    args = (codes.ptr, count.ptr, n, code_bits, bounds.ptr, rt["prefix_n"].ptr, rt["left_child"].ptr,
            rt["has_leaf_left"].ptr, rt["has_leaf_right"].ptr, rt["parent"].ptr, nodes.ptr, cap, ocnt.ptr)
    ws = _ws("sx_build_octree", *args, stream=stream)
    call("sx_build_octree", *args, ws.ptr, C.byref(C.c_size_t(ws.nbytes)), stream)
    return nodes, ocnt


def check_counts(count: DeviceArray | None, cap: int, oct_count: DeviceArray | None = None, oct_cap: int = 0,
                 bad: DeviceArray | None = None, stream=None):
    u = C.c_int64(0)
    o = C.c_int64(0)
    call("sx_check_counts", count.ptr if count else None, cap, oct_count.ptr if oct_count else None, oct_cap,
         bad.ptr if bad else None, C.byref(u), C.byref(o), stream)
    return u.value, o.value


# ---------------------------------------------------------------------------
# Index object.
# ---------------------------------------------------------------------------
class Index:
    """One cloud's index (RT::RadixTree + OT::Octree state) on the GPU."""

    def __init__(self, capacity: int, code_bits: int = 63, root_mode: int = SX_ROOT_TEXTBOOK,
                 oct_factor: float = 3.0):
        h = C.c_void_p()
        call("sx_index_create", C.byref(h), int(capacity), code_bits, root_mode, oct_factor)
        self.handle = h.value
        self.capacity = int(capacity)
        self.code_bits = code_bits
        self._lib = _lib.load()
        self._view = None

    def close(self):
        if getattr(self, "handle", None):
            self._lib.sx_index_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bucket(self, bucket: int):
        call("sx_index_set_bucket", self.handle, bucket)

    def set_algorithms(self, knn: int = _lib.SX_KNN_ALGO_TILES, sort: int = _lib.SX_SORT_HYBRID):
        call("sx_index_set_algorithms", self.handle, knn, sort)

    def build(self, x, y, z, stream=None) -> "Index":
        dx, dy, dz = _dev(x, np.float32), _dev(y, np.float32), _dev(z, np.float32)
        self._inputs = (dx, dy, dz)  # keep alive until the stream has consumed them
        call("sx_index_build", self.handle, dx.ptr, dy.ptr, dz.ptr, len(dx), stream)
        self._view = None
        return self

    def build_codes(self, codes, min_coord: float, range_: float, stream=None) -> "Index":
        """Index of strictly increasing Morton codes under the cube
        [min_coord, min_coord + range_] (no sort, no duplicate removal)."""
        dc = _dev(codes, np.uint64)
        self._inputs = (dc,)
        call("sx_index_build_codes", self.handle, dc.ptr, len(dc), float(min_coord), float(range_), stream)
        self._view = None
        return self

    def to_input_order(self, rows: DeviceArray, stream=None) -> DeviceArray:
        """Rows per unique point (normals, neighbour rows) gathered into input order."""
        v = self._view or self.view(stream)
        row_shape = rows.shape[1:]
        out = DeviceArray((int(v.n_input),) + tuple(row_shape), rows.dtype)
        row_bytes = rows.nbytes // max(rows.shape[0], 1)
        call("sx_index_to_input_order", self.handle, rows.ptr, row_bytes, out.ptr, stream)
        return out

    def knn_points(self, queries, k: int, out: DeviceArray | None = None, stream=None) -> DeviceArray:
        """Exact k nearest index points of arbitrary query points (n x 3 float32)."""
        if isinstance(queries, DeviceArray):
            dq = queries
        else:
            dq = DeviceArray.from_numpy(np.ascontiguousarray(queries, np.float32).reshape(-1, 3))
        n_q = dq.shape[0]
        out = out if out is not None else DeviceArray((n_q, k), np.int32)
        self._queries = dq
        call("sx_knn_points", self.handle, dq.ptr, n_q, k, out.ptr, stream)
        return out

    def view(self, stream=None) -> _lib.IndexView:
        v = _lib.IndexView()
        call("sx_index_get_view", self.handle, C.byref(v), stream)
        self._view = v
        return v

    @property
    def n_unique(self) -> int:
        return int((self._view or self.view()).n_unique)

    def fetch(self) -> dict:
        """Copy every index array to the host (tests)."""
        v = self._view or self.view()
        U, O, N = int(v.n_unique), int(v.n_oct), int(v.n_input)
        return {
            "n_input": N,
            "n_unique": U,
            "n_oct": O,
            "min_coord": float(v.min_coord),
            "max_coord": float(v.max_coord),
            "codes": fetch(v.codes, U, np.uint64),
            "prefix_n": fetch(v.prefix_n, U, np.uint8),
            "left_child": fetch(v.left_child, U, np.int32),
            "has_leaf_left": fetch(v.has_leaf_left, U, np.uint8),
            "has_leaf_right": fetch(v.has_leaf_right, U, np.uint8),
            "parent": fetch(v.parent, U, np.int32),
            "oct_nodes": fetch(v.oct_nodes, O, OTNODE_DTYPE),
            "points": fetch(v.points, 3 * U, np.float32).reshape(U, 3),
            "sorted_perm": fetch(v.sorted_perm, N, np.uint32),
            "rep": fetch(v.rep, U, np.uint32),
            "unique_of_input": fetch(v.unique_of_input, N, np.int32),
            "bounds": fetch(v.d_bounds, 3, np.float32),
        }

    def knn(self, k: int, q_start: int = 0, n_q: int | None = None, mode: int = SX_KNN_PLAIN,
            radius2: DeviceArray | None = None, out: DeviceArray | None = None, stream=None) -> DeviceArray:
        if n_q is None:
            n_q = self.n_unique - q_start
        out = out if out is not None else DeviceArray((n_q, k), np.int32)
        call("sx_knn", self.handle, q_start, n_q, k, mode, radius2.ptr if radius2 is not None else None, out.ptr,
             stream)
        return out

    def normals(self, nn: DeviceArray, k: int, q_start: int = 0, n_q: int | None = None,
                precision: int = SX_PREC_FP32, viewpoint=(0.0, 0.0, 0.0), flags: bool = False, stream=None):
        if n_q is None:
            n_q = self.n_unique - q_start
        out = DeviceArray((n_q, 3), np.float32)
        fl = DeviceArray((max(n_q, 1),), np.uint8) if flags else None
        call("sx_normals", self.handle, nn.ptr, q_start, n_q, k, precision, float(viewpoint[0]),
             float(viewpoint[1]), float(viewpoint[2]), out.ptr, fl.ptr if fl else None, stream)
        return (out, fl) if flags else out

    def estimate_normals(self, k: int, q_start: int = 0, n_q: int | None = None, mode: int = SX_KNN_PLAIN,
                         radius2: DeviceArray | None = None, precision: int = SX_PREC_FP32,
                         nn: DeviceArray | None = None, out: DeviceArray | None = None, stream=None):
        if n_q is None:
            n_q = self.n_unique - q_start
        out = out if out is not None else DeviceArray((n_q, 3), np.float32)
        call("sx_estimate_normals", self.handle, q_start, n_q, k, mode,
             radius2.ptr if radius2 is not None else None, precision, nn.ptr if nn is not None else None, out.ptr,
             stream)
        return out

    def set_knn_stats(self, enable: bool = True):
        call("sx_index_set_knn_stats", self.handle, 1 if enable else 0)

    def knn_candidates(self, reset: bool = True, stream=None) -> int:
        """Candidates the kNN tiles broadcast since the last reset (stats enabled)."""
        r = C.c_int64(0)
        call("sx_index_knn_candidates", self.handle, C.byref(r), 1 if reset else 0, stream)
        return r.value

    def knn_reruns(self, reset: bool = True, stream=None) -> int:
        r = C.c_int64(0)
        call("sx_index_knn_reruns", self.handle, C.byref(r), 1 if reset else 0, stream)
        return r.value


class Batch:
    """Many clouds in one index (sx_batch): cloud c is input points
    [offsets[c], offsets[c+1]).  Per-cloud results equal a standalone Index
    of that cloud; per-slot arrays hold cloud c at [base[c], base[c]+size[c])."""

    def __init__(self, max_points: int, max_clouds: int, code_bits: int = 63):
        h = C.c_void_p()
        call("sx_batch_create", C.byref(h), int(max_points), int(max_clouds), code_bits)
        self.handle = h.value
        self._lib = _lib.load()
        self._view = None

    def close(self):
        if getattr(self, "handle", None):
            self._lib.sx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_clouds(self, offsets, stream=None) -> "Batch":
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        call("sx_batch_set_clouds", self.handle, off.ctypes.data_as(C.POINTER(C.c_int64)), len(off) - 1, stream)
        self.offsets = off
        self._view = None
        return self

    def set_tuning(self, bucket: int = 16, sort: int = _lib.SX_SORT_HYBRID, queue_blocks: int = 0):
        call("sx_batch_set_tuning", self.handle, bucket, sort, queue_blocks)

    def set_octree(self, oct_factor: float = 3.0) -> "Batch":
        """Build every cloud's octree from the next build on (0: stop, free it)."""
        call("sx_batch_set_octree", self.handle, float(oct_factor))
        self._view = None
        return self

    def build(self, x, y, z, stream=None) -> "Batch":
        dx, dy, dz = _dev(x, np.float32), _dev(y, np.float32), _dev(z, np.float32)
        self._inputs = (dx, dy, dz)
        call("sx_batch_build", self.handle, dx.ptr, dy.ptr, dz.ptr, stream)
        self._view = None
        return self

    def view(self, stream=None) -> _lib.BatchView:
        v = _lib.BatchView()
        call("sx_batch_get_view", self.handle, C.byref(v), stream)
        self._view = v
        return v

    def table(self, stream=None) -> dict:
        """Per-cloud base slot, unique count, bounds (host copies)."""
        v = self._view or self.view(stream)
        n = int(v.n_clouds)
        return {
            "n_slots": int(v.n_slots),
            "n_unique": int(v.n_unique),
            "base": fetch(v.cloud_base, n, np.int32),
            "size": fetch(v.cloud_size, n, np.int32),
            "bounds": fetch(v.cloud_bounds, 4 * n, np.float32).reshape(n, 4),
        }

    def fetch(self) -> dict:
        """Copy every per-slot array to the host (tests)."""
        v = self._view or self.view()
        S = int(v.n_slots)
        out = self.table()
        out.update({
            "codes": fetch(v.codes, S, np.uint64),
            "rep": fetch(v.rep, S, np.uint32),
            "points": fetch(v.points, 4 * S, np.float32).reshape(S, 4)[:, :3].copy(),
            "prefix_n": fetch(v.prefix_n, S, np.uint8),
            "left_child": fetch(v.left_child, S, np.int32),
            "parent": fetch(v.parent, S, np.int32),
        })
        if v.oct_nodes:
            n, O = int(v.n_clouds), int(v.n_oct)
            out.update({
                "n_oct": O,
                "oct_nodes": fetch(v.oct_nodes, O, OTNODE_DTYPE),
                "oct_base": fetch(v.cloud_oct_base, n, np.int32),
                "oct_size": fetch(v.cloud_oct_size, n, np.int32),
                "has_leaf_left": fetch(v.has_leaf_left, S, np.uint8),
                "has_leaf_right": fetch(v.has_leaf_right, S, np.uint8),
            })
        return out

    def knn(self, k: int, out: DeviceArray | None = None, stream=None) -> DeviceArray:
        v = self._view or self.view(stream)
        if out is None:
            out = DeviceArray((int(v.n_slots), k), np.int32)
        call("sx_batch_knn", self.handle, k, out.ptr, out.shape[0], stream)
        return out

    def knn_radius(self, k: int, radius2: DeviceArray, out: DeviceArray | None = None, stream=None) -> DeviceArray:
        """kNN seeded by per-slot squared radii (same rows as knn)."""
        v = self._view or self.view(stream)
        if out is None:
            out = DeviceArray((int(v.n_slots), k), np.int32)
        rows = min(out.shape[0], radius2.shape[0])
        call("sx_batch_knn_radius", self.handle, k, radius2.ptr, out.ptr, rows, stream)
        return out

    def radius_features(self, out: DeviceArray | None = None, stream=None) -> DeviceArray:
        """Learned-radius features of every slot (needs set_octree before the build)."""
        v = self._view or self.view(stream)
        if out is None:
            out = DeviceArray((int(v.n_slots), 8), np.float32)
        call("sx_batch_radius_features", self.handle, out.ptr, out.shape[0], stream)
        return out

    def radius_predict(self, model, out: DeviceArray | None = None, stream=None) -> DeviceArray:
        """Squared search radius of every slot from a RadiusModel (features + MLP fused)."""
        v = self._view or self.view(stream)
        if out is None:
            out = DeviceArray((int(v.n_slots),), np.float32)
        call("sx_batch_radius_predict", self.handle, model.handle, out.ptr, out.shape[0], stream)
        return out

    def normals(self, nn: DeviceArray, k: int, precision: int = SX_PREC_FP32, viewpoint=(0.0, 0.0, 0.0),
                out: DeviceArray | None = None, flags: DeviceArray | None = None, stream=None) -> DeviceArray:
        v = self._view or self.view(stream)
        if out is None:
            out = DeviceArray((int(v.n_slots), 3), np.float32)
        rows = min(out.shape[0], nn.shape[0])
        call("sx_batch_normals", self.handle, nn.ptr, k, precision, *viewpoint, out.ptr,
             flags.ptr if flags is not None else None, rows, stream)
        return out

    def deferred(self, reset: bool = True, stream=None) -> int:
        r = C.c_int64(0)
        call("sx_batch_deferred", self.handle, C.byref(r), 1 if reset else 0, stream)
        return r.value

    def graph(self, x: DeviceArray, y: DeviceArray, z: DeviceArray, k: int, nn: DeviceArray, normals: DeviceArray,
              stream) -> "BatchGraph":
        g = C.c_void_p()
        rows = min(nn.shape[0], normals.shape[0])
        call("sx_batch_graph_create", self.handle, x.ptr, y.ptr, z.ptr, k, nn.ptr, normals.ptr, rows, C.byref(g),
             stream)
        self._view = None
        return BatchGraph(g.value, (self, x, y, z, nn, normals))  # the batch outlives its graph

    def autotune(self, k: int, scratch: DeviceArray, buckets=(8, 16, 32), queue_blocks=(0, 2048), reps: int = 2,
                 stream=None) -> dict:
        b = (C.c_int * len(buckets))(*buckets)
        q = (C.c_int * len(queue_blocks))(*queue_blocks)
        ms = (C.c_float * (len(buckets) * len(queue_blocks)))()
        bb, bq = C.c_int(0), C.c_int(0)
        call("sx_batch_autotune", self.handle, k, scratch.ptr, scratch.shape[0], b, len(buckets), q,
             len(queue_blocks), reps, ms, C.byref(bb), C.byref(bq), stream)
        return {"bucket": bb.value, "queue_blocks": bq.value,
                "ms": np.array(list(ms), np.float32).reshape(len(buckets), len(queue_blocks))}


class BatchGraph:
    """A captured batch pipeline (build + kNN + FP32 normals)."""

    def __init__(self, handle, keep):
        self.handle = handle
        self._keep = keep
        self._lib = _lib.load()

    def launch(self, stream):
        call("sx_batch_graph_launch", self.handle, stream)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.sx_batch_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = [
    "Index", "Batch", "BatchGraph", "Workload", "WORKLOADS", "OTNODE_DTYPE", "bbox", "morton_encode", "morton_decode", "sort_pairs",
    "unique", "radix_tree", "octree", "check_counts", "set_sort_spin_limit", "DEFAULT_SORT_SPIN_LIMIT", "gen_sphere", "gen_scene", "gen_gmm_scene",
    "gen_batch_sizes", "gen_degenerate", "fnv64", "SX_KNN_PLAIN", "SX_KNN_RADIUS", "SX_PREC_FP32", "SX_PREC_FP64",
    "SX_ROOT_TEXTBOOK", "SX_ROOT_REFERENCE", "SX_KNN_ALGO_TILES", "SX_KNN_ALGO_LANE", "SX_SORT_ONESWEEP", "SX_SORT_HYBRID",
    "SX_SORT_THREE_KERNEL",
]
