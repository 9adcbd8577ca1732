"""Learned kNN search radius: data, training and inference driver.

The network (8-64-64-1, BF16 weights, FP32 accumulation) and its forward,
backward and Adam kernels live in csrc/mlp.hip; this module only moves data
and calls them through the C ABI.  Training labels are exact: ln of the
squared distance to the k-th neighbour from the plain exact kNN.

Weight file (little endian): magic b"SXRADIUS", int32 version=1, int32 k,
int32 n_params, float32 params[n_params] (the FP32 masters; inference rounds
them to BF16), float32 mu[8], float32 sigma[8], float32 margin.
"""
from __future__ import annotations

import argparse
import ctypes as C
import struct
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import SX_KNN_RADIUS, call
from .api import Index, gen_scene, gen_sphere
from .device import DeviceArray

N_FEATURES = 8
MODEL_DIR = Path(__file__).resolve().parent / "models"
MAGIC = b"SXRADIUS"



class RadiusModel:
    """Device-resident MLP with its normaliser and calibrated log-margin."""

    def __init__(self, k: int = 16, max_batch: int = 65536):
        h = C.c_void_p()
        call("sx_mlp_create", C.byref(h), max_batch)
        self.handle = h.value
        self.k = k
        self.max_batch = max_batch
        self.mu = np.zeros(N_FEATURES, np.float32)
        self.sigma = np.ones(N_FEATURES, np.float32)
        self.margin = 0.0
        self._lib = _lib.load()
        self.n_params = int(self._lib.sx_mlp_num_params())
        self.set_normalizer(self.mu, self.sigma, 0.0)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.sx_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters -------------------------------------------------------
    def init(self, seed: int = 0):
        call("sx_mlp_init", self.handle, seed, None)

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.n_params,)
        call("sx_mlp_set_params", self.handle, p.ctypes.data, None)

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        call("sx_mlp_get_params", self.handle, p.ctypes.data, None)
        return p

    def set_normalizer(self, mu, sigma, margin: float):
        self.mu = np.ascontiguousarray(mu, np.float32)
        self.sigma = np.ascontiguousarray(sigma, np.float32)
        self.margin = float(margin)
        call("sx_mlp_set_normalizer", self.handle, self.mu.ctypes.data, self.sigma.ctypes.data, self.margin, None)

    # compute ------------------------------------------------------------
    def forward(self, feats: DeviceArray, n: int | None = None, out: DeviceArray | None = None) -> DeviceArray:
        n = len(feats) if n is None else n
        out = out if out is not None else DeviceArray((n,), np.float32)
        call("sx_mlp_forward", self.handle, feats.ptr, n, out.ptr, None)
        return out

    def predict_radius2(self, feats: DeviceArray, n: int | None = None, out: DeviceArray | None = None,
                        stream=None) -> DeviceArray:
        n = len(feats) if n is None else n
        out = out if out is not None else DeviceArray((n,), np.float32)
        call("sx_mlp_predict_radius2", self.handle, feats.ptr, n, out.ptr, stream)
        return out

    def predict_index_radius2(self, index, q_start: int = 0, n_q: int | None = None,
                              out: DeviceArray | None = None, stream=None) -> DeviceArray:
        """Features and MLP in one kernel for index points [q_start, q_start + n_q)."""
        n_q = index.n_unique - q_start if n_q is None else n_q
        out = out if out is not None else DeviceArray((n_q,), np.float32)
        call("sx_radius_predict", index.handle, self.handle, q_start, n_q, out.ptr, stream)
        return out

    def backward(self, feats: DeviceArray, targets: DeviceArray, n: int | None = None):
        n = len(feats) if n is None else n
        g = DeviceArray((self.n_params,), np.float32)
        loss = DeviceArray((1,), np.float32)
        call("sx_mlp_backward", self.handle, feats.ptr, targets.ptr, n, g.ptr, loss.ptr, None)
        return g.numpy(), float(loss.numpy()[0]) / n

    def adam_step(self, feats_ptr: int, targets_ptr: int, n: int, lr: float, wd: float = 0.0, loss_ptr=None):
        call("sx_mlp_adam_step", self.handle, feats_ptr, targets_ptr, n, lr, wd, loss_ptr, None)

    # persistence ----------------------------------------------------------
    def save(self, path: Path):
        p = self.params()
        with open(path, "wb") as f:
            f.write(MAGIC)
            f.write(struct.pack("<iii", 1, self.k, self.n_params))
            f.write(p.astype("<f4").tobytes())
            f.write(self.mu.astype("<f4").tobytes())
            f.write(self.sigma.astype("<f4").tobytes())
            f.write(struct.pack("<f", self.margin))

    @classmethod
    def load(cls, path: Path, max_batch: int = 65536) -> "RadiusModel":
        raw = Path(path).read_bytes()
        if raw[:8] != MAGIC:
            raise ValueError(f"{path}: not a radius model file")
        ver, k, n_params = struct.unpack_from("<iii", raw, 8)
        if ver != 1:
            raise ValueError(f"{path}: unsupported version {ver}")
        off = 20
        p = np.frombuffer(raw, "<f4", n_params, off).astype(np.float32)
        off += 4 * n_params
        mu = np.frombuffer(raw, "<f4", 8, off).astype(np.float32)
        sigma = np.frombuffer(raw, "<f4", 8, off + 32).astype(np.float32)
        (margin,) = struct.unpack_from("<f", raw, off + 64)
        m = cls(k, max_batch)
        m.set_params(p)
        m.set_normalizer(mu, sigma, margin)
        return m


def default_model_path(k: int) -> Path:
    return MODEL_DIR / f"radius_k{k}.bin"


def features(index: Index, q_start: int = 0, n_q: int | None = None, out: DeviceArray | None = None,
             stream=None) -> DeviceArray:
    n_q = index.n_unique - q_start if n_q is None else n_q
    out = out if out is not None else DeviceArray((n_q, N_FEATURES), np.float32)
    call("sx_radius_features", index.handle, q_start, n_q, out.ptr, stream)
    return out


def kth_sq_distance(points: np.ndarray, nn: np.ndarray) -> np.ndarray:
    """Squared distance to the last (k-th) neighbour, float32 like the kernel."""
    d = points[nn[:, -1]].astype(np.float32) - points.astype(np.float32)
    return ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)


def training_clouds(n_clouds: int = 32, seed0: int = 100):
    """Sphere-surface clouds in the style of the 100k sphere workload (seeds
    100..), with log-uniform sizes 10k..1M, radius 0.5..50 and noise 0.1-0.5%
    of the radius, plus every fourth cloud a primitive scene (other seeds than
    the benchmark's) so that the features span planar and curved surfaces."""
    g = np.random.default_rng(seed0)
    for i in range(n_clouds):
        n = int(np.exp(g.uniform(np.log(1e4), np.log(1e6))))
        if i % 4 == 3:
            yield f"scene(seed={seed0 + i}, n={n})", gen_scene(seed0 + i, n)
        else:
            rad = float(np.exp(g.uniform(np.log(0.5), np.log(50.0))))
            noise = float(g.uniform(0.001, 0.005))
            x, y, z = gen_sphere(seed0 + i, n, noise)
            c = g.uniform(-100, 100, 3)
            yield (f"sphere(seed={seed0 + i}, n={n}, r={rad:.2f})",
                   ((x * rad + c[0]).astype(np.float32), (y * rad + c[1]).astype(np.float32),
                    (z * rad + c[2]).astype(np.float32)))


def build_dataset(k: int, n_clouds: int = 32, per_cloud: int = 200_000, seed0: int = 100, log=print):
    feats, targets = [], []
    g = np.random.default_rng(seed0 + 999)
    for name, (x, y, z) in training_clouds(n_clouds, seed0):
        ix = Index(len(x), 63).build(x, y, z)
        U = ix.n_unique
        nn = ix.knn(k).numpy()
        pts = ix.fetch()["points"]
        kth = kth_sq_distance(pts, nn)
        f = features(ix).numpy()
        sel = g.choice(U, min(U, per_cloud), replace=False)
        ok = kth[sel] > 0
        feats.append(f[sel][ok])
        targets.append(np.log(kth[sel][ok]).astype(np.float32))
        log(f"[data] {name}: U={U}, kept {ok.sum()}")
        ix.close()
    return np.concatenate(feats), np.concatenate(targets)


def train(k: int, steps: int = 4000, batch: int = 8192, lr: float = 2e-3, seed: int = 0, n_clouds: int = 32,
          quantile: float = 0.9, log=print) -> RadiusModel:
    t0 = time.time()
    X, T = build_dataset(k, n_clouds, log=log)
    g = np.random.default_rng(seed)
    perm = g.permutation(len(X))
    X, T = X[perm], T[perm]
    n_val = min(200_000, len(X) // 10)
    Xv, Tv, Xt, Tt = X[:n_val], T[:n_val], X[n_val:], T[n_val:]
    mu = Xt.mean(0).astype(np.float32)
    sigma = np.maximum(Xt.std(0), 1e-3).astype(np.float32)
    model = RadiusModel(k, max_batch=batch)
    model.init(seed)
    model.set_normalizer(mu, sigma, 0.0)
    dX, dT = DeviceArray.from_numpy(Xt), DeviceArray.from_numpy(Tt)
    loss = DeviceArray((1,), np.float32)
    n_tr = len(Xt)
    log(f"[train] {n_tr} samples, {n_val} validation; data {time.time() - t0:.1f} s")
    for s in range(steps):
        off = (s * batch) % max(1, n_tr - batch)
        cur_lr = lr * 0.5 * (1 + np.cos(np.pi * s / steps))
        model.adam_step(dX.offset(off * N_FEATURES), dT.offset(off), batch, cur_lr, 0.0, loss.ptr)
        if s % 500 == 0 or s == steps - 1:
            log(f"[train] step {s}: batch mse {float(loss.numpy()[0]) / batch:.4f}")
    yv = model.forward(DeviceArray.from_numpy(Xv)).numpy()
    res = Tv - yv
    margin = float(np.quantile(res, quantile))
    model.set_normalizer(mu, sigma, margin)
    log(f"[train] validation rmse {np.sqrt(np.mean(res ** 2)):.4f} (ln r^2), margin q{quantile}: {margin:.4f}; "
        f"total {time.time() - t0:.1f} s")
    return model


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the learned kNN radius on the GPU")
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--clouds", type=int, default=32)
    ap.add_argument("--out", type=Path, default=None)
    a = ap.parse_args(argv)
    model = train(a.k, steps=a.steps, n_clouds=a.clouds)
    out = a.out or default_model_path(a.k)
    out.parent.mkdir(parents=True, exist_ok=True)
    model.save(out)
    print(f"saved {out}")


if __name__ == "__main__":
    sys.exit(main())
