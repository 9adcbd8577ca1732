"""End-to-end normal-estimation benchmark on one MI355X.

One step = the whole hot path on one synthetic cloud already resident in HBM:
cube bounds -> Morton encode -> radix sort -> duplicate removal -> decode ->
Karras radix tree + boxes -> octree -> exact kNN (k=16) -> normals, i.e. the
reference's RadixTree + Octree + NormalEstimation::estimateNormals<k>
(src/main.cu:80-107) on a 8,388,608-point, 200-primitive scene.

Prints ONE JSON line (rank 0, N=1).  `value` = input points / wall time of one
step (device-resident input and outputs).  `roofline` refers to the dominant
stage (kNN: the tile kernel plus the deferred-query kernel), timed with HIP
events around that call on the stream it runs on; `cpu_baseline`
times the CPU restatement of the reference pipeline (oracle/) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "points/sec end-to-end normals at 8M pts k=16; learned vs plain kNN at ≤1° p99"
HBM_PEAK_GBS = 8000.0


def knn_bytes_per_query(k: int) -> int:
    """Algorithmic HBM bytes per query of the kNN kernel: its own point (16 B
    float4), the traversal record of one internal node (64 B; the tree is read
    once in aggregate, U-1 records for U queries) and the k output indices."""
    return 16 + 64 + 4 * k


# Passes over the (key, value) pairs of the 63-bit sort: four onesweep passes
# on the top digits and one pass that sorts every bucket of equal top bits.
SORT_PAIR_PASSES = 5


def pipeline_bytes_per_point(key_bytes: int, passes: int, k: int, r: float) -> float:
    """Byte model of the whole path per input point (137 + 8K + 2P(K+4) + 112r + 8k)."""
    return 137 + 8 * key_bytes + 2 * passes * (key_bytes + 4) + 112 * r + 8 * k


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, workload_n: int, k: int) -> dict:
    """The oracle's restatement of the reference pipeline on a bounded sample
    of the same workload (same generator and seed), on the host cores: one
    untimed warm-up on a small cloud, then the median of three timed runs."""
    from oracle import oracle as O

    import shapeextraction_amd as sx

    n = min(int(os.environ.get("SX_CPU_SAMPLE", 2_097_152)), workload_n)  # ~5 s per run on 16 host threads
    runs = int(os.environ.get("SX_CPU_RUNS", 3))
    x, y, z = sx.Workload(w.name, n, w.k, w.code_bits, w.seed, w.kind).generate()
    O.pipeline(x[:65536].copy(), y[:65536].copy(), z[:65536].copy(), w.code_bits, k, knn_mode=0)
    times, stages = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        U, _, stage = O.pipeline(x, y, z, w.code_bits, k, knn_mode=0)
        times.append(time.perf_counter() - t0)
        stages.append(stage)
    order = np.argsort(times)
    mid = int(order[len(order) // 2])
    dt = times[mid]
    names = ["bounds+encode", "sort", "unique", "radix tree", "octree", "decode", "kNN (32-slot queue)", "normals"]
    return {
        "value": n / dt,
        "unit": "points/s",
        "cores": O.threads(),
        "kind": "port",
        "sample": f"{n}-point {w.kind} cloud from the same generator and seed as the {workload_n}-point "
                  f"workload (1/{workload_n // n} of it; every stage is linear or N log N in the points, the kNN "
                  f"{100 * stages[mid][6] / (1000 * dt):.0f}% of the time and linear); reference pipeline restated "
                  f"on the host (32-slot queue kNN, FP32 Jacobi); median of {runs} runs "
                  f"({', '.join(f'{t:.2f}' for t in times)} s)",
        "stage_ms": {nm: round(float(v), 1) for nm, v in zip(names, stages[mid])},
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--n", type=int, default=8_388_608)
    ap.add_argument("--bucket", type=int, default=32)
    ap.add_argument("--knn", choices=["plain", "learned", "both"], default="both",
                    help="learned/both use shapeextraction_amd/models/radius_k{k}.bin (trained here if missing)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="scene8m",
                    help="scene8m (default, the headline), gmm16m, degenerate4m or sphere100k")
    ap.add_argument("--batched", action="store_true",
                    help="time the batched workload (512 scenes, sizes log-uniform in [1k, 200k]) instead")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run three device-to-device copies of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import shapeextraction_amd as sx
    from shapeextraction_amd import learned

    if sx.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        return 1
    dev = sx.device_name(0)
    log(f"device {dev}")
    if args.copy_probe:
        nb = args.copy_probe << 20
        src, dst = sx.DeviceArray((nb,), np.uint8), sx.DeviceArray((nb,), np.uint8)
        src.fill_(1)
        for _ in range(3):
            sx._lib.call("sx_memcpy_dtod", dst.ptr, src.ptr, nb, None)
        sx.synchronize()
        log(f"copied {nb} bytes x3")
        return 0
    if args.batched:
        return bench_batched(sx, args, dev)
    if args.workload not in sx.WORKLOADS:
        print(f"unknown workload {args.workload}; choose from {sorted(sx.WORKLOADS)}", file=sys.stderr)
        return 2
    w = sx.WORKLOADS[args.workload]
    headline = w.name == "scene8m"
    n, k, bits = (args.n, args.k, 63) if headline else (w.n, w.k, w.code_bits)
    t = time.perf_counter()
    cls = dirs = None
    if headline:
        x, y, z = sx.gen_scene(2, n)
    elif w.kind == "degenerate":
        x, y, z, cls, dirs = sx.gen_degenerate(w.seed, n)
    else:
        x, y, z = w.generate()
    log(f"generated {n} points in {time.perf_counter() - t:.1f} s")
    st = sx.Stream()
    dx, dy, dz = (sx.DeviceArray.from_numpy(a) for a in (x, y, z))
    ix = sx.Index(n, bits)
    ix.set_bucket(args.bucket)

    # first build to learn U (the only host read of a count; untimed)
    ix.build(dx, dy, dz, stream=st.handle)
    v = ix.view(stream=st.handle)
    U, O_nodes = int(v.n_unique), int(v.n_oct)
    log(f"U={U} octree nodes={O_nodes} r={O_nodes / U:.3f}")
    nn = sx.DeviceArray((U, k), np.int32)
    normals = sx.DeviceArray((U, 3), np.float32)
    radius2 = sx.DeviceArray((U,), np.float32)

    model = None
    model_path = learned.default_model_path(k)
    if args.knn in ("learned", "both") and not headline and not model_path.exists():
        log(f"no trained radius model for k={k}: plain kNN only")
    elif args.knn in ("learned", "both"):
        if not model_path.exists():
            log(f"no trained radius model at {model_path}; training one on this GPU")
            model = learned.train(k, log=log)
        else:
            model = learned.RadiusModel.load(model_path)

    def step(mode, events=None):
        """One full pass; events (5) bracket build | features+MLP | kNN | normals."""
        rec = (lambda i: events[i].record(st.handle)) if events else (lambda i: None)
        rec(0)
        sx._lib.call("sx_index_build", ix.handle, dx.ptr, dy.ptr, dz.ptr, n, st.handle)
        rec(1)
        if mode == "learned":
            model.predict_index_radius2(ix, 0, U, radius2, stream=st.handle)
        rec(2)
        sx._lib.call("sx_knn", ix.handle, 0, U, k, sx.SX_KNN_RADIUS if mode == "learned" else sx.SX_KNN_PLAIN,
                     radius2.ptr if mode == "learned" else None, nn.ptr, st.handle)
        rec(3)
        sx._lib.call("sx_normals", ix.handle, nn.ptr, 0, U, k, sx.SX_PREC_FP32, 0.0, 0.0, 0.0, normals.ptr, None,
                     st.handle)
        rec(4)

    def measure(mode):
        ev = [[sx.Event() for _ in range(5)] for _ in range(args.steps)]
        for _ in range(args.warmup):
            step(mode)
        st.sync()
        ix.knn_reruns(reset=True, stream=st.handle)
        t0 = time.perf_counter()
        for s in range(args.steps):
            step(mode, ev[s])
        st.sync()
        t1 = time.perf_counter()
        el = lambda i, j: np.array([e[i].elapsed_ms(e[j]) for e in ev])  # noqa: E731
        v2 = ix.view(stream=st.handle)
        assert int(v2.n_unique) == U and int(v2.n_oct) == O_nodes
        res = {
            "ms_per_step": (t1 - t0) * 1000.0 / args.steps,
            "build_ms": el(0, 1), "radius_ms": el(1, 2), "knn_ms": el(2, 3), "normals_ms": el(3, 4),
            "step_ms": el(0, 4),
            "reruns": ix.knn_reruns(reset=True, stream=st.handle) / max(args.steps, 1),
            "nn": nn.numpy(st.handle), "normals": normals.numpy(st.handle),
        }
        assert np.isfinite(res["normals"]).all()
        log(f"{mode}: {res['ms_per_step']:.2f} ms/step (build {np.median(res['build_ms']):.2f}, radius "
            f"{np.median(res['radius_ms']):.2f}, knn {np.median(res['knn_ms']):.2f}, normals "
            f"{np.median(res['normals_ms']):.2f}), reruns/step {res['reruns']:.0f}")
        return res

    def host_io(runs=5):
        """The plain step with the coordinates copied in from pinned host
        memory and the normals copied back: the rate a caller holding host
        buffers sees (never `value`)."""
        nb_in, nb_out = 4 * n, 12 * U
        bufs = []
        for nb in (nb_in, nb_in, nb_in, nb_out):
            p = ctypes.c_void_p()
            sx._lib.call("sx_host_malloc", ctypes.byref(p), nb)
            bufs.append(p)
        try:
            for a, p in zip((x, y, z), bufs):
                ctypes.memmove(p, a.ctypes.data, nb_in)
            times, h2d, d2h = [], [], []
            for _ in range(runs + 1):
                e = [sx.Event() for _ in range(3)]
                st.sync()
                t0 = time.perf_counter()
                e[0].record(st.handle)
                for d, p in zip((dx, dy, dz), bufs):
                    sx._lib.call("sx_memcpy_htod", d.ptr, p, nb_in, st.handle)
                e[1].record(st.handle)
                step("plain")
                e[2].record(st.handle)
                sx._lib.call("sx_memcpy_dtoh", bufs[3], normals.ptr, nb_out, st.handle)  # synchronises
                times.append(time.perf_counter() - t0)
                h2d.append(e[0].elapsed_ms(e[1]))
                d2h.append((times[-1] * 1000.0) - e[0].elapsed_ms(e[2]))
            times, h2d, d2h = times[1:], h2d[1:], d2h[1:]  # first run warms the copy path
        finally:
            for p in bufs:
                sx._lib.call("sx_host_free", p)
        ms = float(np.median(times)) * 1000.0
        res = {"ms_per_step": ms, "points_per_s": n / (ms / 1000.0), "h2d_ms": float(np.median(h2d)),
               "d2h_ms": float(np.median(d2h)), "bytes_in": 3 * nb_in, "bytes_out": nb_out,
               "note": "pinned host xyz -> HBM, plain step, normals -> pinned host; median of "
                       f"{runs} runs (d2h includes the wait for the step's last kernel)"}
        log(f"host buffers: {ms:.2f} ms/step (h2d {res['h2d_ms']:.2f}, d2h {res['d2h_ms']:.2f})")
        return res

    modes = ["plain"] + (["learned"] if model is not None else [])
    results = {m: measure(m) for m in modes}
    pcie = host_io()
    if "learned" in results:
        assert np.array_equal(results["learned"]["nn"], results["plain"]["nn"]), "learned kNN differs from plain"
        assert np.array_equal(results["learned"]["normals"], results["plain"]["normals"])
    # both modes give bit-identical neighbours and normals; the headline is
    # the plain pipeline, the learned one is reported beside it
    # (learned_vs_plain)
    head = "plain"
    R = results[head]
    ms_per_step = R["ms_per_step"]
    value = n / (ms_per_step / 1000.0)
    kb = knn_bytes_per_query(k) * U
    knn_avg = float(R["knn_ms"].mean())
    achieved = kb / (knn_avg / 1000.0) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "knn_scene8m_traffic.json"
    if tf.exists() and n == 8_388_608 and k == 16:
        traffic = json.loads(tf.read_text()).get("hbm_bytes_per_call", {}).get(head)
    r = O_nodes / U
    pipe_bytes = pipeline_bytes_per_point(8, SORT_PAIR_PASSES, k, r) * n

    def stage(res):
        return {
            "ms_per_step": res["ms_per_step"],
            "points_per_s": n / (res["ms_per_step"] / 1000.0),
            "index_build_ms": float(np.median(res["build_ms"])),
            "radius_ms": float(np.median(res["radius_ms"])),
            "knn_ms": float(np.median(res["knn_ms"])),
            "normals_ms": float(np.median(res["normals_ms"])),
            "knn_reruns_per_step": res["reruns"],
        }

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"scene8m: {n} float32 points, 200 random primitives (planes/spheres/cylinders, "
                         f"sizes 1-20) in [0,100]^3, noise 0.01, 1% outliers, seed 2; 63-bit Morton codes; "
                         f"k={k}; FP32 normals") if headline else
                        (f"{w.name}: {n} float32 points, {w.kind} generator, seed {w.seed}; {bits}-bit Morton "
                         f"codes; k={k}; FP32 normals"),
            "knn": head,
            "bucket": args.bucket,
            "n_unique": U,
            "n_octree_nodes": O_nodes,
            "device": dev,
        },
        "modes": {m: stage(results[m]) for m in results},
        "learned_vs_plain": (results["plain"]["ms_per_step"] / results["learned"]["ms_per_step"]
                             if "learned" in results else None),
        "roofline": {
            "kernel": f"sx_knn ({head}): knn_tile_kernel<{k},{'true' if head == 'learned' else 'false'}> + "
                      f"knn_group_kernel<{k}>, HIP events around the call on its stream",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": kb,
        },
        "valu_roofline": valu_roofline(head, knn_avg, n, k),
        "pipeline_roofline": {
            "bytes_per_point": pipeline_bytes_per_point(8, SORT_PAIR_PASSES, k, r),
            "achieved_GBs": pipe_bytes / (ms_per_step / 1000.0) / 1e9,
            "frac": pipe_bytes / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS,
        },
        "host_buffers": pcie,
    }
    if cls is not None:
        line["per_class_angular_error_deg"] = per_class_errors(sx, ix, nn, normals, cls, dirs, k, st)
    else:
        line["normals_vs_fp64"] = normal_errors(sx, ix, nn, normals, k, st)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(w, n, k)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


def valu_roofline(mode: str, knn_ms: float, n: int, k: int):
    """The kNN's actual ceiling: VALU issue.  Wave instructions per call from
    the committed SQ_INSTS_VALU profile of this workload, divided by the kNN
    time measured in this run.  `peak` is the dual issue rate measured for
    two-source FP32/integer ops (the ceiling for any instruction mix);
    `frac_single_rate` compares against 256 CUs x 4 SIMDs x 2.4 GHz / 4
    cycles, the rate of the compares and selects that dominate the kNN."""
    f = ROOT / "profiles" / "knn_scene8m_valu.json"
    if not (f.exists() and n == 8_388_608 and k == 16):
        return None
    prof = json.loads(f.read_text())
    per_call = prof["valu_wave_instructions_per_call"][mode]
    achieved = per_call / (knn_ms / 1000.0)
    single = prof["peak_wave_instructions_per_s"]
    peak = prof["peak_dual_rate_wave_instructions_per_s"]
    return {"kernel": "sx_knn (tile + deferred-query kernels)", "bound": "valu", "achieved": achieved, "peak": peak,
            "unit": "wave-instructions/s", "frac": achieved / peak, "peak_single_rate": single,
            "frac_single_rate": achieved / single, "instructions_per_call": per_call,
            "peak_note": prof.get("peak_note", ""), "source": "profiles/knn_scene8m_valu.json (SQ_INSTS_VALU)"}


def normal_errors(sx, ix, nn, normals, k, st) -> dict:
    """The metric's precision gate, checked after the timed steps: the timed
    FP32 normals against the FP64 path on the same neighbours (FP64 covariance
    + FP64 Jacobi on the GPU, itself pinned to the CPU FP64 PCA in the tests),
    angle up to sign."""
    n32 = normals.numpy(st.handle).astype(np.float64)
    n64 = ix.normals(nn, k, precision=sx.SX_PREC_FP64, stream=st.handle).numpy(st.handle).astype(np.float64)
    err = np.degrees(np.arccos(np.clip(np.abs(np.sum(n32 * n64, axis=1)), 0.0, 1.0)))
    return {"reference": "FP64 covariance + FP64 Jacobi (GPU), same neighbours", "median_deg": float(np.median(err)),
            "p99_deg": float(np.percentile(err, 99)), "max_deg": float(err.max()),
            "gate_p99_le_1deg": bool(np.percentile(err, 99) <= 1.0)}


def per_class_errors(sx, ix, nn, normals, cls, dirs, k, st) -> dict:
    """Per structure class: FP32 normals against the FP64 path (FP64
    covariance + FP64 Jacobi on the GPU, itself pinned to the CPU FP64 PCA in
    the tests).  A line's normal is only defined up to a rotation about the
    line, so for lines the error is the deviation from perpendicular."""
    n32 = normals.numpy(st.handle).astype(np.float64)
    n64 = ix.normals(nn, k, precision=sx.SX_PREC_FP64, stream=st.handle).numpy(st.handle).astype(np.float64)
    rep = ix.fetch()["rep"]
    ucls, udir = cls[rep], dirs[rep].astype(np.float64)
    vs64 = np.degrees(np.arccos(np.clip(np.abs(np.sum(n32 * n64, axis=1)), 0.0, 1.0)))
    perp32 = np.degrees(np.arcsin(np.clip(np.abs(np.sum(n32 * udir, axis=1)), 0.0, 1.0)))
    perp64 = np.degrees(np.arcsin(np.clip(np.abs(np.sum(n64 * udir, axis=1)), 0.0, 1.0)))
    out = {}
    for c, name in enumerate(("lines", "corners", "slabs")):
        m = ucls == c
        if not m.any():
            continue
        if c == 0:  # both precisions on the same decoded points: the spread is the data's
            out[name] = {"n": int(m.sum()), "measure": "deviation from perpendicular to the line (deg)",
                         "fp32_median": float(np.median(perp32[m])), "fp32_p99": float(np.percentile(perp32[m], 99)),
                         "fp64_median": float(np.median(perp64[m])), "fp64_p99": float(np.percentile(perp64[m], 99))}
        else:
            a = vs64[m]
            out[name] = {"n": int(m.sum()), "measure": "FP32 vs FP64 normal (deg)", "median": float(np.median(a)),
                         "p99": float(np.percentile(a, 99)), "max": float(a.max())}
    return out


def bench_batched(sx, args, dev) -> int:
    """512 independent scenes (seed 4*1000+i, sizes log-uniform in [1k, 200k])
    through one captured graph: build + kNN (k=16) + FP32 normals."""
    k, n_clouds = 16, 512
    sizes = sx.gen_batch_sizes(4, n_clouds)
    off = np.zeros(n_clouds + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    xs, ys, zs = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.float32)
    t = time.perf_counter()
    for i in range(n_clouds):
        x, y, z = sx.gen_scene(4 * 1000 + i, int(sizes[i]))
        xs[off[i]:off[i + 1]], ys[off[i]:off[i + 1]], zs[off[i]:off[i + 1]] = x, y, z
    log(f"generated {n_clouds} clouds, {n} points in {time.perf_counter() - t:.1f} s")
    st = sx.Stream()
    dx, dy, dz = (sx.DeviceArray.from_numpy(a) for a in (xs, ys, zs))
    b = sx.Batch(n, n_clouds)
    b.set_clouds(off)
    b.build(dx, dy, dz, stream=st.handle)
    tab = b.table(stream=st.handle)
    S = tab["n_slots"]
    nn = sx.DeviceArray((S, k), np.int32)
    normals = sx.DeviceArray((S, 3), np.float32)
    tune = b.autotune(k, nn, buckets=(8, 16, 32), queue_blocks=(0, 2048), reps=2, stream=st.handle)
    log(f"autotune: bucket {tune['bucket']}, queue blocks {tune['queue_blocks']}, ms {tune['ms'].round(3).tolist()}")

    def eager():  # the same three calls as the graph, no host synchronisation
        sx._lib.call("sx_batch_build", b.handle, dx.ptr, dy.ptr, dz.ptr, st.handle)
        sx._lib.call("sx_batch_knn", b.handle, k, nn.ptr, S, st.handle)
        sx._lib.call("sx_batch_normals", b.handle, nn.ptr, k, sx.SX_PREC_FP32, 0.0, 0.0, 0.0, normals.ptr, None, S,
                     st.handle)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        st.sync()
        e0, e1 = sx.Event(), sx.Event()
        e0.record(st.handle)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        e1.record(st.handle)
        st.sync()
        return e0.elapsed_ms(e1) / steps, (time.perf_counter() - t0) * 1000.0 / steps

    eager_ms, eager_wall = timed(eager, args.steps, args.warmup)
    graph = b.graph(dx, dy, dz, k, nn, normals, st.handle)
    graph_ms, graph_wall = timed(lambda: graph.launch(st.handle), args.steps, args.warmup)
    b.deferred(reset=True, stream=st.handle)
    graph.launch(st.handle)
    deferred = b.deferred(reset=True, stream=st.handle)
    nrm = normals.numpy(st.handle)
    valid = np.zeros(S, bool)
    for c in range(n_clouds):
        valid[tab["base"][c]:tab["base"][c] + tab["size"][c]] = True
    assert np.isfinite(nrm[valid]).all()
    log(f"eager {eager_ms:.2f} ms (wall {eager_wall:.2f}), graph {graph_ms:.2f} ms (wall {graph_wall:.2f}), "
        f"deferred {deferred}")
    graph.close()

    # the same batch with every cloud's octree: plain pipeline, and the learned
    # radius (features + MLP, radius-seeded kNN) on top of it
    from shapeextraction_amd import learned as L

    model = L.RadiusModel.load(L.default_model_path(k))
    r2 = sx.DeviceArray((S,), np.float32)
    nn_l = sx.DeviceArray((S, k), np.int32)

    def build_oct():
        sx._lib.call("sx_batch_build", b.handle, dx.ptr, dy.ptr, dz.ptr, st.handle)

    build_ms, _ = timed(build_oct, args.steps, args.warmup)  # octrees still off
    b.set_octree(3.0)

    def plain_oct():
        build_oct()
        sx._lib.call("sx_batch_knn", b.handle, k, nn.ptr, S, st.handle)
        sx._lib.call("sx_batch_normals", b.handle, nn.ptr, k, sx.SX_PREC_FP32, 0.0, 0.0, 0.0, normals.ptr, None, S,
                     st.handle)

    def learned_oct():
        build_oct()
        sx._lib.call("sx_batch_radius_predict", b.handle, model.handle, r2.ptr, S, st.handle)
        sx._lib.call("sx_batch_knn_radius", b.handle, k, r2.ptr, nn_l.ptr, S, st.handle)
        sx._lib.call("sx_batch_normals", b.handle, nn_l.ptr, k, sx.SX_PREC_FP32, 0.0, 0.0, 0.0, normals.ptr, None, S,
                     st.handle)

    build_oct()
    b.view(stream=st.handle)  # capacity check of the octrees
    oct_nodes = int(b._view.n_oct)
    oct_build_ms, _ = timed(build_oct, args.steps, args.warmup)
    oct_plain_ms, _ = timed(plain_oct, args.steps, args.warmup)
    oct_learned_ms, _ = timed(learned_oct, args.steps, args.warmup)
    same = np.array_equal(nn.numpy(st.handle)[valid], nn_l.numpy(st.handle)[valid])
    assert same, "learned rows differ from plain rows"
    log(f"build alone {build_ms:.2f} ms; with octrees ({oct_nodes} nodes): build {oct_build_ms:.2f} ms, plain {oct_plain_ms:.2f} ms, "
        f"learned {oct_learned_ms:.2f} ms (rows equal: {same})")
    line = {
        "metric": "points/sec batched normals, 512 clouds (log-uniform 1k-200k points), k=16",
        "value": n / (graph_wall / 1000.0),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": graph_wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batch512: {n_clouds} scenes (200 primitives, noise 0.01, 1% outliers, seeds 4000+i), "
                        f"sizes log-uniform in [1000, 200000], {n} points; 63-bit codes; k={k}; FP32 normals",
            "n_points": n, "n_unique": tab["n_unique"], "n_slots": S,
            "cloud_sizes": {"min": int(sizes.min()), "median": int(np.median(sizes)), "max": int(sizes.max())},
            "device": dev,
        },
        "per_batch_ms": {"graph_events": graph_ms, "graph_wall": graph_wall, "eager_events": eager_ms,
                         "eager_wall": eager_wall},
        "graph_speedup_vs_eager": eager_wall / graph_wall,
        "autotune": {"bucket": tune["bucket"], "queue_blocks": tune["queue_blocks"],
                     "candidates_ms": tune["ms"].tolist()},
        "deferred_queries": deferred,
        "build_ms": build_ms,
        "with_octrees_ms": {"octree_nodes": oct_nodes, "build": oct_build_ms, "plain": oct_plain_ms,
                            "learned": oct_learned_ms, "learned_rows_equal_plain": bool(same)},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
