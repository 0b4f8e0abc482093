#!/usr/bin/env python3
"""Throughput of the trace-transform hot path on one MI355X.

A step is one forward pass of the path over one synthetic image already in
HBM: sinograms for every requested T-functional (tt_sinograms) followed by
the circus functions of every sinogram (tt_circus); for the large workload,
whose P-functionals include Hermite, the Hermite functions run on the
nearest orthonormal sinogram of each sinogram with its centre
(tt_nearest_orthonormal_sinogram, c++/src/transform.cpp:75-92) and P1-P3 on
the raw sinogram.  ``value`` is sinogram
lines per second, a line being one (angle, column) trace with all requested
T-functionals evaluated (lines per image = A * n).

Prints ONE JSON line.  Extra objects:
  roofline      the dominant kernel (the T0-T5 line kernel), its algorithmic
                bytes per launch over its average launch time, timed with HIP
                events on the stream it runs on, against 8 TB/s HBM3E;
  cpu_baseline  the oracle's OpenMP port of the reference pipeline
                (sinograms + circus + zscore, c++/src/transform.cpp:49-112)
                timed on this host, 1 warm-up + timed runs as in
                c++/src/demo.cpp:222-248.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # One 1024x1024 image, sum of 32 Gaussian blobs (seed 2), padded to 1449,
    # 360 angles at 1 deg, T0-T7; P1-P3 on each raw sinogram and Hermite
    # orders 1-5 on its nearest orthonormal sinogram with that sinogram's
    # centre (no 255x255 stretch).
    "large-1024-blob-T0-T7-P1-P3-H1-H5": dict(
        side=1024, kind="blob", seed=2, step=1.0, tcodes=[0, 1, 2, 3, 4, 5, 6, 7],
        pfuncs=[(1,), (2,), (3,)] + [(0, k, -1) for k in range(1, 6)],
        hermite_on="nos"),
    # The same 1024x1024 blob image, T0-T5 and P1-P3 only: the line kernel
    # at the large size (no sort, no orthonormal step).
    "large-1024-blob-T0-T5-P1-P3": dict(
        side=1024, kind="blob", seed=2, step=1.0, tcodes=[0, 1, 2, 3, 4, 5],
        pfuncs=[(1,), (2,), (3,)]),
    # One 256x256 U[0,1) image (seed 1) padded to 363, 360 angles at 1 deg,
    # T0-T5, P1-P3.
    "plain-256-uniform-T0-T5-P1-P3": dict(
        side=256, kind="uniform", seed=1, step=1.0, tcodes=[0, 1, 2, 3, 4, 5],
        pfuncs=[(1,), (2,), (3,)]),
    # 256 blob images with sides uniform in 64..512 (seed 3), each padded to
    # its own n, 180 angles at 2 deg, T0-T5, P1-P3: one batched graph replay
    # per step, beside the plain per-image path.
    "batched-256-blob-T0-T5-P1-P3": dict(
        batch=256, lo=64, hi=512, kind="blob", seed=3, step=2.0,
        tcodes=[0, 1, 2, 3, 4, 5], pfuncs=[(1,), (2,), (3,)]),
    # Forward + adjoint of one 512x512 blob image (seed 4) padded to 725, 360
    # angles at 1 deg, T0/T3/T4/T5, P1-P3, loss = sum of squared circus
    # differences against a second image (seed 5): a step is sinograms,
    # circus, dC = 2 (C - C2) and the backward pass to the image.
    "adjoint-512-blob-T0-T3-T5-P1-P3": dict(
        side=512, kind="blob", seed=4, target_seed=5, step=1.0, tcodes=[0, 3, 4, 5],
        pfuncs=[(1,), (2,), (3,)], adjoint=True),
    # Batched forward + adjoint: 64 blob images with sides uniform in
    # 256..1024 (seed 6), 720 angles at 0.5 deg, T0/T3/T4/T5, P1-P3, targets
    # from a second set of images; a step is one batched forward replay, the
    # loss gradient and one batched backward replay.
    "batched-adjoint-64-blob-T0-T3-T5-P1-P3": dict(
        batch=64, lo=256, hi=1024, kind="blob", seed=6, step=0.5, tcodes=[0, 3, 4, 5],
        pfuncs=[(1,), (2,), (3,)], adjoint=True),
}
DEFAULT_WORKLOAD = "large-1024-blob-T0-T7-P1-P3-H1-H5"


def make_image(w, side=None, seed=None):
    from tracetransform_amd import inputs
    side = w["side"] if side is None else side
    seed = w["seed"] if seed is None else seed
    if w["kind"] == "uniform":
        return inputs.uniform_image(side, seed)
    return inputs.blob_image(side, seed)


def split_pfuncs(w):
    """(regular P-functionals, Hermite orders) of a workload."""
    return ([p for p in w["pfuncs"] if p[0] != 0],
            [p[1] for p in w["pfuncs"] if p[0] == 0])


def sino_bytes(n, A, nt):
    """Algorithmic bytes of one sinogram launch: one 4*n^2 pass over the image
    per angle, the image read once and the sinograms written once."""
    return 4.0 * n * n * A + 4.0 * n * n + 4.0 * n * A * nt


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def cpu_threads():
    threads = int(os.environ.get("TT_CPU_THREADS", "16"))
    return max(1, min(threads, os.cpu_count() or 1))


def cpu_transform_nos(tt_oracle, img, n, A, w, threads):
    """Reference getTransform with P1-P3 on the raw sinograms and Hermite on
    each nearest orthonormal sinogram (numpy FP64 SVD), then z-scores."""
    reg, orders = split_pfuncs(w)
    sino, _ = tt_oracle.sinograms(img, n, A, w["step"], w["tcodes"], threads)
    sig = []
    for t in range(len(w["tcodes"])):
        if reg:
            c, _ = tt_oracle.circus(sino[t], n, A, reg, threads)
            sig.extend(tt_oracle.zscore(row) for row in c)
        if orders:
            nos, center = tt_oracle.nearest_orthonormal_sinogram(sino[t])
            c, _ = tt_oracle.circus(nos, nos.shape[1], A,
                                    [(0, k, center) for k in orders], threads)
            sig.extend(tt_oracle.zscore(row) for row in c)
    return sig


def cpu_baseline(img_cm, w, n, A, iters, side):
    """OpenMP port of the reference pipeline on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tt_oracle
    threads = cpu_threads()
    img_cm = tt_oracle.pad(img_cm.ravel(), side, side)
    nos = w.get("hermite_on") == "nos"

    def once():
        if nos:
            cpu_transform_nos(tt_oracle, img_cm, n, A, w, threads)
        else:
            tt_oracle.transform(img_cm, n, A, w["step"], w["tcodes"], w["pfuncs"], threads)

    once()
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        once()
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    what = ("sinograms+circus P1-P3+NOS (numpy FP64 SVD)+Hermite+zscore" if nos
            else "sinograms+circus+zscore")
    return {"value": A * n / med, "unit": "lines/s", "cores": threads,
            "kind": "port",
            "sample": f"full workload ({A} angles x {n} lines), 1 warm-up + "
                      f"{iters} timed runs of {what}, median"}


def cpu_baseline_batch(imgs_cm, sides, w, A, budget_s=20.0):
    """Per-image reference pipeline on this host over a prefix of the batch
    (images in batch order) until ~budget_s of CPU time, one warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tt_oracle
    threads = cpu_threads()
    done_lines, spent, used = 0, 0.0, 0
    warm = tt_oracle.pad(imgs_cm[0].ravel(), int(sides[0]), int(sides[0]))
    n0, _, _ = tt_oracle.pad_geometry(int(sides[0]), int(sides[0]))
    tt_oracle.transform(warm, n0, A, w["step"], w["tcodes"], w["pfuncs"], threads)
    for img, side in zip(imgs_cm, sides):
        side = int(side)
        n, _, _ = tt_oracle.pad_geometry(side, side)
        p = tt_oracle.pad(img.ravel(), side, side)
        t0 = time.perf_counter()
        tt_oracle.transform(p, n, A, w["step"], w["tcodes"], w["pfuncs"], threads)
        spent += time.perf_counter() - t0
        done_lines += A * n
        used += 1
        if spent >= budget_s:
            break
    return {"value": done_lines / spent, "unit": "lines/s", "cores": threads,
            "kind": "port",
            "sample": f"first {used} of {len(imgs_cm)} batch images processed one "
                      f"after another (sinograms+circus+zscore), {spent:.1f} s"}


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB "
                         "(calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def calibrate_copy(mib):
    """One timed device copy of a known size, for scaling HBM counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n,
            "copy_ms": round(dt * 1e3, 4), "copy_gbs": round(8 * n / dt / 1e9, 1)}


def read_kernels(ctx, steps):
    from tracetransform_amd import _lib
    kinds = (("sino_t05", _lib.TT_KERNEL_SINO_T05), ("sino_sort", _lib.TT_KERNEL_SINO_SORT),
             ("circus", _lib.TT_KERNEL_CIRCUS), ("aux", _lib.TT_KERNEL_AUX),
             ("batch_graph", _lib.TT_KERNEL_BATCH), ("gather", _lib.TT_KERNEL_GATHER),
             ("circus_vjp", _lib.TT_KERNEL_CIRCUS_VJP),
             ("batch_backward_graph", _lib.TT_KERNEL_BATCH_BACKWARD),
             ("nos", _lib.TT_KERNEL_NOS))
    return {name: ctx.profile_read(k) for name, k in kinds}


def timed(ctx, step, steps, warmup):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    read_kernels(ctx, 0)
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.profile(False)
    return (t1 - t0) * 1e3 / steps, read_kernels(ctx, steps)


def roofline_of(kern, bytes_launch):
    ms_k, cnt = kern["sino_t05"]
    avg_ms = ms_k / max(cnt, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "kernel": "sino_t05_kernel",
            "avg_launch_ms": round(avg_ms, 4),
            "algorithmic_bytes_per_launch": bytes_launch}


def run_single(args, w):
    import numpy as np
    import torch
    from tracetransform_amd import core

    img_cm = make_image(w)
    side = w["side"]
    n, _, _ = core.pad_geometry(side, side)
    A = core.angle_count(w["step"])
    reg, orders = split_pfuncs(w)
    if orders and w.get("hermite_on") != "nos":
        raise ValueError("Hermite P-functionals need the orthonormal sinogram")
    nt = len(w["tcodes"])
    dev = torch.device("cuda", 0)
    d_raw = torch.from_numpy(np.ascontiguousarray(img_cm)).to(dev)
    d_img = core.pad(d_raw, side, side)
    sino = torch.empty((nt, A, n), dtype=torch.float32, device=dev)
    circ = torch.empty((nt, len(reg), A), dtype=torch.float32, device=dev)
    circ_h = torch.empty((nt, len(orders), A), dtype=torch.float32, device=dev)
    nos_ws = torch.empty((nt, A * 3 * n), dtype=torch.float32, device=dev)
    centers = [None] * nt
    ctx = core.get_context(0)

    def step():
        core.sinograms(d_img, A, w["step"], w["tcodes"], out=sino)
        if reg:
            core.circus(sino, reg, out=circ)
        if orders:
            nos, centers[:] = core.nearest_orthonormal_sinograms(sino, out=nos_ws)
            for t in range(nt):
                core.circus(nos[t], [(0, k, centers[t]) for k in orders],
                            out=circ_h[t:t + 1])

    ms_step, kern = timed(ctx, step, args.steps, args.warmup)
    lines = A * n
    roof = roofline_of(kern, sino_bytes(n, A, min(nt, 6)))
    roof["traffic"] = load_traffic(args.workload)
    out = {
        "metric": "sinogram lines/s", "value": round(lines / (ms_step * 1e-3), 1),
        "unit": "lines/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "image": f"{side}x{side} {w['kind']} "
                   f"seed {w['seed']}", "padded_n": n, "angles": A,
                   "angle_step_deg": w["step"], "tfunctionals": w["tcodes"],
                   "pfunctionals": [p[0] for p in reg] + [f"H{k}" for k in orders],
                   "hermite_on": ("nearest orthonormal sinogram, centres "
                                  f"{sorted(set(centers))}") if orders else None,
                   "lines_per_step": lines},
        "roofline": roof,
        "kernel_ms_per_step": {k: round(v[0] / max(args.steps, 1), 4)
                               for k, v in kern.items() if v[1]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(img_cm, w, n, A, args.cpu_iters, side)
    return out


def run_batched(args, w):
    import numpy as np
    import torch
    from tracetransform_amd import core, inputs

    sides = inputs.batch_sides(w["batch"], w["lo"], w["hi"], w["seed"])
    imgs_cm = [make_image(w, side=int(sd), seed=w["seed"] * 1000 + k)
               for k, sd in enumerate(sides)]
    dev = torch.device("cuda", 0)
    imgs = []
    for img, sd in zip(imgs_cm, sides):
        d_raw = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
        imgs.append(core.pad(d_raw, int(sd), int(sd)))
    A = core.angle_count(w["step"])
    ns = [im.shape[0] for im in imgs]
    lines = A * sum(ns)
    ctx = core.get_context(0)
    batch = core.Batch(imgs, A, w["step"], w["tcodes"], w["pfuncs"])
    batch.run(autotune=bool(args.autotune))  # plans, tunes and captures
    torch.cuda.synchronize()
    ms_step, kern = timed(ctx, batch.run, args.steps, args.warmup)
    shape = batch.launch_shape()

    # plain per-image path: the same kernels launched image by image
    nt, npf = len(w["tcodes"]), len(w["pfuncs"])
    sinos = [torch.empty((nt, A, n), dtype=torch.float32, device=dev) for n in ns]
    circs = [torch.empty((nt, npf, A), dtype=torch.float32, device=dev) for _ in ns]

    def plain():
        for im, so, co in zip(imgs, sinos, circs):
            core.sinograms(im, A, w["step"], w["tcodes"], out=so)
            core.circus(so, w["pfuncs"], out=co)

    ms_plain, kern_plain = timed(ctx, plain, max(2, args.steps // 2), 1)
    gbytes = sum(sino_bytes(n, A, nt) for n in ns)
    ms_graph, cnt = kern["batch_graph"]
    avg_graph = ms_graph / max(cnt, 1)
    achieved = gbytes / (avg_graph * 1e-3) / 1e9
    out = {
        "metric": "sinogram lines/s", "value": round(lines / (ms_step * 1e-3), 1),
        "unit": "lines/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "images": len(ns),
                   "sides": f"{int(min(sides))}..{int(max(sides))}",
                   "padded_n": f"{min(ns)}..{max(ns)}", "angles": A,
                   "angle_step_deg": w["step"], "tfunctionals": w["tcodes"],
                   "pfunctionals": [p[0] for p in w["pfuncs"]],
                   "lines_per_step": lines, "launch_shape": list(shape),
                   "autotune": bool(args.autotune)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "batched graph (transpose + line + circus kernels)",
                     "avg_launch_ms": round(avg_graph, 4),
                     "algorithmic_bytes_per_launch": gbytes},
        "plain_per_image": {"ms_per_step": round(ms_plain, 4),
                            "value": round(lines / (ms_plain * 1e-3), 1),
                            "batched_speedup": round(ms_plain / ms_step, 3)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(imgs_cm, sides, w, A)
    return out


def adjoint_bytes(n, A, nt, npf):
    """Algorithmic bytes of forward + adjoint for one image: three 4*n^2
    passes per angle (forward, recompute, transposed rotation), the image
    and its gradient, sinograms written and read, circus and its gradient."""
    return (12.0 * n * n * A + 8.0 * n * n + 8.0 * n * A * nt
            + 8.0 * A * nt * npf)


def gather_bytes(n, A):
    """Algorithmic bytes of one transposed-rotation launch: one 4*n^2 pass
    over the rotated-domain gradient per angle, the gradient written once."""
    return 4.0 * n * n * A + 4.0 * n * n


def cpu_baseline_adjoint(imgs, targets, ns, w, A, budget_s=20.0):
    """The oracle's FP32 forward (OpenMP) and FP64 analytic backward (OpenMP)
    per image on this host, over a prefix of the images until ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import tt_oracle
    threads = cpu_threads()
    done_lines, spent, used = 0, 0.0, 0
    for img, tgt, n in zip(imgs, targets, ns):
        t0 = time.perf_counter()
        sino, aux = tt_oracle.sinograms(img, n, A, w["step"], w["tcodes"], threads)
        circ, src = [], []
        for t in range(len(w["tcodes"])):
            c, a = tt_oracle.circus(sino[t], n, A, w["pfuncs"], threads)
            circ.append(c)
            src.append(a["p2_src"])
        dc = 2.0 * (np.stack(circ).astype(np.float64) - tgt)
        tt_oracle.backward_f64(img, n, A, w["step"], w["tcodes"], w["pfuncs"], aux,
                               np.stack(src), dc, threads)
        spent += time.perf_counter() - t0
        done_lines += A * n
        used += 1
        if spent >= budget_s:
            break
    return {"value": done_lines / spent, "unit": "lines/s", "cores": threads,
            "kind": "port",
            "sample": f"first {used} of {len(imgs)} images one after another: FP32 "
                      f"forward + FP64 analytic backward of the oracle, {spent:.1f} s"}


def run_adjoint(args, w):
    """Forward + adjoint of one image (plain path: it is one image)."""
    import numpy as np
    import torch
    from tracetransform_amd import core

    side = w["side"]
    n, _, _ = core.pad_geometry(side, side)
    A = core.angle_count(w["step"])
    nt, npf = len(w["tcodes"]), len(w["pfuncs"])
    dev = torch.device("cuda", 0)
    img_cm = make_image(w)
    tgt_cm = make_image(w, seed=w["target_seed"])
    d_img = core.pad(torch.from_numpy(np.ascontiguousarray(img_cm)).to(dev), side, side)
    d_tgt = core.pad(torch.from_numpy(np.ascontiguousarray(tgt_cm)).to(dev), side, side)
    target = core.circus(core.sinograms(d_tgt, A, w["step"], w["tcodes"]), w["pfuncs"])
    sino = torch.empty((nt, A, n), dtype=torch.float32, device=dev)
    circ = torch.empty((nt, npf, A), dtype=torch.float32, device=dev)
    dcirc = torch.empty_like(circ)
    grad = [None]

    def step():
        core.sinograms(d_img, A, w["step"], w["tcodes"], out=sino)
        core.circus(sino, w["pfuncs"], out=circ)
        torch.sub(circ, target, out=dcirc)
        dcirc.mul_(2.0)
        grad[0] = core.backward(d_img, A, w["step"], w["tcodes"], w["pfuncs"], dcirc)

    ctx = core.get_context(0)
    ms_step, kern = timed(ctx, step, args.steps, args.warmup)
    lines = A * n
    ms_g, cnt_g = kern["gather"]
    avg_g = ms_g / max(cnt_g, 1)
    gb = gather_bytes(n, A)
    total = adjoint_bytes(n, A, nt, npf)
    out = {
        "metric": "sinogram lines/s", "value": round(lines / (ms_step * 1e-3), 1),
        "unit": "lines/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "image": f"{side}x{side} {w['kind']} "
                   f"seed {w['seed']}, target seed {w['target_seed']}", "padded_n": n,
                   "angles": A, "angle_step_deg": w["step"], "tfunctionals": w["tcodes"],
                   "pfunctionals": [p[0] for p in w["pfuncs"]], "lines_per_step": lines,
                   "step": "forward + loss gradient + backward"},
        "roofline": {"bound": "hbm", "achieved": round(gb / (avg_g * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gb / (avg_g * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(args.workload),
                     "kernel": "gather_kernel (transposed rotation, with line records)",
                     "avg_launch_ms": round(avg_g, 4),
                     "algorithmic_bytes_per_launch": gb},
        "step_roofline": {"algorithmic_bytes": total,
                          "achieved_gbs": round(total / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(total / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_ms_per_step": {k: round(v[0] / max(args.steps, 1), 4)
                               for k, v in kern.items() if v[1]},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tt_oracle
        p_img = tt_oracle.pad(img_cm.ravel(), side, side)
        p_tgt = tt_oracle.pad(tgt_cm.ravel(), side, side)
        s2, _ = tt_oracle.sinograms(p_tgt, n, A, w["step"], w["tcodes"], cpu_threads())
        tgt = np.stack([tt_oracle.circus(s2[t], n, A, w["pfuncs"], cpu_threads())[0]
                        for t in range(nt)]).astype(np.float64)
        out["cpu_baseline"] = cpu_baseline_adjoint([p_img], [tgt], [n], w, A)
    return out


def run_batched_adjoint(args, w):
    """Batched forward + adjoint replays beside the plain per-image path."""
    import numpy as np
    import torch
    from tracetransform_amd import core, inputs

    sides = inputs.batch_sides(w["batch"], w["lo"], w["hi"], w["seed"])
    dev = torch.device("cuda", 0)
    imgs_cm, imgs, tgts = [], [], []
    for k, sd in enumerate(sides):
        sd = int(sd)
        a = make_image(w, side=sd, seed=w["seed"] * 1000 + k)
        b = make_image(w, side=sd, seed=w["seed"] * 1000 + 500 + k)
        imgs_cm.append(a)
        imgs.append(core.pad(torch.from_numpy(np.ascontiguousarray(a)).to(dev), sd, sd))
        tgts.append(core.pad(torch.from_numpy(np.ascontiguousarray(b)).to(dev), sd, sd))
    A = core.angle_count(w["step"])
    ns = [im.shape[0] for im in imgs]
    nt, npf = len(w["tcodes"]), len(w["pfuncs"])
    lines = A * sum(ns)
    ctx = core.get_context(0)
    target = torch.stack([core.circus(core.sinograms(t, A, w["step"], w["tcodes"]),
                                      w["pfuncs"]) for t in tgts])
    del tgts
    fwd = core.Batch(imgs, A, w["step"], w["tcodes"], w["pfuncs"], save_for_backward=True)
    bwd = core.BatchBackward(imgs, A, w["step"], w["tcodes"], w["pfuncs"], forward=fwd)

    def step():
        fwd.run()
        torch.sub(fwd.circus_all, target, out=bwd.dcircus_all)
        bwd.dcircus_all.mul_(2.0)
        bwd.run()

    fwd.run(autotune=bool(args.autotune))
    step()
    torch.cuda.synchronize()
    ms_step, kern = timed(ctx, step, args.steps, args.warmup)

    sinos = [torch.empty((nt, A, n), dtype=torch.float32, device=dev) for n in ns]
    circs = torch.empty((len(ns), nt, npf, A), dtype=torch.float32, device=dev)
    dcircs = torch.empty_like(circs)

    def plain():
        for k, im in enumerate(imgs):
            core.sinograms(im, A, w["step"], w["tcodes"], out=sinos[k])
            core.circus(sinos[k], w["pfuncs"], out=circs[k])
        torch.sub(circs, target, out=dcircs)
        dcircs.mul_(2.0)
        for k, im in enumerate(imgs):
            core.backward(im, A, w["step"], w["tcodes"], w["pfuncs"], dcircs[k])

    ms_plain, _ = timed(ctx, plain, max(2, args.steps // 2), 1)
    total = sum(adjoint_bytes(n, A, nt, npf) for n in ns)
    ms_f, cnt_f = kern["batch_graph"]
    ms_b, cnt_b = kern["batch_backward_graph"]
    avg_b = ms_b / max(cnt_b, 1)
    bb = sum(8.0 * n * n * A + 8.0 * n * n for n in ns)  # recompute + gather passes
    out = {
        "metric": "sinogram lines/s", "value": round(lines / (ms_step * 1e-3), 1),
        "unit": "lines/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "images": len(ns),
                   "sides": f"{int(min(sides))}..{int(max(sides))}",
                   "padded_n": f"{min(ns)}..{max(ns)}", "angles": A,
                   "angle_step_deg": w["step"], "tfunctionals": w["tcodes"],
                   "pfunctionals": [p[0] for p in w["pfuncs"]], "lines_per_step": lines,
                   "step": "batched forward + loss gradient + batched backward",
                   "launch_shape": list(fwd.launch_shape())},
        "roofline": {"bound": "hbm", "achieved": round(bb / (avg_b * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bb / (avg_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(args.workload),
                     "kernel": "batched backward graph (recompute, circus VJP, records, "
                               "gather)",
                     "avg_launch_ms": round(avg_b, 4),
                     "algorithmic_bytes_per_launch": bb},
        "step_roofline": {"algorithmic_bytes": total,
                          "achieved_gbs": round(total / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(total / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "graph_ms_per_step": {"forward": round(ms_f / max(cnt_f, 1), 4),
                              "backward": round(avg_b, 4)},
        "plain_per_image": {"ms_per_step": round(ms_plain, 4),
                            "value": round(lines / (ms_plain * 1e-3), 1),
                            "batched_speedup": round(ms_plain / ms_step, 3)},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tt_oracle
        tgt_np = target.cpu().numpy().astype(np.float64)
        p_imgs = [tt_oracle.pad(a.ravel(), int(sd), int(sd)) for a, sd in
                  zip(imgs_cm[:8], sides[:8])]
        out["cpu_baseline"] = cpu_baseline_adjoint(p_imgs, list(tgt_np[:8]), ns[:8], w, A)
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path "
              "shards by image or angle without a collective in the forward pass, "
              "see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib:
        print(json.dumps(calibrate_copy(args.calibrate_copy_mib)), flush=True)
        return
    w = WORKLOADS[args.workload]
    if w.get("adjoint"):
        out = run_batched_adjoint(args, w) if "batch" in w else run_adjoint(args, w)
    else:
        out = run_batched(args, w) if "batch" in w else run_single(args, w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
