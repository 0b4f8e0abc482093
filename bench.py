#!/usr/bin/env python3
"""Throughput of the trace-transform hot path on one MI355X.

Default workload: the batched forward + adjoint of 64 blob images with sides
256..1024 at 720 angles (0.5 deg), T0/T3/T4/T5 and P1-P3: a step is one
batched forward (sinograms, circus functions, the line data kept for the
backward), the loss gradient dC = 2 (C - C_target) and one batched backward
to the 64 image gradients.  Other workloads (--workload) time the plain
forward of one image, the large all-T image with the orthonormal Hermite
leg, the 256-image batched forward and the single-image adjoint.  ``value``
is sinogram lines per second, a line being one (image, angle, column) trace
with all requested T-functionals evaluated (lines per image = A * n).

Prints ONE JSON line.  Extra objects:
  roofline      the kernel with the largest time per step: its algorithmic
                bytes per step (one 4 n^2 pass over an image per angle, the
                streaming model) over its time per step, timed with HIP events
                the library records around it on the stream it runs on,
                against 8 TB/s HBM3E;
  step_roofline the whole step's algorithmic bytes over the step time;
  cpu_baseline  the oracle's OpenMP port of the reference pipeline
                (c++/src/transform.cpp:49-112; openmp/src/sinogram.cpp:110,
                openmp/src/circus.cpp:129) timed on this host, 1 warm-up + at
                least 5 timed runs as in c++/src/demo.cpp:222-248.
The steps are timed twice: with the graphs replayed (``ms_per_step``,
``value``), then with per-kernel timing on (the batched entry points then
launch their stages one by one, ``profiled_ms_per_step``).
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
DEFAULT_WORKLOAD = "batched-adjoint-64-blob-T0-T3-T5-P1-P3"


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


def load_traffic(workload, kind):
    """Measured HBM bytes per launch of one kernel kind in one workload
    (rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by a device copy's reading),
    from profiles/hbm_traffic.json, or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload, {}).get(kind)
    except (OSError, ValueError, AttributeError):
        return None


def cpu_threads():
    """OpenMP threads of the CPU port: OMP_NUM_THREADS (the host share a GPU
    box gives one GPU's jobs), else every core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, int(env) if env.isdigit() else (os.cpu_count() or 1))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_record(value, unit, kind, sample, threads):
    return {"value": value, "unit": unit, "cores": threads, "kind": kind,
            "sample": sample, "nproc": os.cpu_count(), "cpu_model": cpu_model()}


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
    return cpu_record(A * n / med, "lines/s", "port",
                      f"full workload ({A} angles x {n} lines), 1 warm-up + {iters} timed "
                      f"runs of {what}, median (spread {times[0]:.3f}..{times[-1]:.3f} s)",
                      threads)


def cpu_baseline_batch(imgs_cm, sides, w, A, budget_s=15.0, min_runs=5):
    """Per-image reference pipeline on this host over the batch's images in
    batch order (the reference processes images one after another,
    c++/src/demo.cpp:179), one warm-up, until at least min_runs images and
    budget_s seconds."""
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
        if used >= min_runs and spent >= budget_s:
            break
    return cpu_record(done_lines / spent, "lines/s", "port",
                      f"first {used} of {len(imgs_cm)} batch images processed one after "
                      f"another (sinograms+circus+zscore), 1 warm-up, {spent:.1f} s",
                      threads)


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-iters", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain per-image comparison of batched workloads")
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


KINDS = ("sino_t05", "sino_sort", "circus", "aux", "gather", "circus_vjp", "nos",
         "p3_dft", "records")


def read_kernels(ctx):
    from tracetransform_amd import _lib
    ids = {"sino_t05": _lib.TT_KERNEL_SINO_T05, "sino_sort": _lib.TT_KERNEL_SINO_SORT,
           "circus": _lib.TT_KERNEL_CIRCUS, "aux": _lib.TT_KERNEL_AUX,
           "gather": _lib.TT_KERNEL_GATHER, "circus_vjp": _lib.TT_KERNEL_CIRCUS_VJP,
           "nos": _lib.TT_KERNEL_NOS, "p3_dft": _lib.TT_KERNEL_P3_DFT,
           "records": _lib.TT_KERNEL_RECORDS}
    return {name: ctx.profile_read(ids[name]) for name in KINDS}


def timed(ctx, step, steps, warmup):
    """ms per step with the graphs replayed, then the per-kernel times of
    the same number of steps with kernel timing on: (ms, profiled ms,
    {kind: (ms, launches)})."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    read_kernels(ctx)
    ctx.profile(True)
    step()  # the first timed-stage call of a batch builds nothing new
    torch.cuda.synchronize()
    read_kernels(ctx)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms_prof = (time.perf_counter() - t0) * 1e3 / steps
    kern = read_kernels(ctx)
    ctx.profile(False)
    return ms, ms_prof, kern


def roofline_of(kern, steps, models):
    """Roofline of the kernel with the largest time per step among those
    with a byte model: models = {kind: algorithmic bytes per step}."""
    best = None
    for kind, nbytes in models.items():
        ms, cnt = kern.get(kind, (0.0, 0))
        if cnt and (best is None or ms > best[1]):
            best = (kind, ms, cnt, nbytes)
    if best is None:
        return None
    kind, ms, cnt, nbytes = best
    per_step_ms = ms / steps
    achieved = nbytes / (per_step_ms * 1e-3) / 1e9
    launches = cnt / steps
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": KERNEL_NAMES[kind], "kind": kind, "ms_per_step": round(per_step_ms, 4),
            "launches_per_step": launches,
            "avg_launch_ms": round(per_step_ms / launches, 4),
            "algorithmic_bytes_per_launch": nbytes / launches}


KERNEL_NAMES = {
    "sino_t05": "sino_t05_kernel / sino_t05_batched_kernel (T0-T5 line kernel)",
    "sino_sort": "sino_sort_kernel (T6/T7, up to three passes)",
    "gather": "gather_kernel / gather_batched_kernel (transposed rotation)",
    "circus": "circus kernels (P1-P3, Hermite; P2 wave sort)",
    "circus_vjp": "circus_vjp kernels (adjoint of P1, P2, Hermite)",
}


def per_kernel(kern, steps):
    return {k: round(v[0] / max(steps, 1), 4) for k, v in kern.items() if v[1]}


def base_line(args, value, ms, ms_prof, config):
    return {"metric": "sinogram lines/s", "value": round(value, 1), "unit": "lines/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 4), "profiled_ms_per_step": round(ms_prof, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "config": config}


def step_roofline(nbytes, ms, what):
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"algorithmic_bytes": nbytes, "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "model": what}


def forward_models(ns, A, nt, sort_lines=None):
    """Byte models of the forward kernels, per step."""
    m = {"sino_t05": sum(sino_bytes(n, A, min(nt, 6)) for n in ns)}
    if sort_lines:
        # T6/T7 re-sample the lines' tails: at most one 4 n^2 pass per angle
        m["sino_sort"] = sum(4.0 * n * n * A + 8.0 * n * A for n in ns)
    return m


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

    ms, ms_prof, kern = timed(ctx, step, args.steps, args.warmup)
    lines = A * n
    models = forward_models([n], A, nt, sort_lines=any(c >= 6 for c in w["tcodes"]))
    roof = roofline_of(kern, args.steps, models)
    if roof:
        roof["traffic"] = load_traffic(args.workload, roof["kind"])
    out = base_line(args, lines / (ms * 1e-3), ms, ms_prof, {
        "workload": args.workload, "image": f"{side}x{side} {w['kind']} seed {w['seed']}",
        "padded_n": n, "angles": A, "angle_step_deg": w["step"],
        "tfunctionals": w["tcodes"],
        "pfunctionals": [p[0] for p in reg] + [f"H{k}" for k in orders],
        "hermite_on": ("nearest orthonormal sinogram, centres "
                       f"{sorted(set(centers))}") if orders else None,
        "lines_per_step": lines})
    out["roofline"] = roof
    out["step_roofline"] = step_roofline(sum(models.values()), ms,
                                         "one 4 n^2 pass per angle per line kernel")
    out["kernel_ms_per_step"] = per_kernel(kern, args.steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(img_cm, w, n, A, args.cpu_iters, side)
    return out


def batch_images(w, save_seed_offset=None):
    import numpy as np
    import torch
    from tracetransform_amd import core, inputs
    sides = inputs.batch_sides(w["batch"], w["lo"], w["hi"], w["seed"])
    dev = torch.device("cuda", 0)
    imgs_cm, imgs = [], []
    for k, sd in enumerate(sides):
        a = make_image(w, side=int(sd), seed=w["seed"] * 1000 + k)
        imgs_cm.append(a)
        imgs.append(core.pad(torch.from_numpy(np.ascontiguousarray(a)).to(dev),
                             int(sd), int(sd)))
    return sides, imgs_cm, imgs


def run_batched(args, w):
    import torch
    from tracetransform_amd import core

    sides, imgs_cm, imgs = batch_images(w)
    dev = imgs[0].device
    A = core.angle_count(w["step"])
    ns = [im.shape[0] for im in imgs]
    lines = A * sum(ns)
    ctx = core.get_context(0)
    batch = core.Batch(imgs, A, w["step"], w["tcodes"], w["pfuncs"])
    batch.run(autotune=bool(args.autotune))  # plans, tunes and captures
    torch.cuda.synchronize()
    ms, ms_prof, kern = timed(ctx, batch.run, args.steps, args.warmup)
    shape = batch.launch_shape()
    nt = len(w["tcodes"])
    models = forward_models(ns, A, nt)
    out = base_line(args, lines / (ms * 1e-3), ms, ms_prof, {
        "workload": args.workload, "images": len(ns),
        "sides": f"{int(min(sides))}..{int(max(sides))}",
        "padded_n": f"{min(ns)}..{max(ns)}", "angles": A, "angle_step_deg": w["step"],
        "tfunctionals": w["tcodes"], "pfunctionals": [p[0] for p in w["pfuncs"]],
        "lines_per_step": lines, "launch_shape": list(shape[:2]),
        "tiles": core.batched_tiles(0), "autotune": bool(args.autotune),
        "step": "one batched forward (graph replay)"})
    out["roofline"] = roofline_of(kern, args.steps, models)
    if out["roofline"]:
        out["roofline"]["traffic"] = load_traffic(args.workload, out["roofline"]["kind"])
    out["step_roofline"] = step_roofline(sum(models.values()), ms,
                                         "one 4 n^2 pass per angle and image")
    out["kernel_ms_per_step"] = per_kernel(kern, args.steps)
    if not args.no_plain:
        # plain per-image path: the same kernels launched image by image
        npf = len(w["pfuncs"])
        sinos = [torch.empty((nt, A, n), dtype=torch.float32, device=dev) for n in ns]
        circs = [torch.empty((nt, npf, A), dtype=torch.float32, device=dev) for _ in ns]

        def plain():
            for im, so, co in zip(imgs, sinos, circs):
                core.sinograms(im, A, w["step"], w["tcodes"], out=so)
                core.circus(so, w["pfuncs"], out=co)

        ms_plain, _, _ = timed(ctx, plain, max(2, args.steps // 2), 1)
        out["plain_per_image"] = {"ms_per_step": round(ms_plain, 4),
                                  "value": round(lines / (ms_plain * 1e-3), 1),
                                  "batched_speedup": round(ms_plain / ms, 3)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(imgs_cm, sides, w, A)
    return out


def adjoint_bytes(n, A, nt, npf, passes=3):
    """Algorithmic bytes of forward + adjoint for one image: `passes` 4 n^2
    passes per angle (the streaming model counts three: forward, recompute,
    transposed rotation), the image and its gradient, sinograms written and
    read, circus and its gradient."""
    return (4.0 * passes * n * n * A + 8.0 * n * n + 8.0 * n * A * nt
            + 8.0 * A * nt * npf)


def gather_bytes(n, A):
    """Algorithmic bytes of one transposed rotation: one 4 n^2 pass over the
    rotated-domain gradient per angle, the gradient written once."""
    return 4.0 * n * n * A + 4.0 * n * n


def cpu_baseline_adjoint(imgs, targets, ns, w, A, budget_s=15.0, min_runs=5):
    """The oracle's FP32 forward (OpenMP) and FP64 analytic backward (OpenMP)
    per image on this host, images in batch order, one warm-up on the first,
    until at least min_runs images and budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import tt_oracle
    threads = cpu_threads()

    def one(img, tgt, n):
        sino, aux = tt_oracle.sinograms(img, n, A, w["step"], w["tcodes"], threads)
        circ, src = [], []
        for t in range(len(w["tcodes"])):
            c, a = tt_oracle.circus(sino[t], n, A, w["pfuncs"], threads)
            circ.append(c)
            src.append(a["p2_src"])
        dc = 2.0 * (np.stack(circ).astype(np.float64) - tgt)
        tt_oracle.backward_f64(img, n, A, w["step"], w["tcodes"], w["pfuncs"], aux,
                               np.stack(src), dc, threads)

    one(imgs[0], targets[0], ns[0])
    done_lines, spent, used = 0, 0.0, 0
    for img, tgt, n in zip(imgs, targets, ns):
        t0 = time.perf_counter()
        one(img, tgt, n)
        spent += time.perf_counter() - t0
        done_lines += A * n
        used += 1
        if used >= min_runs and spent >= budget_s:
            break
    return cpu_record(done_lines / spent, "lines/s", "port",
                      f"first {used} of {len(imgs)} images one after another, 1 warm-up: "
                      f"FP32 forward + FP64 analytic backward of the oracle, {spent:.1f} s",
                      threads)


def adjoint_models(ns, A, nt):
    m = forward_models(ns, A, nt)
    m["gather"] = sum(gather_bytes(n, A) for n in ns)
    return m


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
    ms, ms_prof, kern = timed(ctx, step, args.steps, args.warmup)
    lines = A * n
    models = adjoint_models([n], A, nt)
    # the single-image backward recomputes the forward: the line kernel runs
    # twice per step
    models["sino_t05"] *= 2
    out = base_line(args, lines / (ms * 1e-3), ms, ms_prof, {
        "workload": args.workload,
        "image": f"{side}x{side} {w['kind']} seed {w['seed']}, target seed "
                 f"{w['target_seed']}", "padded_n": n, "angles": A,
        "angle_step_deg": w["step"], "tfunctionals": w["tcodes"],
        "pfunctionals": [p[0] for p in w["pfuncs"]], "lines_per_step": lines,
        "step": "forward + loss gradient + backward (the backward recomputes the forward)"})
    out["roofline"] = roofline_of(kern, args.steps, models)
    if out["roofline"]:
        out["roofline"]["traffic"] = load_traffic(args.workload, out["roofline"]["kind"])
    out["step_roofline"] = step_roofline(adjoint_bytes(n, A, nt, npf), ms,
                                         "three 4 n^2 passes per angle")
    out["kernel_ms_per_step"] = per_kernel(kern, args.steps)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tt_oracle
        p_img = tt_oracle.pad(img_cm.ravel(), side, side)
        p_tgt = tt_oracle.pad(tgt_cm.ravel(), side, side)
        s2, _ = tt_oracle.sinograms(p_tgt, n, A, w["step"], w["tcodes"], cpu_threads())
        tgt = np.stack([tt_oracle.circus(s2[t], n, A, w["pfuncs"], cpu_threads())[0]
                        for t in range(nt)]).astype(np.float64)
        out["cpu_baseline"] = cpu_baseline_adjoint([p_img] * 6, [tgt] * 6, [n] * 6, w, A,
                                                   budget_s=0.0)
    return out


def run_batched_adjoint(args, w):
    """Batched forward + adjoint replays beside the plain per-image path."""
    import numpy as np
    import torch
    from tracetransform_amd import core

    sides, imgs_cm, imgs = batch_images(w)
    dev = imgs[0].device
    tgts = []
    for k, sd in enumerate(sides):
        b = make_image(w, side=int(sd), seed=w["seed"] * 1000 + 500 + k)
        tgts.append(core.pad(torch.from_numpy(np.ascontiguousarray(b)).to(dev),
                             int(sd), int(sd)))
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
    ms, ms_prof, kern = timed(ctx, step, args.steps, args.warmup)
    models = adjoint_models(ns, A, nt)
    out = base_line(args, lines / (ms * 1e-3), ms, ms_prof, {
        "workload": args.workload, "images": len(ns),
        "sides": f"{int(min(sides))}..{int(max(sides))}",
        "padded_n": f"{min(ns)}..{max(ns)}", "angles": A, "angle_step_deg": w["step"],
        "tfunctionals": w["tcodes"], "pfunctionals": [p[0] for p in w["pfuncs"]],
        "lines_per_step": lines,
        "step": "batched forward (keeps the line data) + loss gradient + batched "
                "backward (reads it: no forward recompute), graph replays",
        "launch_shape": list(fwd.launch_shape()[:2]), "tiles": core.batched_tiles(0),
        "autotune": bool(args.autotune)})
    out["roofline"] = roofline_of(kern, args.steps, models)
    if out["roofline"]:
        out["roofline"]["traffic"] = load_traffic(args.workload, out["roofline"]["kind"])
    three = sum(adjoint_bytes(n, A, nt, npf) for n in ns)
    two = sum(adjoint_bytes(n, A, nt, npf, passes=2) for n in ns)
    out["step_roofline"] = step_roofline(three, ms, "three 4 n^2 passes per angle "
                                         "(forward, recompute, transposed rotation)")
    out["step_roofline"]["frac_two_passes"] = round(two / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    out["kernel_ms_per_step"] = per_kernel(kern, args.steps)
    if not args.no_plain:
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

        steps_plain = max(2, args.steps // 2)
        torch.cuda.synchronize()
        plain()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps_plain):
            plain()
        torch.cuda.synchronize()
        ms_plain = (time.perf_counter() - t0) * 1e3 / steps_plain
        out["plain_per_image"] = {"ms_per_step": round(ms_plain, 4),
                                  "value": round(lines / (ms_plain * 1e-3), 1),
                                  "batched_speedup": round(ms_plain / ms, 3)}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import tt_oracle
        tgt_np = target.cpu().numpy().astype(np.float64)
        p_imgs = [tt_oracle.pad(a.ravel(), int(sd), int(sd)) for a, sd in
                  zip(imgs_cm[:12], sides[:12])]
        out["cpu_baseline"] = cpu_baseline_adjoint(p_imgs, list(tgt_np[:12]), ns[:12], w, A)
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
