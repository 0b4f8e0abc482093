#!/usr/bin/env python3
"""Throughput of the trace-transform hot path on one MI355X.

A step is one forward pass of the path over one synthetic image already in
HBM: sinograms for every requested T-functional (tt_sinograms) followed by
the circus functions of every sinogram (tt_circus).  ``value`` is sinogram
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
    # One 256x256 U[0,1) image (seed 1) padded to 363, 360 angles at 1 deg,
    # T0-T5, P1-P3.
    "plain-256-uniform-T0-T5-P1-P3": dict(
        side=256, kind="uniform", seed=1, step=1.0, tcodes=[0, 1, 2, 3, 4, 5],
        pfuncs=[(1,), (2,), (3,)]),
}
DEFAULT_WORKLOAD = "plain-256-uniform-T0-T5-P1-P3"


def make_image(w):
    from tracetransform_amd import inputs
    side = w["side"]
    if w["kind"] == "uniform":
        return inputs.uniform_image(side, w["seed"])
    return inputs.blob_image(side, w["seed"])


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


def cpu_baseline(img_cm, w, n, A, iters):
    """OpenMP port of the reference pipeline on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tt_oracle
    threads = int(os.environ.get("TT_CPU_THREADS", "16"))
    threads = max(1, min(threads, os.cpu_count() or 1))
    side = w["side"]
    img_cm = tt_oracle.pad(img_cm.ravel(), side, side)
    tt_oracle.transform(img_cm, n, A, w["step"], w["tcodes"], w["pfuncs"], threads)
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        tt_oracle.transform(img_cm, n, A, w["step"], w["tcodes"], w["pfuncs"], threads)
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return {"value": A * n / med, "unit": "lines/s", "cores": threads,
            "kind": "port",
            "sample": f"full workload ({A} angles x {n} lines), 1 warm-up + "
                      f"{iters} timed runs of sinograms+circus+zscore, median"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-iters", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path "
              "shards by image or angle without a collective in the forward pass, "
              "see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from tracetransform_amd import _lib, core

    w = WORKLOADS[args.workload]
    img_cm = make_image(w)
    side = w["side"]
    n, _, _ = core.pad_geometry(side, side)
    A = core.angle_count(w["step"])
    nt, npf = len(w["tcodes"]), len(w["pfuncs"])
    dev = torch.device("cuda", 0)
    d_raw = torch.from_numpy(np.ascontiguousarray(img_cm)).to(dev)
    d_img = core.pad(d_raw, side, side)
    sino = torch.empty((nt, A, n), dtype=torch.float32, device=dev)
    circ = torch.empty((nt, npf, A), dtype=torch.float32, device=dev)
    ctx = core.get_context(0)

    def step():
        core.sinograms(d_img, A, w["step"], w["tcodes"], out=sino)
        core.circus(sino, w["pfuncs"], out=circ)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    for k in range(4):
        ctx.profile_read(k)
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.profile(False)
    kern = {name: ctx.profile_read(k) for name, k in
            (("sino_t05", _lib.TT_KERNEL_SINO_T05), ("sino_sort", _lib.TT_KERNEL_SINO_SORT),
             ("circus", _lib.TT_KERNEL_CIRCUS), ("aux", _lib.TT_KERNEL_AUX))}
    ms_step = (t1 - t0) * 1e3 / args.steps
    lines = A * n
    value = lines / (ms_step * 1e-3)

    ms_k, cnt = kern["sino_t05"]
    avg_ms = ms_k / max(cnt, 1)
    bytes_launch = sino_bytes(n, A, nt)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": load_traffic(args.workload),
                "kernel": "sino_t05_kernel", "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": bytes_launch}

    out = {
        "metric": "sinogram lines/s", "value": round(value, 1), "unit": "lines/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "image": f"{side}x{side} {w['kind']} "
                   f"seed {w['seed']}", "padded_n": n, "angles": A,
                   "angle_step_deg": w["step"], "tfunctionals": w["tcodes"],
                   "pfunctionals": [p[0] for p in w["pfuncs"]],
                   "lines_per_step": lines},
        "roofline": roofline,
        "kernel_ms_per_step": {k: round(v[0] / max(args.steps, 1), 4)
                               for k, v in kern.items() if v[1]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(img_cm, w, n, A, args.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
