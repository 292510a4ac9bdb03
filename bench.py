#!/usr/bin/env python3
"""Headline benchmark: DUGKS step throughput (MLUPS) and HBM roofline.

Workload: the large single-GPU Re_tau = 180 channel, 256 x 192 x 256 cells,
tanh-stretched wall-normal (y) mesh, synthetic log-law + Fourier initial
state, constant-flow-rate forcing -- stepped in FP32 (value), FP32 compute
with FP16-deviation storage, and FP64.  A "step" is one DUGKS time step of
the whole grid.  Inputs are resident in HBM before timing starts.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dugks-gpu_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public spec
BYTES_PER_VALUE = {"fp64": 8, "fp32": 4, "fp16": 2}
METRIC = "MLUPS & %HBM peak: FP64 vs FP32/FP16-store at stated error; adjoint/fwd cost"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_mode(dugks_amd, configs, torch, precision, args):
    prec = {"fp64": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32,
            "fp16": dugks_amd.PREC_FP16S}[precision]
    kernel = {"auto": dugks_amd.KERNEL_AUTO, "fused": dugks_amd.KERNEL_FUSED,
              "twopass": dugks_amd.KERNEL_TWOPASS}[args.kernel]
    text = configs.large_channel(precision=precision)
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.time()
    eng = dugks_amd.Engine(text, precision=prec, kernel=kernel, device=0, stream=stream)
    log(f"{precision}: init {time.time() - t0:.1f}s, grid {eng.shape}")
    ncell = eng.ncell
    if args.warmup:
        eng.step(args.warmup, istep0=1)
    torch.cuda.synchronize()
    # timed region: K steps, synchronised on both sides
    t1 = time.perf_counter()
    eng.step(args.steps, istep0=args.warmup + 1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    # kernel time of the step kernel, hipEvents on the launch stream
    kt = eng.time_steps(min(args.steps, 50), istep0=args.warmup + args.steps + 1)
    torch.cuda.synchronize()
    S = BYTES_PER_VALUE[precision]
    alg_bytes = 38 * S * ncell  # read f once + write f once, 19 values each
    mlups = ncell * args.steps / wall / 1e6
    achieved = alg_bytes / (kt["mean_ms"] * 1e-3) / 1e9
    s = eng.global_sums()
    ok = all(math.isfinite(v) for v in s) and s[0] > 0
    res = dict(mlups=mlups, ms_per_step=wall * 1e3 / args.steps, kernel_ms=kt["mean_ms"],
               kernel_ms_min=kt["min_ms"], launches_per_step=int(kt["launches_per_step"]),
               alg_bytes_per_launch=alg_bytes, achieved_gbs=achieved,
               hbm_frac=achieved / HBM_PEAK_GBS, launches=int(kt["launches_per_step"]), finite=bool(ok))
    log(f"{precision}: {mlups:.0f} MLUPS, kernel {kt['mean_ms']:.3f} ms, {achieved:.0f} GB/s algorithmic")
    eng.close()
    return res


def cpu_baseline(configs, threads):
    """FP64 oracle (the reference algorithm restated in C++/OpenMP) on a
    z-slab of the same workload: 256 x 192 x 16 cells, same mesh and dt."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import dugks_amd
    from pyoracle import Oracle
    text = configs.turbulent_channel(nx=256, ny=192, nz=16)
    dims, p = dugks_amd.derive(text)
    w = dugks_amd.config_spacing(text, 1)
    f0 = dugks_amd.initial_state(text)
    o = Oracle(dims, (1, 0, 1), (4 * math.pi, 2.0, 4 * math.pi / 3 * 16 / 256), dt=p["dt"],
               visc=p["visc"], famb=((180 * 3.0e-5) ** 2, 0, 0), spacing=(None, w, None),
               threads=threads)
    o.set_f(f0)
    o.step(1)  # warm-up
    n = 3
    t = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t
    cells = int(np.prod(dims))
    return dict(value=cells * n / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=f"FP64 oracle, {n} steps of a 256x192x16 z-slab of the workload "
                       f"({cells} cells, same stretched mesh and dt), {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--modes", default="fp32,fp16,fp64")
    ap.add_argument("--kernel", default="auto", choices=["auto", "fused", "twopass"])
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the domain is not sharded); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import dugks_amd
    from dugks_amd import configs

    modes = [m for m in args.modes.split(",") if m]
    results = {}
    for m in modes:
        results[m] = run_mode(dugks_amd, configs, torch, m, args)
    primary = "fp32" if "fp32" in results else modes[0]
    r = results[primary]
    cpu = None
    if not args.no_cpu:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(configs, threads)
        log(f"cpu baseline {cpu['value']:.2f} MLUPS")
    out = {
        "metric": METRIC,
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f32-compute/f16-store"}[primary],
        "data": "synthetic",
        "config": {"workload": "Re_tau=180 channel 256x192x256, tanh-stretched y (alpha 0.85), "
                               "constant flow rate, log-law + Fourier initial state",
                   "grid": [256, 192, 256], "cells": 256 * 192 * 256, "precision": primary,
                   "step_kernel": "fused" if r["launches_per_step"] <= 2 else "twopass"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "modes": {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}
                  for m, res in results.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
