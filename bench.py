#!/usr/bin/env python3
"""Headline benchmark: DUGKS step throughput (MLUPS) and HBM roofline.

Workload: the large single-GPU Re_tau = 180 channel, 256 x 192 x 256 cells,
tanh-stretched wall-normal (y) mesh, synthetic log-law + Fourier initial
state, constant-flow-rate forcing -- stepped in FP32 (value), in FP32
arithmetic with FP16 non-equilibrium storage (FP32 rho, u and FP16
f - feq) and with FP16 deviation storage, and in FP64.  A "step" is one
DUGKS time step of the whole grid.  Inputs are resident in HBM before timing
starts.

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
# stored bytes of one cell's state: 19 values, or (fp16n) rho and u in FP32
# plus 19 FP16 values
BYTES_PER_CELL = {"fp64": 152, "fp32": 76, "fp16": 38, "fp16n": 54}
METRIC = "MLUPS & %HBM peak: FP64 vs FP32/FP16-store at stated error; adjoint/fwd cost"


# sources of the step kernel: a traffic figure measured on other sources is stale
STEP_KERNEL_SOURCES = ("step_fused.hip", "fused_common.h", "device_common.h", "dugks_internal.h", "lattice.h")


def step_kernel_hash():
    """First 16 hex digits of the SHA-256 of the step kernel's sources."""
    import hashlib
    h = hashlib.sha256()
    for name in STEP_KERNEL_SOURCES:
        with open(os.path.join(ROOT, "dugks-gpu_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(precision):
    """HBM bytes per launch of the step kernel on this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed under profiles/
    (each counter scaled by its reading for a copy kernel of known size),
    with the source hash it was measured on: (bytes, hash, current).  bytes
    is None when no summary for this precision exists or when it was measured
    on other kernel sources than the current ones (current False)."""
    path = os.path.join(ROOT, "profiles", "k_fused_traffic_256x192x256.txt")
    if not os.path.exists(path):
        return None, None, False
    value, stamp = None, None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == "source_hash":
            stamp = parts[1]
        if len(parts) >= 3 and parts[0] == "traffic" and parts[1] == precision:
            value = int(float(parts[2]))
    current = stamp is not None and stamp == step_kernel_hash()
    return (value if current else None), stamp, current


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def field_stats(eng, np):
    """rho, u and the wall-normal (y) profiles of mean u_x, rms velocities and
    <u'v'> of the current state, FP64 on the host."""
    rho, ux, uy, uz = eng.macro()  # [k][j][i]
    prof = {}
    mx = ux.mean(axis=(0, 2))
    prof["U"] = mx
    for name, a in (("urms", ux), ("vrms", uy), ("wrms", uz)):
        prof[name] = np.sqrt(np.maximum((a * a).mean(axis=(0, 2)) - a.mean(axis=(0, 2)) ** 2, 0))
    prof["uv"] = (ux * uy).mean(axis=(0, 2)) - mx * uy.mean(axis=(0, 2))
    return dict(rho=rho, u=(ux, uy, uz), prof=prof)


def field_errors(ref, got, np):
    """Errors of a reduced-precision state against the FP64 state at the same
    step: max and RMS of rho and |u| (relative to the FP64 maxima) and of the
    y-profiles (relative to each FP64 profile's maximum)."""
    d_rho = got["rho"] - ref["rho"]
    du2 = sum((g - r) ** 2 for g, r in zip(got["u"], ref["u"]))
    umax = float(np.sqrt(sum(r * r for r in ref["u"]).max()))
    out = {"rho_max": float(np.abs(d_rho).max() / np.abs(ref["rho"]).max()),
           "rho_rms": float(np.sqrt((d_rho ** 2).mean()) / np.abs(ref["rho"]).max()),
           "u_max": float(np.sqrt(du2.max()) / umax),
           "u_rms": float(np.sqrt(du2.mean()) / umax)}
    for k in ("U", "urms", "vrms", "wrms", "uv"):
        scale = float(np.abs(ref["prof"][k]).max())
        out[f"{k}_profile_max"] = float(np.abs(got["prof"][k] - ref["prof"][k]).max() / scale) if scale else 0.0
    return out


def run_mode(dugks_amd, configs, torch, precision, args, np=None):
    prec = {"fp64": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32,
            "fp16": dugks_amd.PREC_FP16S, "fp16n": dugks_amd.PREC_FP16N}[precision]
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
    alg_bytes = 2 * BYTES_PER_CELL[precision] * ncell  # read the state once + write it once
    mlups = ncell * args.steps / wall / 1e6
    achieved = alg_bytes / (kt["mean_ms"] * 1e-3) / 1e9
    s = eng.global_sums()
    ok = all(math.isfinite(v) for v in s) and s[0] > 0
    stats = field_stats(eng, np) if np is not None else None
    res = dict(mlups=mlups, ms_per_step=wall * 1e3 / args.steps, kernel_ms=kt["mean_ms"],
               kernel_ms_min=kt["min_ms"], launches_per_step=int(kt["launches_per_step"]),
               alg_bytes_per_launch=alg_bytes, achieved_gbs=achieved,
               hbm_frac=achieved / HBM_PEAK_GBS, launches=int(kt["launches_per_step"]), finite=bool(ok))
    log(f"{precision}: {mlups:.0f} MLUPS, kernel {kt['mean_ms']:.3f} ms, {achieved:.0f} GB/s algorithmic")
    eng.close()
    return res, stats


def adjoint_leg(dugks_amd, configs, torch, T, precision="fp32"):
    """Forward + adjoint over T steps of the same 256x192x256 channel at
    constant flow rate (the sweep differentiates the flow-rate controller):
    wall-shear loss, sqrt(T) checkpoints in HBM, gradient with respect to a
    per-cell body force.  Reports the per-step adjoint (face pass and gather)
    and whole-sweep time ratios to the forward, and the sweep's device
    memory."""
    prec = {"fp64": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32}[precision]
    text = configs.large_channel(precision=precision)
    eng = dugks_amd.Engine(text, precision=prec, device=0, stream=torch.cuda.current_stream().cuda_stream)
    nz, ny, nx = eng.shape[2], eng.shape[1], eng.shape[0]
    f0 = torch.empty((19, nz, ny, nx), dtype=torch.float64, device="cuda")
    eng.get_f_device(f0)
    eng.step(2, istep0=1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.step(T, istep0=3)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t
    # one adjoint step (face cotangents, gather), with lam in FP64 and, for
    # the FP32 modes, in FP32 (what the sweep uses)
    g = torch.zeros((3, nz, ny, nx), dtype=torch.float64, device="cuda")
    t_adj = {}
    for lt in (("float64", "float32") if precision != "fp64" else ("float64",)):
        lam = torch.randn((19, nz, ny, nx), dtype=getattr(torch, lt), device="cuda")
        lam2 = torch.empty_like(lam)
        try:
            eng.adjoint_step(lam, lam2, g)
        except dugks_amd.DugksError:
            continue  # FP32 lam not supported on this configuration
        torch.cuda.synchronize()
        nadj = 10
        t = time.perf_counter()
        for _ in range(nadj):
            eng.adjoint_step(lam, lam2, g)
        torch.cuda.synchronize()
        t_adj[lt] = (time.perf_counter() - t) / nadj
        del lam, lam2
    lam_type = "float32" if "float32" in t_adj else "float64"
    S = {"fp64": 8, "fp32": 4}[precision]
    lb = 4 if lam_type == "float32" else 8
    rec = 2 * 3 * 13 * S  # compact records (the z-marching gather's configurations)
    # storage in the arithmetic type: the face pass hands rho and j (4 values)
    # to the gather, which then does not read f
    state = 19 * S + 2 * 4 * S
    adj_bytes = eng.ncell * (3 * 19 * lb + state + rec + 2 * 3 * 8)
    eng.set_f_device(f0)
    del f0
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    t = time.perf_counter()
    J, peak = eng.forward_adjoint(T, g)
    torch.cuda.synchronize()
    t_fa = time.perf_counter() - t
    pool = eng.sweep_bytes()
    gmax = float(g.abs().max())
    ok = math.isfinite(J) and math.isfinite(gmax)
    eng.close()
    fwd_step = t_fwd / T
    res = dict(precision=precision, T=T, interval=int(math.ceil(math.sqrt(T))), loss="wall shear",
               forcing="constant flow rate, controller differentiated",
               forward_ms_per_step=round(fwd_step * 1e3, 4),
               adjoint_lambda=lam_type,
               adjoint_step_ms=round(t_adj[lam_type] * 1e3, 4),
               adjoint_over_forward_step=round(t_adj[lam_type] / fwd_step, 3),
               adjoint_step_ms_fp64_lambda=round(t_adj["float64"] * 1e3, 4),
               # algorithmic bytes of one adjoint step (per cell: lam' read twice and lam
               # written, f read once and rho, j written and read, 13-value face records
               # of 3 axes written and read, the FP64 force gradient read and written)
               adjoint_alg_bytes_per_step=adj_bytes,
               adjoint_achieved_gbs=round(adj_bytes / t_adj[lam_type] / 1e9, 1),
               fwd_adj_sweep_s=round(t_fa, 3),
               sweep_over_forward=round(t_fa / t_fwd, 3),
               # device memory in use at the end of the sweep (hipMemGetInfo: the
               # engine's state, face records and checkpoint pool, and torch's
               # cached blocks) and the checkpoint pool alone
               sweep_device_in_use_bytes=peak, sweep_pool_bytes=pool,
               device_free_before=free0, device_total=total,
               finite=bool(ok))
    log(f"adjoint {precision}: step {t_adj[lam_type]*1e3:.2f} ms ({lam_type} lam; {t_adj['float64']*1e3:.2f} ms with float64) "
        f"vs forward {fwd_step*1e3:.3f} ms; "
        f"{T}-step sweep {t_fa:.2f} s ({t_fa/t_fwd:.1f}x forward), pool {pool/1e9:.1f} GB, "
        f"{peak/1e9:.1f} GB in use")
    return res


def host_cpu():
    """CPU model, the physical cores of the CPUs this process may run on, and
    the host's CPU share for this job (OMP_NUM_THREADS when the launcher sets
    it): (model, physical_cores, share).  Cores are counted from the sysfs
    topology (package and core id of each allowed CPU), else from
    /proc/cpuinfo when it carries those ids, else (or when those name a
    single core for several CPUs) one per allowed CPU."""
    model = "unknown"
    info = {}
    cpu, phys, core = None, None, None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh.read().splitlines() + [""]:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    cpu = int(v)
                elif k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not line.strip() and cpu is not None:
                    info[cpu] = (phys, core)
                    cpu, phys, core = None, None, None
    except OSError:
        pass
    allowed = sorted(os.sched_getaffinity(0))

    def sysfs(c):
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        with open(base + "physical_package_id") as a, open(base + "core_id") as b:
            return a.read().strip(), b.read().strip()

    try:
        cores = len({sysfs(c) for c in allowed})
    except OSError:
        ids = [info.get(c, (None, None)) for c in allowed]
        if ids and all(p is not None and k is not None for p, k in ids):
            cores = len(set(ids))
        else:
            cores = len(allowed)
    if cores == 1 and len(allowed) > 1:
        # a virtual machine that reports one core id for every CPU
        cores = len(allowed)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return model, cores, share


def cpu_baseline(configs, threads, cpu_host, target_s=15.0):
    """FP64 oracle (the reference algorithm restated in C++/OpenMP) on a
    z-slab of the same workload: 256 x 192 x 32 cells, same mesh and dt,
    stepped in batches of 5 until about target_s seconds of CPU work, one
    OpenMP thread per physical core (OMP_PROC_BIND=close, set before the
    OpenMP runtime starts)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import dugks_amd
    from pyoracle import Oracle
    nzs = 32
    text = configs.turbulent_channel(nx=256, ny=192, nz=nzs)
    dims, p = dugks_amd.derive(text)
    w = dugks_amd.config_spacing(text, 1)
    f0 = dugks_amd.initial_state(text)
    o = Oracle(dims, (1, 0, 1), (4 * math.pi, 2.0, 4 * math.pi / 3 * nzs / 256), dt=p["dt"],
               visc=p["visc"], famb=((180 * 3.0e-5) ** 2, 0, 0), spacing=(None, w, None),
               threads=threads)
    o.set_f(f0)
    o.step(1)  # warm-up
    n = 0
    t = time.perf_counter()
    while time.perf_counter() - t < target_s and n < 400:
        o.step(5)
        n += 5
    dt = time.perf_counter() - t
    cells = int(np.prod(dims))
    model, phys, share = cpu_host
    return dict(value=cells * n / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                cpu_model=model, physical_cores=phys, cpu_share=share or None,
                omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                sample=f"FP64 oracle, {n} steps of a 256x192x{nzs} z-slab of the workload "
                       f"({cells} cells, same stretched mesh and dt), {dt:.1f} s, "
                       f"{threads} OpenMP threads on {model}")


def study_tgv(dugks_amd, configs, torch, np, steps, every=10):
    """3D Taylor-Green vortex, 128^3 periodic, u0 = 0.05, Re = 1600: kinetic
    energy E(t) (the device loss functional, 0.5 <rho |u|^2>) every `every`
    steps and the dissipation rate eps = -dE/dt (central differences), FP32
    against FP64."""
    series = {}
    rates = {}
    for m in ("fp64", "fp32"):
        prec = {"fp64": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32}[m]
        eng = dugks_amd.Engine(configs.taylor_green(n=128, precision=m, nsteps=steps), precision=prec, device=0,
                               stream=torch.cuda.current_stream().cuda_stream)
        E = [eng.loss(dugks_amd.LOSS_KINETIC_ENERGY)]
        t0 = time.perf_counter()
        for k in range(steps // every):
            eng.step(every, istep0=1 + k * every)
            E.append(eng.loss(dugks_amd.LOSS_KINETIC_ENERGY))
        torch.cuda.synchronize()
        rates[m] = eng.ncell * steps / (time.perf_counter() - t0) / 1e6
        dt = eng.params["dt"]
        eng.close()
        series[m] = np.array(E), dt
        log(f"tgv {m}: E {E[0]:.6e} -> {E[-1]:.6e}")
    E64, dt = series["fp64"]
    E32, _ = series["fp32"]
    eps64 = -(E64[2:] - E64[:-2]) / (2 * every * dt)
    eps32 = -(E32[2:] - E32[:-2]) / (2 * every * dt)
    return {"workload": "Taylor-Green vortex 128^3, u0=0.05, Re=1600", "steps": steps,
            "sample_every": every,
            "E_max_rel_err": float(np.abs(E32 - E64).max() / np.abs(E64).max()),
            "eps_max_rel_err": float(np.abs(eps32 - eps64).max() / np.abs(eps64).max()),
            "E_final": {"fp64": float(E64[-1]), "fp32": float(E32[-1])},
            "eps_final": {"fp64": float(eps64[-1]), "fp32": float(eps32[-1])},
            "mlups_incl_sampling": {k: round(v, 1) for k, v in rates.items()}}


def study_channel(dugks_amd, configs, torch, np, steps):
    """Re_tau = 180 channel 128^3 (tanh-stretched y, constant flow rate,
    synthetic log-law + Fourier state): mean-profile, rms and Reynolds-stress
    errors of FP32, FP16 non-equilibrium and FP16 deviation storage against
    FP64 after `steps` steps."""
    stats, rates = {}, {}
    for m in ("fp64", "fp64_twin", "fp32", "fp16n", "fp16"):
        prec = {"fp64": dugks_amd.PREC_FP64, "fp64_twin": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32,
                "fp16": dugks_amd.PREC_FP16S, "fp16n": dugks_amd.PREC_FP16N}[m]
        eng = dugks_amd.Engine(configs.turbulent_channel(precision=m.split("_")[0], nsteps=steps), precision=prec,
                               device=0, stream=torch.cuda.current_stream().cuda_stream)
        if m == "fp64_twin":
            # the FP64 run from the initial state perturbed at the size of an
            # FP32 rounding (3e-8 relative, seeded): how far the flow itself
            # carries a rounding-sized difference over the run
            f0 = torch.empty(19 * eng.ncell, dtype=torch.float64, device="cuda")
            eng.get_f_device(f0)
            gen = torch.Generator(device="cuda")
            gen.manual_seed(20261203)
            f0 *= 1 + 3e-8 * torch.randn(f0.shape, generator=gen, dtype=torch.float64, device="cuda")
            eng.set_f_device(f0)
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(1000, steps - done)
            eng.step(k, istep0=1 + done)
            done += k
            torch.cuda.synchronize()
            log(f"channel {m}: {done} steps")
        rates[m] = eng.ncell * steps / (time.perf_counter() - t0) / 1e6
        stats[m] = field_stats(eng, np)
        eng.close()
    out = {"workload": "Re_tau=180 channel 128^3, tanh y (alpha 0.85), constant flow rate", "steps": steps,
           "against": "fp64", "mlups": {k: round(v, 1) for k, v in rates.items()}}
    for m in ("fp64_twin", "fp32", "fp16n", "fp16"):
        out[m] = {k: float(f"{v:.3e}") for k, v in field_errors(stats["fp64"], stats[m], np).items()}
    return out


def study_recon(dugks_amd, configs, torch, np, steps):
    """Trilinear against gradient reconstruction on the 256 x 192 x 256
    Re_tau = 180 channel with a uniform wall-normal mesh (the gradient form
    needs uniform axes), on the fused and the two-pass kernels: time per
    step over `steps` steps from the synthetic state, the largest relative
    difference between the FP64 fused and two-pass states after them, and
    the FP32 and FP16 non-equilibrium field errors against FP64."""
    prec = {"fp64": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32, "fp16n": dugks_amd.PREC_FP16N}
    kern = {"fused": dugks_amd.KERNEL_FUSED, "twopass": dugks_amd.KERNEL_TWOPASS}
    out = {"workload": "Re_tau=180 channel 256x192x256, uniform y, constant flow rate", "steps": steps,
           "runs": {}, "fused_vs_twopass_fp64": {}, "errors_vs_fp64": {}}
    for recon in ("tri", "grad"):
        stats, f64 = {}, {}
        for m in ("fp64", "fp32", "fp16n"):
            for k in ("fused", "twopass"):
                if k == "twopass" and m == "fp16n":
                    continue
                text = configs.large_channel(precision=m, nsteps=steps, mesh_y="uniform", recon=recon)
                eng = dugks_amd.Engine(text, precision=prec[m], kernel=kern[k], device=0,
                                       stream=torch.cuda.current_stream().cuda_stream)
                t = eng.time_steps(steps)
                ms = t["total_ms"] / steps
                out["runs"][f"{recon}/{m}/{k}"] = {"ms_per_step": round(ms, 4),
                                                   "glups": round(eng.ncell / ms / 1e6, 3)}
                log(f"recon {recon} {m} {k}: {ms:.3f} ms per step, {eng.ncell / ms / 1e6:.2f} GLUPS")
                if m == "fp64":
                    f64[k] = torch.empty(19 * eng.ncell, dtype=torch.float64, device="cuda")
                    eng.get_f_device(f64[k])
                if k == "fused":
                    stats[m] = field_stats(eng, np)
                eng.close()
        d = (f64["fused"] - f64["twopass"]).abs().max() / f64["twopass"].abs().max()
        out["fused_vs_twopass_fp64"][recon] = float(f"{float(d):.3e}")
        f64.clear()
        out["errors_vs_fp64"][recon] = {m: {k: float(f"{v:.3e}") for k, v in
                                            field_errors(stats["fp64"], stats[m], np).items()}
                                        for m in ("fp32", "fp16n")}
    return out


def study_adjoint(dugks_amd, configs, torch, np, T, ndirs=32, flow=False):
    """Adjoint workload: transitional channel 64x64x32 (seeded perturbed
    laminar state, Re = 2000), J = time-mean wall shear stress over T steps,
    control G ~ N(0, 1e-6) per cell (seeded).  sqrt(T)-checkpointed
    gradient in FP64 and FP32, against central differences of J along
    `ndirs` seeded random directions (eps per precision), FP32 against FP64,
    and the timing of the sweep against T forward steps.  flow: driven at
    constant flow rate instead of constant force (the sweep differentiates
    the controller; every run restarts it from the same force)."""
    kw = dict(flow_rate="auto", flow_axis=0) if flow else {}
    out = {"workload": "transitional channel 64x64x32, Re=2000, wall-shear loss, "
                       + ("constant flow rate" if flow else "constant force"), "T": T, "directions": ndirs}
    grads, fds = {}, {}
    for m, eps in (("fp64", 1.0), ("fp32", 100.0)):
        prec = {"fp64": dugks_amd.PREC_FP64, "fp32": dugks_amd.PREC_FP32}[m]
        eng = dugks_amd.Engine(configs.adjoint_channel(precision=m, **kw), precision=prec, device=0,
                               stream=torch.cuda.current_stream().cuda_stream)
        nx, ny, nz = eng.shape
        f0 = torch.empty((19, nz, ny, nx), dtype=torch.float64, device="cuda")
        eng.get_f_device(f0)
        F0 = eng.ambient_force()
        G = 1e-6 * np.random.default_rng(2024).standard_normal((3, nz, ny, nx))
        Gd = torch.from_numpy(G).cuda()
        gd = torch.empty_like(Gd)

        def restart(ctrl):
            eng.set_ambient_force(F0)
            eng.set_f_device(f0)
            eng.set_cell_force_device(ctrl)

        def run_J(ctrl):
            restart(ctrl)
            tot = 0.0
            for n in range(T):
                eng.step(1, istep0=1 + n)
                tot += eng.loss(dugks_amd.LOSS_WALL_SHEAR)
            return tot / T

        restart(Gd)
        torch.cuda.synchronize()
        t = time.perf_counter()
        J, peak = eng.forward_adjoint(T, gd)
        torch.cuda.synchronize()
        t_sweep = time.perf_counter() - t
        restart(Gd)
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.step(T, istep0=1)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t
        grad = gd.cpu().numpy()
        grads[m] = grad
        rng = np.random.default_rng(7)
        errs = []
        fds[m] = []
        for d in range(ndirs):
            v = 1e-6 * rng.standard_normal(G.shape)
            vd = torch.from_numpy(v).cuda()
            jp = run_J(Gd + eps * vd)
            jm = run_J(Gd - eps * vd)
            fd = (jp - jm) / (2 * eps)
            an = float((grad * v).sum())
            errs.append(abs(fd - an) / abs(an))
            fds[m].append((fd, v))
        out[m] = {"J": J, "eps_fd": eps, "fd_rel_err_max": float(max(errs)),
                  "fd_rel_err_median": float(np.median(errs)), "sweep_s": round(t_sweep, 3),
                  "forward_s": round(t_fwd, 4), "sweep_over_forward": round(t_sweep / t_fwd, 2),
                  "sweep_device_bytes": peak}
        log(f"adjoint study {m}: FD max rel err {max(errs):.2e}, sweep {t_sweep:.2f} s "
            f"({t_sweep / t_fwd:.1f}x forward)")
        eng.close()
    g64, g32 = grads["fp64"], grads["fp32"]
    out["fp32_vs_fp64_grad_max_rel"] = float(np.abs(g32 - g64).max() / np.abs(g64).max())
    # the FP32 gradient against the FP64 central differences (the FP32 forward
    # is too noisy for its own differences at this perturbation size)
    out["fp32_grad_vs_fp64_fd_max_rel"] = float(max(abs(fd - float((g32 * v).sum())) / abs(fd)
                                                    for fd, v in fds["fp64"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--modes", default="fp32,fp16n,fp16,fp64")
    ap.add_argument("--kernel", default="auto", choices=["auto", "fused", "twopass"])
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0: one per physical core of this "
                         "process's CPU set, capped at the job's CPU share OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--adj-steps", type=int, default=1000,
                    help="steps of the forward+adjoint sweep on the large channel (0: skip)")
    ap.add_argument("--adj-modes", default="fp32,fp64",
                    help="arithmetic of the adjoint legs (the first is reported as 'adjoint', the "
                         "others as 'adjoint_<mode>')")
    ap.add_argument("--study", choices=["tgv", "channel", "adjoint", "adjoint_flow", "recon"],
                    help="instead of the benchmark: the reduced-precision error study on the "
                         "Taylor-Green vortex or the 128^3 channel, the adjoint gradient study "
                         "at constant force or constant flow rate, or trilinear against gradient "
                         "reconstruction on the uniform 256x192x256 channel (prints one JSON line)")
    ap.add_argument("--study-steps", type=int, default=0,
                    help="steps of the study (default 1000 / 10000 / 400 / 100)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB read + 1 GiB write elementwise kernel 5 times "
                         "(the reference point for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the domain is not sharded); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    # the host's cores, read before any OpenMP runtime starts: with
    # OMP_PROC_BIND set, the first parallel region (torch's or the oracle's,
    # which share libgomp) binds this thread to one core, and the affinity
    # mask then names that core only
    cpu_host = host_cpu()
    # the CPU baseline's OpenMP threads stay on their cores; libgomp reads
    # this when it starts, which is after this point
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import torch
    if args.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(5):
            torch.add(x, 0.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"read_bytes": x.numel() * 4, "write_bytes": y.numel() * 4,
                                               "launches": 5}}), flush=True)
        return
    import dugks_amd
    from dugks_amd import configs
    if args.study:
        import numpy as np
        if args.study == "tgv":
            res = study_tgv(dugks_amd, configs, torch, np, args.study_steps or 1000)
        elif args.study == "recon":
            res = study_recon(dugks_amd, configs, torch, np, args.study_steps or 100)
        elif args.study in ("adjoint", "adjoint_flow"):
            res = study_adjoint(dugks_amd, configs, torch, np, args.study_steps or 400,
                                flow=args.study == "adjoint_flow")
        else:
            res = study_channel(dugks_amd, configs, torch, np, args.study_steps or 10000)
        print(json.dumps({"study": args.study, **res}), flush=True)
        return

    import numpy as np
    modes = [m for m in args.modes.split(",") if m]
    results, stats = {}, {}
    want_err = "fp64" in modes and len(modes) > 1 and not args.no_errors
    for m in modes:
        results[m], stats[m] = run_mode(dugks_amd, configs, torch, m, args, np if want_err else None)
    errors = None
    if want_err:
        nstep = args.warmup + args.steps + min(args.steps, 50)
        errors = {"steps": nstep, "against": "fp64"}
        for m in modes:
            if m != "fp64":
                errors[m] = {k: float(f"{v:.3e}") for k, v in field_errors(stats["fp64"], stats[m], np).items()}
                log(f"{m} vs fp64 after {nstep} steps: u max {errors[m]['u_max']:.2e}, "
                    f"U profile {errors[m]['U_profile_max']:.2e}")
    stats.clear()
    adjoint, adjoint_more = None, {}
    if args.adj_steps > 0:
        for i, am in enumerate(m for m in args.adj_modes.split(",") if m):
            res_a = adjoint_leg(dugks_amd, configs, torch, args.adj_steps, precision=am)
            if i == 0:
                adjoint = res_a
            else:
                adjoint_more["adjoint_" + am] = res_a
    primary = "fp32" if "fp32" in results else modes[0]
    r = results[primary]
    traffic = measured_traffic(primary)
    cpu = None
    if not args.no_cpu:
        _, phys, share = cpu_host
        threads = args.cpu_threads or (min(phys, share) if share else phys)
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(configs, threads, cpu_host)
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
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f32-compute/f16-store",
                  "fp16n": "f32-compute/f32-moments+f16-neq-store"}[primary],
        "data": "synthetic",
        "config": {"workload": "Re_tau=180 channel 256x192x256, tanh-stretched y (alpha 0.85), "
                               "constant flow rate, log-law + Fourier initial state",
                   "grid": [256, 192, 256], "cells": 256 * 192 * 256, "precision": primary,
                   "step_kernel": "fused" if r["launches_per_step"] <= 2 else "twopass"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic[0],
                     "traffic_source_hash": traffic[1], "kernel_source_hash": step_kernel_hash()},
        "cpu_baseline": cpu,
        "precision_errors": errors,
        "adjoint": adjoint,
        **adjoint_more,
        "modes": {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}
                  for m, res in results.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
