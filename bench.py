"""Benchmark: batched forward + adjoint gradient of the 2D acoustic propagator.

Workload (one "step" = one full gradient evaluation):
  350 x 1700 grid (+20 absorbing layers per side -> 390 x 1740 padded), dx 5 m,
  dt 0.5 ms, 4000 time steps, 64 shots evenly spaced along the surface (z = 2),
  1700 receivers at z = 2, 15 Hz Ricker, velocity = depth-linear 1500-4500 m/s
  background + anisotropic Gaussian random field (lh 300 m, lv 50 m, std 200,
  seed 0) clipped to [1500, 4700], Gardner density, L2 misfit against data
  modelled in a smoothed copy of the model, forward checkpoints every 50 steps.
  The step is forward (seismograms + checkpoints) + misfit + hand-written
  adjoint (recompute + backward sweep + imaging condition) -> vp.grad.

Metric: Gcell-updates/s = shots * padded cells * nt / step time.

Also reported (one JSON line): the roofline of the dominant kernel measured
live with events on the launch stream (in one instrumented step after the
timed ones), the job-level HBM roofline fraction at the algorithmic 80 + 24/k
bytes per cell-update, the reference CPU path (the PyTorch restatement of the
reference operator with autograd, on the host's cores) timed on a bounded
sample of the same workload, and beside it the FP32 C++ oracle.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s fwd+adjoint (64-shot batch vs per-shot) and % of HBM roofline"
_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """One line on stderr when a phase of the run ends (the JSON line stays
    the only output on stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def workload(n_shots=64, nz=350, nx=1700, nt=4000, nabc=20, dx=5.0, dt=5e-4, f0=15.0):
    from geobrain_amd import synthetic
    from geobrain_amd.physics import wave
    vp = synthetic.grf_model(nz, nx, dx, seed=0)
    return dict(
        nz=nz, nx=nx, nt=nt, nabc=nabc, dx=dx, dt=dt, vp=vp, vp_obs=synthetic.smooth_model(vp, 0.05),
        rho=synthetic.gardner(vp),
        damp=wave.bc_pml(nx, nz, dx, dx, nabc, float(vp.max()), False).astype(np.float32),
        wavelet=np.tile(synthetic.ricker_padded(f0, dt, nt), (n_shots, 1)),
        src_x=synthetic.surface_shots(nx, n_shots), src_z=np.full(n_shots, 2, np.int64),
        rcv_x=np.arange(nx, dtype=np.int64), rcv_z=np.full(nx, 2, np.int64))


def eager_forward(nz, nx, nabc, free_surface, dt, dz, src_z, src_x, src_v, rcv_z, rcv_x,
                  damp, vp, rho, checkpoint_segments=1):
    """Plain PyTorch restatement of the reference operator (one ATen launch per
    slice operation, autograd through reentrant torch.utils.checkpoint chunks),
    as geobrain/physics/wave/kernels/acoustic_2d.py:123-501 runs it.  It is the
    comparison baseline timed beside the HIP path ("plain per-shot GPU path");
    the product never calls it."""
    import torch
    from torch.utils.checkpoint import checkpoint
    dev = vp.device
    NZ, NX = nz + 2 * nabc, nx + 2 * nabc
    f = nabc if free_surface else 1
    ii = torch.arange(NZ, device=dev).sub(nabc).clamp(0, nz - 1)
    jj = torch.arange(NX, device=dev).sub(nabc).clamp(0, nx - 1)
    c = vp[ii][:, jj]
    den = rho[ii][:, jj]
    a1 = den * c * c * dt / dz
    k1 = damp * dt
    a2 = dt / (den * dz)
    k2 = torch.zeros_like(damp)
    k2[:, 1:NX - 2] = 0.5 * (damp[:, 1:NX - 2] + damp[:, 2:NX - 1]) * dt
    k3 = torch.zeros_like(damp)
    k3[f:NZ - 2, :] = 0.5 * (damp[f:NZ - 2, :] + damp[f + 1:NZ - 1, :]) * dt
    S = src_z.shape[0]
    sz, sx = src_z.to(dev) + nabc, src_x.to(dev) + nabc
    rz, rx = rcv_z.to(dev) + nabc, rcv_x.to(dev) + nabc
    shots = torch.arange(S, device=dev)
    c1, c2 = 9.0 / 8.0, -1.0 / 24.0
    P = (slice(None), slice(f + 1, NZ - 2), slice(2, NX - 2))
    X = (slice(None), slice(f, NZ - 1), slice(1, NX - 2))
    Z = (slice(None), slice(f, NZ - 2), slice(1, NX - 1))
    inner = (slice(nabc, nabc + nz), slice(nabc, nabc + nx))

    def chunk_fn(a1, a2, k1, k2, k3, w, p, vx, vz):
        p, vx, vz = p.clone(), vx.clone(), vz.clone()
        n = w.shape[-1]
        rec = [torch.zeros((S, n, rz.shape[0]), device=dev) for _ in range(3)]
        en = [torch.zeros((nz, nx), device=dev) for _ in range(3)]
        for it in range(n):
            p[P] = (1.0 - k1[P[1:]]) * p[P] - a1[P[1:]] * (
                c1 * (vx[:, f + 1:NZ - 2, 2:NX - 2] - vx[:, f + 1:NZ - 2, 1:NX - 3]
                      + vz[:, f + 1:NZ - 2, 2:NX - 2] - vz[:, f:NZ - 3, 2:NX - 2])
                + c2 * (vx[:, f + 1:NZ - 2, 3:NX - 1] - vx[:, f + 1:NZ - 2, 0:NX - 4]
                        + vz[:, f + 2:NZ - 1, 2:NX - 2] - vz[:, f - 1:NZ - 4, 2:NX - 2]))
            upd = dt * (w[it] if w.dim() == 1 else w[:, it])
            p[shots, sz, sx] = p[shots, sz, sx] + upd
            if free_surface:
                p[:, f - 1, :] = -p[:, f + 1, :]
            vx[X] = (1.0 - k2[X[1:]]) * vx[X] - a2[X[1:]] * (
                c1 * (p[:, f:NZ - 1, 2:NX - 1] - p[:, f:NZ - 1, 1:NX - 2])
                + c2 * (p[:, f:NZ - 1, 3:NX] - p[:, f:NZ - 1, 0:NX - 3]))
            vz[Z] = (1.0 - k3[Z[1:]]) * vz[Z] - a2[Z[1:]] * (
                c1 * (p[:, f + 1:NZ - 1, 1:NX - 1] - p[:, f:NZ - 2, 1:NX - 1])
                + c2 * (p[:, f + 2:NZ, 1:NX - 1] - p[:, f - 1:NZ - 3, 1:NX - 1]))
            if free_surface:
                vz[:, f - 1, :] = vz[:, f, :]
            for r, fld in zip(rec, (p, vx, vz)):
                r[:, it, :] = fld[:, rz, rx]
            for e, fld in zip(en, (p, vx, vz)):
                e += torch.sum(fld * fld, dim=0)[inner].detach()
        return (p, vx, vz, *rec, *en)

    p = torch.zeros((S, NZ, NX), device=dev)
    vx = torch.zeros((S, NZ, NX - 1), device=dev)
    vz = torch.zeros((S, NZ - 1, NX), device=dev)
    recs, ens = [[], [], []], [torch.zeros((nz, nx), device=dev) for _ in range(3)]
    for w in torch.chunk(src_v, checkpoint_segments, dim=-1):
        if checkpoint_segments > 1 and torch.is_grad_enabled():
            res = checkpoint(chunk_fn, a1, a2, k1, k2, k3, w, p, vx, vz, use_reentrant=True)
        else:
            res = chunk_fn(a1, a2, k1, k2, k3, w, p, vx, vz)
        p, vx, vz = res[:3]
        for lst, r in zip(recs, res[3:6]):
            lst.append(r)
        for e, r in zip(ens, res[6:9]):
            e += r.detach()
    out = {k: torch.cat(v, dim=1) for k, v in zip(("p", "vx", "vz"), recs)}
    for k, e in zip(("forward_wavefield_p", "forward_wavefield_vx", "forward_wavefield_vz"), ens):
        out[k] = e
    return out


def host_cpu_info():
    """Host CPU facts for the baseline record: model name, logical CPUs, physical
    cores (distinct (package, core) pairs in /proc/cpuinfo) and the CPUs this
    process may run on (its affinity mask, which a container's CPU share sets)."""
    model, cores = None, set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model is None:
                    model = val
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    core = val
                elif not key and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(),
            "physical_cores": len(cores) or None, "allowed_cpus": allowed}


def cpu_threads(info):
    """Threads for the CPU baselines: one per physical core this process may
    use, capped by OMP_NUM_THREADS when the environment sets it."""
    n = min(v for v in (info["physical_cores"], info["allowed_cpus"]) if v)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(w, shots=8, steps=100, k=50):
    """The reference CPU path: ``eager_forward`` (the PyTorch restatement of
    acoustic_2d.py:123-501, one ATen op per slice operation) on the host CPU
    with autograd through reentrant checkpoint chunks of k steps, as
    acoustic_2d.py:448-460 runs it, for a bounded sample of the benchmark
    workload (the first `shots` shots x the first `steps` steps); the rate is
    Gcell-updates/s of forward + gradient, which the sample measures directly
    (the whole 64-shot x 4000-step job at that rate is stated beside it)."""
    import torch
    info = host_cpu_info()
    n = cpu_threads(info)
    prev = torch.get_num_threads()
    torch.set_num_threads(n)
    try:
        T = lambda a, **kw: torch.as_tensor(a, **kw)  # noqa: E731
        d_obs = None
        args = dict(nz=w["nz"], nx=w["nx"], nabc=w["nabc"], free_surface=False, dt=w["dt"],
                    dz=w["dx"], src_z=T(w["src_z"][:shots]), src_x=T(w["src_x"][:shots]),
                    src_v=T(w["wavelet"][:shots, :steps]), rcv_z=T(w["rcv_z"]),
                    rcv_x=T(w["rcv_x"]), damp=T(w["damp"]), rho=T(w["rho"]),
                    checkpoint_segments=max(1, steps // k))
        with torch.no_grad():
            d_obs = eager_forward(vp=T(w["vp_obs"]), **args)["p"]

        def once():
            vp = T(w["vp"]).requires_grad_(True)
            t0 = time.perf_counter()
            out = eager_forward(vp=vp, **args)
            ((out["p"] - d_obs) ** 2).sum().backward()
            return time.perf_counter() - t0

        once()  # warm-up: thread pool and first-touch allocation
        runs = [once() for _ in range(3)]
        t = float(np.median(runs))
    finally:
        torch.set_num_threads(prev)
    NZ, NX = w["nz"] + 2 * w["nabc"], w["nx"] + 2 * w["nabc"]
    rate = shots * NZ * NX * steps / t
    full = w["wavelet"].shape[0] * NZ * NX * w["wavelet"].shape[1]
    return {"value": round(rate / 1e9, 5), "unit": "Gcell-updates/s", "cores": n,
            "kind": "reference",
            "sample": f"PyTorch eager restatement of the reference operator (acoustic_2d.py) on "
                      f"the host CPU, forward + autograd gradient through checkpoint chunks of "
                      f"{k} steps, first {shots} of {w['wavelet'].shape[0]} shots x first "
                      f"{steps} of {w['wavelet'].shape[1]} steps of the benchmark workload, "
                      f"median of 3 timed runs after a warm-up ({t:.2f} s each); the full job "
                      f"would take {full / rate / 3600:.2f} h at this rate (extrapolated)",
            "runs_s": [round(r, 3) for r in runs],
            "cpu": info, "torch_threads": n}


def cpu_oracle_baseline(w, steps=2500, shots=1):
    """The FP32 C++ oracle (OpenMP) gradient on `shots` shot(s) x `steps` time
    steps of the same grid: the build's own CPU port, reported beside the
    reference path, never as it."""
    import oracle
    oracle.build()
    sl = slice(0, shots)
    args = dict(nz=w["nz"], nx=w["nx"], nabc=w["nabc"], free_surface=False, dt=w["dt"],
                dz=w["dx"], src_z=w["src_z"][sl], src_x=w["src_x"][sl],
                src_v=w["wavelet"][sl, :steps], rcv_z=w["rcv_z"], rcv_x=w["rcv_x"],
                vp=w["vp"], rho=w["rho"], damp=w["damp"])
    t0 = time.perf_counter()
    f = oracle.forward(**args, precision="f32")["p"]
    g = np.random.default_rng(0).standard_normal(f.shape).astype(np.float32)
    oracle.gradient(**args, g_p=g, chunk=50, precision="f32")
    t = time.perf_counter() - t0
    cells = shots * (w["nz"] + 2 * w["nabc"]) * (w["nx"] + 2 * w["nabc"]) * steps
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(cells / t / 1e9, 5), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"FP32 C++ oracle (OpenMP), forward + checkpointed adjoint gradient, "
                      f"{shots} shot x {steps} steps of the {w['nz']}x{w['nx']} grid, "
                      f"{t:.1f} s wall"}


def per_shot_gpu_baseline(w, d_obs, steps=200, k=50):
    """The plain per-shot GPU path: the PyTorch restatement of the reference
    operator (eager ATen launches, autograd through reentrant checkpoint
    chunks of k steps) on the GPU for one shot, `steps` time steps, timed with
    events after one warm-up."""
    import torch
    dev = torch.device("cuda:0")
    T = lambda a, **kw: torch.as_tensor(a, device=dev, **kw)  # noqa: E731
    args = dict(nz=w["nz"], nx=w["nx"], nabc=w["nabc"], free_surface=False, dt=w["dt"],
                dz=w["dx"], src_z=T(w["src_z"][:1]), src_x=T(w["src_x"][:1]),
                src_v=T(w["wavelet"][:1, :steps]), rcv_z=T(w["rcv_z"]), rcv_x=T(w["rcv_x"]),
                damp=T(w["damp"]), rho=T(w["rho"]), checkpoint_segments=max(1, steps // k))
    obs = d_obs[:1, :steps]
    times = []
    for _ in range(2):
        vp = T(w["vp"]).requires_grad_(True)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = eager_forward(vp=vp, **args)
        ((out["p"] - obs) ** 2).sum().backward()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
    cells = (w["nz"] + 2 * w["nabc"]) * (w["nx"] + 2 * w["nabc"]) * steps
    return {"value": round(cells / times[-1] / 1e9, 4), "unit": "Gcell-updates/s",
            "sample": f"PyTorch eager restatement of the reference operator on the GPU, "
                      f"1 shot x {steps} steps, checkpoint every {k}, forward + autograd "
                      f"gradient, {times[-1]:.2f} s"}


def uneven_batch(steps: int, warmup: int, tune: bool, launch=None) -> dict:
    """The uneven-batch workload: 96 independent problems from seed 7 (nz in
    [200, 600], nx in [200, 1200], nt in [1000, 4000], 1-8 shots, f0 in
    [8, 20] Hz, layered Gaussian-random velocity, receivers on every column),
    forward modelling of all of them as one HIP-graph launch; launch choice
    from the autotuner.  Returns throughput, per-problem latency and the batch
    kernel's live per-launch time."""
    import torch
    from geobrain_amd import _native, autotune, synthetic
    from geobrain_amd import batch as B
    dev = torch.device("cuda:0")
    specs = synthetic.uneven_batch(96, seed=7)
    problems = [B.problem_inputs(sp, dev) for sp in specs]
    if launch is None:
        launch = autotune.tune_batch(problems, verbose=True) if tune else B.BatchLaunch()
    bf = B.BatchedForward(problems, launch)
    # live per-launch time of the batch kernel (an eager run; the graph's
    # launches are the same)
    _native.set_kernel_timing(True)
    bf.run()
    torch.cuda.synchronize()
    kt = _native.kernel_times().get("fwd2_batch_kernel", (0, 0.0))
    _native.set_kernel_timing(False)
    bf.capture()
    for _ in range(warmup):
        bf.replay()
    torch.cuda.synchronize()
    times, lats = [], []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        bf.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
        lats.append(bf.latencies_ms())
    t = float(np.median(times))
    lat = np.median(np.stack(lats), axis=0)
    nts = np.array([sp["nt"] for sp in specs])
    return {"metric": "Gcell-updates/s, forward, uneven batch in one HIP graph",
            "value": round(bf.cells / t / 1e9, 3), "unit": "Gcell-updates/s",
            "graph_replay_s": round(t, 4), "replays": steps,
            "workload": "96 problems from seed 7: nz 200-600, nx 200-1200, nt 1000-4000, "
                        "1-8 shots, receivers on every column; forward (seismograms)",
            "cell_updates": bf.cells, "shots": int(sum(sp["n_shots"] for sp in specs)),
            "launch": {"shots_per_block": launch.shots_per_block,
                       "steps_per_launch": launch.steps_per_launch,
                       "tile_rows": launch.tile_rows},
            "latency_ms": {"min": round(float(lat.min()), 2),
                           "median": round(float(np.median(lat)), 2),
                           "p90": round(float(np.percentile(lat, 90)), 2),
                           "max": round(float(lat.max()), 2),
                           "per_step_us_median": round(float(np.median(lat / nts * 1e3)), 2)},
            "batch_kernel": {"launches": kt[0],
                             "avg_launch_us": round(kt[1] / max(kt[0], 1) * 1e3, 2)}}


def uneven_gradient(steps: int, warmup: int, tune: bool, launch=None) -> dict:
    """The uneven-batch workload as a gradient: the same 96 problems from seed
    7, misfit sum((p - d_obs)^2) against data modelled in 2%-slower velocity
    models, forward with checkpoints + adjoint sources + recompute + adjoint
    of all problems captured in one HIP graph (BatchedGradient), launch choice
    from the autotuner.  Returns forward+adjoint throughput, per-problem
    latency and the batch kernels' live per-launch times."""
    import torch
    from geobrain_amd import _native, autotune, synthetic
    from geobrain_amd import batch as B
    dev = torch.device("cuda:0")
    specs = synthetic.uneven_batch(96, seed=7)
    problems = [B.problem_inputs(sp, dev) for sp in specs]
    obs = B.BatchedForward([dict(a, vp=a["vp"] * 0.98) for a in problems])
    d_obs = [o["p"].clone() for o in obs.run()]
    torch.cuda.synchronize()
    del obs
    torch.cuda.empty_cache()
    if launch is None:
        launch = autotune.tune_batch_gradient(problems, verbose=True) if tune else B.BatchLaunch()
    bg = B.BatchedGradient(problems, launch)
    # live per-launch times of the batch kernels (an eager step; the graph's
    # launches are the same)
    _native.set_kernel_timing(True)
    bg.misfit_step(d_obs)
    torch.cuda.synchronize()
    kt = _native.kernel_times()
    _native.set_kernel_timing(False)
    bg.capture(lambda: bg.misfit_step(d_obs))
    for _ in range(warmup):
        bg.replay()
    torch.cuda.synchronize()
    times, lats = [], []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        bg.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
        lats.append(bg.latencies_ms())
    t = float(np.median(times))
    lat = np.median(np.stack(lats), axis=0)
    nts = np.array([sp["nt"] for sp in specs])
    return {"metric": "Gcell-updates/s, forward + adjoint gradient, uneven batch in one HIP graph",
            "value": round(bg.cells / t / 1e9, 3), "unit": "Gcell-updates/s",
            "graph_replay_s": round(t, 4), "replays": steps,
            "workload": "96 problems from seed 7: nz 200-600, nx 200-1200, nt 1000-4000, "
                        "1-8 shots, receivers on every column; gradient of the seismogram "
                        "misfit against data from 2%-slower models",
            "cell_updates": bg.cells, "shots": int(sum(sp["n_shots"] for sp in specs)),
            "launch": {"shots_per_block": launch.shots_per_block, "tile_rows": launch.tile_rows},
            "workspace_gb": round(bg.ws.numel() / 1e9, 1),
            "latency_ms": {"min": round(float(lat.min()), 2),
                           "median": round(float(np.median(lat)), 2),
                           "p90": round(float(np.percentile(lat, 90)), 2),
                           "max": round(float(lat.max()), 2),
                           "per_step_us_median": round(float(np.median(lat / nts * 1e3)), 2)},
            "kernels": {k: {"launches": n, "avg_launch_us": round(ms / n * 1e3, 2)}
                        for k, (n, ms) in kt.items()}}


def workload_3d(n_shots=8, n=200, nt=1000, nabc=20, dx=10.0, dt=8e-4, f0=10.0):
    """3D forward: n^3 grid (+nabc per side), depth-linear 1500-4500 m/s plus a
    3D Gaussian random field (lh 200 m, lv 50 m, std 200, seed 0) clipped to
    [1500, 4700], Gardner density, shots along the y = n/2 line at z = 2,
    receivers every 4th cell of the z = 2 plane, 10 Hz Ricker."""
    from geobrain_amd import synthetic
    from geobrain_amd.physics import wave
    vp = synthetic.grf_model_3d(n, n, n, dx, seed=0)
    rr = np.arange(0, n, 4, dtype=np.int64)
    ry, rx = np.meshgrid(rr, rr, indexing="ij")
    return dict(
        nz=n, ny=n, nx=n, nt=nt, nabc=nabc, dx=dx, dt=dt, vp=vp, rho=synthetic.gardner(vp),
        damp=wave.bc_pml_3d(n, n, n, dx, dx, dx, nabc, float(vp.max()), False).astype(np.float32),
        wavelet=np.tile(synthetic.ricker_padded(f0, dt, nt), (n_shots, 1)),
        src_x=synthetic.surface_shots(n, n_shots), src_y=np.full(n_shots, n // 2, np.int64),
        src_z=np.full(n_shots, 2, np.int64), rcv_x=rx.reshape(-1), rcv_y=ry.reshape(-1),
        rcv_z=np.full(rx.size, 2, np.int64))


def forward_3d(steps: int = 3, warmup: int = 1, cpu: bool = True, cpu_steps: int = 1300,
               nt: int = 1000, grad: bool = True, launch=None) -> dict:
    """Times acoustic_forward_kernel_3d (all nt steps = one bench step) on
    ``workload_3d``; roofline of the time step at its algorithmic 32 B per
    cell-update (p, vx, vy, vz each read and written once).  With ``grad``,
    also one vp gradient of the same problem.  ``launch`` pins a
    ``kernels.LaunchChoice3D`` (default: the library's choice)."""
    import torch
    from geobrain_amd import _native
    from geobrain_amd.physics import wave
    from geobrain_amd.physics.wave import kernels as K
    K.set_launch_choice_3d(launch)
    dev = torch.device("cuda:0")
    w = workload_3d(nt=nt)
    S, nt = w["wavelet"].shape
    T = lambda a, **kw: torch.as_tensor(a, device=dev, **kw)  # noqa: E731
    args = dict(nx=w["nx"], ny=w["ny"], nz=w["nz"], dx=w["dx"], dy=w["dx"], dz=w["dx"], nt=nt,
                dt=w["dt"], nabc=w["nabc"], free_surface=False, src_x=w["src_x"],
                src_y=w["src_y"], src_z=w["src_z"], src_n=S, src_v=T(w["wavelet"]),
                rcv_x=w["rcv_x"], rcv_y=w["rcv_y"], rcv_z=w["rcv_z"], rcv_n=w["rcv_x"].size,
                damp=T(w["damp"]), vp=T(w["vp"]), rho=T(w["rho"]))
    for _ in range(warmup):
        wave.acoustic_forward_kernel_3d(**args)
    torch.cuda.synchronize()
    _native.set_kernel_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = wave.acoustic_forward_kernel_3d(**args)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kt = _native.kernel_times()
    _native.set_kernel_timing(False)
    pad = [w["nz"] + 2 * w["nabc"], w["ny"] + 2 * w["nabc"], w["nx"] + 2 * w["nabc"]]
    cells = S * pad[0] * pad[1] * pad[2]
    step_kernels = [k for k in kt if k in ("p3_step_kernel", "v3_step_kernel",
                                           "fwd3_march_kernel<forward>", "fwd3f_kernel")]
    kernels = {k: {"launches": n, "total_s": round(ms / 1e3, 4),
                   "avg_launch_us": round(ms / n * 1e3, 2)} for k, (n, ms) in kt.items()}
    # one time step = one launch of each step kernel
    step_us = sum(kernels[k]["avg_launch_us"] for k in step_kernels)
    b_step = 32.0 * cells
    achieved = b_step / (step_us * 1e-6) / 1e9
    # measured HBM bytes per time step (rocprofv3 FETCH_SIZE / WRITE_SIZE) from
    # the committed profile of this workload, when it has one
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_3d_{S}shot_{pad[0]}cube.json")
    if os.path.exists(tpath) and pad[0] == pad[1] == pad[2]:
        with open(tpath) as fh:
            tdoc = json.load(fh)
        if tdoc.get("padded") == pad and all(k in tdoc["kernels"] for k in step_kernels):
            traffic = sum(tdoc["kernels"][k]["traffic"] for k in step_kernels)
    rec = {
        "metric": "Gcell-updates/s, 3D forward (seismograms + energy maps)",
        "value": round(cells * nt / wall / 1e9, 3), "unit": "Gcell-updates/s",
        "ms_per_step": round(wall * 1e3, 2), "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"3D forward: {S} shots, {w['nz']}^3 grid (+{w['nabc']} PML), "
                               f"{nt} steps, {w['rcv_x'].size} receivers",
                   "shots": S, "grid": [w["nz"], w["ny"], w["nx"]], "padded": pad, "nt": nt},
        "roofline": {"bound": "hbm", "kernel": "+".join(step_kernels), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "avg_time_step_us": round(step_us, 2),
                     "bytes_per_time_step": b_step},
        "kernels": kernels,
        "launch": dataclasses.asdict(K._launch3),
        "check": {"p_max": float(out["p"].abs().max())},
        "cpu_oracle": None,
    }
    if grad:
        # gradient of a seismogram misfit w.r.t. vp: forward with checkpoints,
        # recompute, adjoint, chain rule (one timed call after one warm-up)
        d_obs = out["p"].detach().clone().mul_(0.9)
        del out
        vp = args["vp"].clone().requires_grad_(True)

        def grad_step():
            o = wave.acoustic_forward_kernel_3d(**dict(args, vp=vp))
            ((o["p"] - d_obs) ** 2).sum().backward()

        grad_step()
        torch.cuda.synchronize()
        _native.set_kernel_timing(True)
        t0 = time.perf_counter()
        grad_step()
        torch.cuda.synchronize()
        gwall = time.perf_counter() - t0
        gkt = _native.kernel_times()
        _native.set_kernel_timing(False)
        rec["gradient"] = {
            "value": round(cells * nt / gwall / 1e9, 3), "unit": "Gcell-updates/s",
            "seconds": round(gwall, 3), "grad_vp_max": float(vp.grad.abs().max()),
            "kernels": {k: {"launches": n, "avg_launch_us": round(ms / n * 1e3, 2)}
                        for k, (n, ms) in gkt.items()}}
        del vp, d_obs
        torch.cuda.empty_cache()
    if cpu:
        import oracle
        oracle.build()
        sl = slice(0, 1)
        t0 = time.perf_counter()
        oracle.forward3d(nz=w["nz"], ny=w["ny"], nx=w["nx"], nabc=w["nabc"], free_surface=False,
                         dt=w["dt"], dz=w["dx"], src_z=w["src_z"][sl], src_y=w["src_y"][sl],
                         src_x=w["src_x"][sl], src_v=w["wavelet"][sl, :cpu_steps],
                         rcv_z=w["rcv_z"], rcv_y=w["rcv_y"], rcv_x=w["rcv_x"], vp=w["vp"],
                         rho=w["rho"], damp=w["damp"], precision="f32")
        t = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rec["cpu_oracle"] = {
            "value": round(pad[0] * pad[1] * pad[2] * cpu_steps / t / 1e9, 5),
            "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP32 C++ 3D oracle (OpenMP), forward, 1 shot x {cpu_steps} steps of "
                      f"the same grid, {t:.1f} s wall"}
    return rec


def fwi_precision(n_iter: int = 20, shots: int = 128, nz: int = 400, nx: int = 2000,
                  nt: int = 5000) -> dict:
    """The precision workload: FWI of a 400 x 2000 layered Gaussian-random model
    (dx 5 m, dt 0.5 ms, 5000 steps, 128 surface shots, 15 Hz Ricker, receivers on
    every column) from its Gaussian-smoothed (sigma 20 cells) version, 20 Adam
    iterations at lr 10 m/s with the example-09 driver (trace-normalised misfit,
    bound clip and Gardner refresh), run twice: checkpoints stored in FP32 and
    in BF16 (power-of-two FP32 scale per field and shot; compute FP32).  Reports
    the BF16 run's first gradient and final model against the FP32-storage run."""
    import torch
    from scipy.ndimage import gaussian_filter
    from geobrain_amd import synthetic
    from geobrain_amd.fwi import FWI, trace_l2_misfit
    from geobrain_amd.physics import wave
    dx, dt, f0 = 5.0, 5e-4, 15.0
    vp_true = synthetic.grf_model(nz, nx, dx, seed=5)
    vp_init = gaussian_filter(vp_true.astype(np.float64), sigma=20, mode="nearest").astype(np.float32)
    grid = wave.GridConfig(nx=nx, nz=nz, dx=dx, dz=dx)
    boundary = wave.BoundaryConfig(type="pml", n_layers=20, free_surface=False)
    w = synthetic.ricker_padded(f0, dt, nt)
    source = wave.Source(nt=nt, dt=dt, f0=f0)
    for x in synthetic.surface_shots(nx, shots):
        source.add_source(x=int(x), z=2, wavelet=w)
    receiver = wave.Receiver(nt=nt, dt=dt)
    receiver.add_receivers(x=np.arange(nx), z=np.full(nx, 2), rcv_type="pr")
    survey = wave.Survey(source=source, receiver=receiver)
    bounds = (float(vp_true.min()), float(vp_true.max()))
    model_true = wave.AcousticModel(grid=grid, boundary=boundary, vp=vp_true,
                                    rho=synthetic.gardner(vp_true), vp_grad=False, device="cuda")
    prop = wave.AcousticPropagator(model=model_true, survey=survey, device="cuda")
    t0 = time.perf_counter()
    with torch.no_grad():
        d_obs = prop.forward()["p"].detach().clone()
    torch.cuda.synchronize()
    t_obs = time.perf_counter() - t0
    del prop, model_true
    runs = {}
    for name, cdtype in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
        model = wave.AcousticModel(grid=grid, boundary=boundary, vp=vp_init.copy(),
                                   rho=synthetic.gardner(vp_init), vp_grad=True, rho_grad=False,
                                   vp_bound=bounds, auto_update_rho=True, device="cuda")
        prop = wave.AcousticPropagator(model=model, survey=survey, device="cuda")
        prop.checkpoint_dtype = cdtype
        fwi = FWI(prop, model, d_obs, lr=10.0, checkpoint_segments=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        first_grad = None
        for it in range(n_iter):  # FWI.step, keeping the first gradient
            fwi.optimizer.zero_grad()
            result = prop.forward(checkpoint_segments=1)
            loss = trace_l2_misfit(result["p"], d_obs)
            loss.backward()
            if first_grad is None:
                first_grad = model.vp.grad.detach().clone()
            fwi.optimizer.step()
            fwi.scheduler.step()
            model.forward()
            fwi.history.append(loss.item())
            print(f"  fwi {name} iteration {it}: loss {fwi.history[-1]:.6e} "
                  f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        runs[name] = dict(grad=first_grad, vp=model.vp.detach().clone(), losses=list(fwi.history),
                          seconds=time.perf_counter() - t0)
        del fwi, prop, model
        torch.cuda.empty_cache()

    def rel(a, b):
        return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()))

    vt = torch.as_tensor(vp_true, device="cuda")
    cells = shots * (nz + 40) * (nx + 40) * nt
    return {"metric": "BF16-checkpoint FWI: gradient and final-model rel-L2 vs FP32 storage",
            "workload": f"FWI {n_iter} Adam iterations (lr 10): {shots} shots, {nz}x{nx} grid, "
                        f"{nt} steps, dx 5 m; start = true model smoothed with sigma 20 cells",
            "grad_rel_l2_bf16_vs_fp32": rel(runs["bf16"]["grad"], runs["fp32"]["grad"]),
            "model_rel_l2_bf16_vs_fp32": rel(runs["bf16"]["vp"], runs["fp32"]["vp"]),
            "model_rel_l2_vs_true": {k: rel(runs[k]["vp"], vt) for k in runs},
            "initial_model_rel_l2_vs_true": rel(torch.as_tensor(vp_init, device="cuda"), vt),
            "loss_first_last": {k: [runs[k]["losses"][0], runs[k]["losses"][-1]] for k in runs},
            "seconds_per_iteration": {k: round(runs[k]["seconds"] / n_iter, 3) for k in runs},
            "gradient_gcells": {k: round(cells / (runs[k]["seconds"] / n_iter) / 1e9, 2)
                                for k in runs},
            "observed_data_s": round(t_obs, 2),
            "fp64_oracle": "reduced-size comparison in tests/test_gpu_precision.py"}


def launch3d(spec):
    """``shots_per_block,step_kernel,z_chunk`` -> kernels.LaunchChoice3D (None: default)."""
    if not spec:
        return None
    from geobrain_amd.physics.wave import kernels as K
    return K.LaunchChoice3D(*[int(v) for v in spec.split(",")])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baselines (reference path and C++ oracle)")
    ap.add_argument("--cpu-shots", type=int, default=8,
                    help="shots in the reference-path CPU baseline's sample")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="time steps in the reference-path CPU baseline's sample")
    ap.add_argument("--oracle-steps", type=int, default=2500,
                    help="time steps of the C++ oracle timing (one shot)")
    ap.add_argument("--no-tune", action="store_true", help="library default launch choice")
    ap.add_argument("--no-eager", action="store_true", help="skip the per-shot PyTorch GPU baseline")
    ap.add_argument("--launch", type=str, default=None,
                    help="pin a launch choice instead of autotuning: comma-separated "
                         "tile_rows,shots_per_block,march,bwd_tile_rows,bwd_shots_per_block,"
                         "steps_per_launch[,fwd_march]")
    ap.add_argument("--workload", choices=["gradient", "fwi-bf16", "fwd3d"], default="gradient",
                    help="gradient: the 64-shot batched gradient (default, plus the uneven "
                         "batch and the 3D forward); fwi-bf16: the BF16-checkpoint FWI "
                         "precision study; fwd3d: the 3D forward alone")
    ap.add_argument("--fwi-iters", type=int, default=20)
    ap.add_argument("--batch-launch", type=str, default=None,
                    help="pin the uneven batch's launch choice instead of autotuning: "
                         "shots_per_block,steps_per_launch,tile_rows")
    ap.add_argument("--grad-batch-launch", type=str, default=None,
                    help="pin the uneven batch gradient's launch choice instead of autotuning: "
                         "shots_per_block,tile_rows")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-batch forward workload (96 problems, one HIP graph)")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D forward workload")
    ap.add_argument("--nt3d", type=int, default=1000, help="time steps of the 3D workload")
    ap.add_argument("--launch3d", type=str, default=None,
                    help="3D launch choice shots_per_block,step_kernel,z_chunk (default: library's)")
    ap.add_argument("--no-grad3d", action="store_true", help="skip the 3D gradient timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy three times (the reference "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 4 << 28}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots shard across GPUs with one gradient all-reduce per iteration, "
              "not implemented here)", file=sys.stderr)
        sys.exit(2)

    import torch
    from geobrain_amd.physics import wave

    if args.workload == "fwi-bf16":
        print(json.dumps(fwi_precision(n_iter=args.fwi_iters)), flush=True)
        return
    if args.workload == "fwd3d":
        rec = forward_3d(steps=args.steps, warmup=args.warmup, cpu=not args.no_cpu, nt=args.nt3d,
                         grad=not args.no_grad3d, launch=launch3d(args.launch3d))
        rec.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                    "higher_is_better": True, "scaling": "weak", "vs_baseline": None})
        print(json.dumps(rec), flush=True)
        return
    dev = torch.device("cuda:0")
    w = workload(n_shots=args.shots, nt=args.nt)
    S, nt = args.shots, args.nt
    NZ, NX = w["nz"] + 2 * w["nabc"], w["nx"] + 2 * w["nabc"]
    k = 50
    segments = -(-nt // k)
    T = lambda a, **kw: torch.as_tensor(a, device=dev, **kw)  # noqa: E731
    common = dict(nx=w["nx"], nz=w["nz"], dx=w["dx"], dz=w["dx"], nt=nt, dt=w["dt"],
                  nabc=w["nabc"], free_surface=False, src_x=w["src_x"], src_z=w["src_z"],
                  src_n=S, src_v=T(w["wavelet"]), rcv_x=w["rcv_x"], rcv_z=w["rcv_z"],
                  rcv_n=w["nx"], damp=T(w["damp"]), rho=T(w["rho"]))
    with torch.no_grad():
        d_obs = wave.acoustic_forward_kernel(vp=T(w["vp_obs"]), **common)["p"]
    vp = T(w["vp"]).requires_grad_(True)
    stream = torch.cuda.current_stream(dev)
    choice = None
    if args.launch:
        from geobrain_amd.physics.wave import kernels as K
        choice = K.LaunchChoice(*[int(v) for v in args.launch.split(",")])
        K.set_launch_choice(choice)
    elif not args.no_tune:
        from geobrain_amd import autotune
        choice = autotune.tune(dict(common, vp=vp, nt=nt), nt_trial=60, verbose=True)

    def step(ev=None):
        vp.grad = None
        if ev:
            ev[0].record(stream)
        out = wave.acoustic_forward_kernel(vp=vp, checkpoint_segments=segments, **common)
        if ev:
            ev[1].record(stream)
        loss = ((out["p"] - d_obs) ** 2).sum()
        if ev:
            ev[2].record(stream)
        loss.backward()
        if ev:
            ev[3].record(stream)
        return loss

    progress("headline workload ready" + ("" if choice is None else f", launch {choice}"))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    progress(f"{args.warmup} warm-up gradient(s) done")
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    # timed steps: no instrumentation (the events above are four per step, on
    # the launch stream, outside the library)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    progress(f"{args.steps} timed gradient(s): {wall / args.steps:.3f} s each")
    t_fwd = np.mean([e[0].elapsed_time(e[1]) for e in evs]) / 1e3
    t_bwd = np.mean([e[2].elapsed_time(e[3]) for e in evs]) / 1e3
    # one more step, instrumented: the library brackets each time-step kernel
    # launch with HIP events on the stream it launches on (torch's current
    # stream); per-kernel times and the roofline come from it
    from geobrain_amd import _native
    _native.set_kernel_timing(True)
    ti = time.perf_counter()
    step()
    torch.cuda.synchronize()
    wall_instr = time.perf_counter() - ti
    ktimes = _native.kernel_times()
    _native.set_kernel_timing(False)
    cells = S * NZ * NX * nt
    step_s = wall / args.steps
    value = cells / step_s / 1e9
    # algorithmic bytes per cell and time step of each kernel: state read +
    # written (p, vx, vz or their adjoints, 24 B), plus the bracket D written by
    # the recompute / read by the adjoint (4 B); two-step kernels do two steps
    per_cell_step = {"fwd_step_kernel<forward>": 24.0, "fwd_step_kernel<recompute>": 28.0,
                     "adj_step_kernel": 28.0, "recomp_march_kernel": 28.0}
    kernels = {}
    for name, (launches, ms) in ktimes.items():
        # adj2m_kernel / fwd2m_kernel: the z-marching two-step adjoint / forward;
        # fwd2_mixed_kernel: the two-phase two-step forward / recompute
        base = name.replace("fwd2_mixed_kernel", "fwd_step_kernel").replace(
            "fwd2_", "fwd_").replace("adj2_", "adj_").replace(
            "adj2m_kernel", "adj_step_kernel").replace("fwd2m_kernel", "fwd_step_kernel")
        steps_per = 2 if ("fwd2" in name or "adj2" in name) else 1
        b = per_cell_step[base] * steps_per * S * NZ * NX
        avg_s = ms / launches / 1e3
        kernels[name] = {"launches": launches, "total_s": round(ms / 1e3, 4),
                         "avg_launch_us": round(avg_s * 1e6, 2), "bytes_per_launch": b,
                         "achieved_gbs": round(b / avg_s / 1e9, 1)}
    dom = max(kernels, key=lambda n: kernels[n]["total_s"])
    achieved = kernels[dom]["achieved_gbs"]
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by
    # a copy kernel) from the committed profile of this workload, when it has one
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         f"hbm_traffic_{S}shot_{w['nz']}x{w['nx']}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tdoc = json.load(fh)
        if tdoc.get("padded") == [NZ, NX] and dom in tdoc.get("kernels", {}):
            traffic = tdoc["kernels"][dom]["traffic"]
    job_bytes = (80.0 + 24.0 / k) * cells
    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"batched gradient: {S} shots, {w['nz']}x{w['nx']} grid "
                               f"(+{w['nabc']} PML), {nt} steps, checkpoint every {k}",
                   "shots": S, "grid": [w["nz"], w["nx"]], "padded": [NZ, NX], "nt": nt,
                   "checkpoint_every": k,
                   "launch": None if choice is None else dataclasses.asdict(choice)},
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_us": kernels[dom]["avg_launch_us"],
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"]},
        "kernels": kernels,
        "instrumented_step_s": round(wall_instr, 4),
        "phases": {"forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4),
                   "forward_gcells": round(cells / t_fwd / 1e9, 3),
                   "backward_gcells": round(cells / t_bwd / 1e9, 3),
                   "job_hbm_frac": round(job_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
                   "job_bytes_per_cell_update": 80.0 + 24.0 / k},
        "cpu_baseline": None,
    }
    if not args.no_eager:
        base = per_shot_gpu_baseline(w, d_obs)
        rec["per_shot_gpu"] = base
        progress("per-shot GPU baseline done")
        rec["speedup_vs_per_shot_gpu"] = round(value / base["value"], 1)
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(w, shots=args.cpu_shots, steps=args.cpu_steps)
        rec["cpu_oracle"] = cpu_oracle_baseline(w, steps=args.oracle_steps)
        rec["speedup_vs_cpu_baseline"] = round(value / rec["cpu_baseline"]["value"], 1)
        progress("CPU baselines done")
    if not args.no_uneven:
        del d_obs, vp, common
        torch.cuda.empty_cache()
        blaunch = None
        if args.batch_launch:
            from geobrain_amd import batch as B
            blaunch = B.BatchLaunch(*[int(v) for v in args.batch_launch.split(",")])
        rec["uneven_batch"] = uneven_batch(steps=3, warmup=1, tune=not args.no_tune,
                                           launch=blaunch)
        progress("uneven batch forward done")
        torch.cuda.empty_cache()
        glaunch = None
        if args.grad_batch_launch:
            from geobrain_amd import batch as B
            spb, rows = (int(v) for v in args.grad_batch_launch.split(","))
            glaunch = B.BatchLaunch(shots_per_block=spb, tile_rows=rows)
        rec["uneven_gradient"] = uneven_gradient(steps=2, warmup=1, tune=not args.no_tune,
                                                 launch=glaunch)
        progress("uneven batch gradient done")
    if not args.no_3d:
        torch.cuda.empty_cache()
        rec["forward_3d"] = forward_3d(steps=2, warmup=1, cpu=not args.no_cpu, nt=args.nt3d,
                                       grad=not args.no_grad3d, launch=launch3d(args.launch3d))
        progress("3D workload done")
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
