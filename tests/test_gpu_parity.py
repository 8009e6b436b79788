"""GPU parity of the HIP propagator against the reference's golden vectors and
the CPU oracle.  Bar: relative L2 <= 1e-5 for FP32 seismograms, energy maps
and gradients (the tolerance is written in each assert); the error against the
FP64 oracle is printed for each result."""

import ctypes

import numpy as np
import pytest
import torch

import oracle
from geobrain_amd import _native
from geobrain_amd.physics import wave
from geobrain_amd.physics.wave import kernels as K
from tests._fixtures import (FWD_CASES, GRAD_CASES, OUT_KEYS, adjoint_sources, chunk_len, fd_c1,
                             load, oracle_args, rel_l2)

pytestmark = pytest.mark.gpu
TOL = 1e-5
DEV = torch.device("cuda:0")


def _t(a, **kw):
    return torch.as_tensor(np.asarray(a), device=DEV, **kw)


def run(z, m, *, vp=None, segments=None, snapshot_every=None, grad=(), launch=None, wavelet=None):
    """Runs acoustic_forward_kernel on a fixture's inputs; returns (outputs, tensors)."""
    K.set_launch_choice(launch)
    try:
        vp_t = _t(z["in_vp"] if vp is None else vp, dtype=torch.float32).clone()
        rho_t = _t(z["in_rho"], dtype=torch.float32).clone()
        w_t = _t(z["in_wavelet"] if wavelet is None else wavelet, dtype=torch.float32).clone()
        vp_t.requires_grad_("vp" in grad)
        rho_t.requires_grad_("rho" in grad)
        w_t.requires_grad_("w" in grad)
        out = wave.acoustic_forward_kernel(
            nx=m["nx"], nz=m["nz"], dx=m["dx"], dz=m["dx"], nt=m["nt"], dt=m["dt"],
            nabc=m["nabc"], free_surface=m["free_surface"],
            src_x=_t(z["in_src_x"]), src_z=_t(z["in_src_z"]), src_n=len(z["in_src_x"]),
            src_v=w_t, rcv_x=_t(z["in_rcv_x"]), rcv_z=_t(z["in_rcv_z"]),
            rcv_n=len(z["in_rcv_x"]), damp=_t(z["in_damp"], dtype=torch.float32),
            vp=vp_t, rho=rho_t,
            checkpoint_segments=m["segments"] if segments is None else segments,
            snapshot_every=m["snapshot_every"] if snapshot_every is None else snapshot_every)
    finally:
        K.set_launch_choice(None)
    return out, dict(vp=vp_t, rho=rho_t, w=w_t)


def np_(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", FWD_CASES + GRAD_CASES)
def test_forward_matches_reference(name):
    print(f"[case] forward {name}", flush=True)
    z, m = load(name)
    out, _ = run(z, m)
    torch.cuda.synchronize()
    o64 = oracle.forward(**oracle_args(z, m), chunk=chunk_len(m["nt"], m["segments"]),
                         snapshot_every=m["snapshot_every"], precision="f64")
    n = 0
    for k in OUT_KEYS:
        if "out_" + k in z:
            got = np_(out[k])
            e_ref = rel_l2(got, z["out_" + k])
            e64 = rel_l2(got, o64[k])
            same = np.array_equal(got, z["out_" + k])
            print(f"  {k}: rel-L2 vs reference {e_ref:.2e} (bitwise={same}), vs FP64 oracle {e64:.2e}")
            assert e_ref <= TOL, k
            n += 1
    if "out_p_sub" in z:
        assert rel_l2(np_(out["p"])[:, ::10, ::10], z["out_p_sub"]) <= TOL
    assert n >= 3 or "out_p_sub" in z


@pytest.mark.parametrize("name", GRAD_CASES)
def test_gradient_matches_reference_autograd(name):
    print(f"[case] gradient {name}", flush=True)
    z, m = load(name)
    want = ["vp"] + (["rho"] if "grad_rho" in z else []) + (["w"] if "grad_w" in z else [])
    obs = None
    if "g_seed" not in m:
        with torch.no_grad():
            obs, _ = run(z, m, vp=z["in_vp_obs"], segments=1)
            obs = obs["p"]
    out, ts = run(z, m, grad=want)
    if "g_seed" in m:
        g = adjoint_sources(z, m)
        loss = sum((_t(gk) * out[k]).sum() for gk, k in zip(g, ("p", "vx", "vz")))
    else:
        loss = ((out["p"] - obs) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    if m.get("loss") is not None:
        print(f"  loss {float(loss):.7e} vs reference {m['loss']:.7e}")
    args = oracle_args(z, m)
    if "g_seed" in m:
        g64 = adjoint_sources(z, m)
    else:
        f64 = oracle.forward(**args, precision="f64")["p"]
        o64 = oracle.forward(**dict(args, vp=z["in_vp_obs"]), precision="f64")["p"]
        g64 = (2.0 * (f64 - o64), None, None)
    gv64, gr64, gw64 = oracle.gradient(**args, g_p=g64[0], g_vx=g64[1], g_vz=g64[2],
                                       chunk=chunk_len(m["nt"], m["segments"]), precision="f64",
                                       wavelet_shared=z["in_wavelet"].ndim == 1)
    for key, gkey, g64k in (("vp", "grad_vp", gv64), ("rho", "grad_rho", gr64), ("w", "grad_w", gw64)):
        if key in want:
            got = np_(ts[key].grad)
            e_ref = rel_l2(got, z[gkey])
            e64 = rel_l2(got, g64k)
            print(f"  grad {key}: rel-L2 vs reference autograd {e_ref:.2e}, vs FP64 oracle {e64:.2e}"
                  f" (reference vs FP64: {rel_l2(z[gkey], g64k):.2e})")
            assert e_ref <= TOL, key


@pytest.mark.parametrize("tile_rows,spb", [(16, 1), (32, 0), (32, 2), (16, 3), (24, 0), (24, 2)])
def test_launch_choices_do_not_change_results(tile_rows, spb):
    print(f"[case] launch tile_rows={tile_rows} spb={spb}", flush=True)
    z, m = load("tiny2_grad")
    base, tb = run(z, m, grad=("vp", "w"))
    alt, ta = run(z, m, grad=("vp", "w"), launch=K.LaunchChoice(tile_rows, spb))
    for k in ("p", "vx", "vz"):
        np.testing.assert_array_equal(np_(base[k]), np_(alt[k]))
    for k in ("forward_wavefield_p", "forward_wavefield_vx", "forward_wavefield_vz"):
        assert rel_l2(np_(alt[k]), np_(base[k])) <= 1e-6
    g = adjoint_sources(z, m)
    for out, ts in ((base, tb), (alt, ta)):
        sum((_t(gk) * out[k]).sum() for gk, k in zip(g, ("p", "vx", "vz"))).backward()
    assert rel_l2(np_(ta["vp"].grad), np_(tb["vp"].grad)) <= 1e-6
    np.testing.assert_array_equal(np_(ta["w"].grad), np_(tb["w"].grad))


def test_batched_shots_equal_single_shots():
    print("[case] batch vs single shots", flush=True)
    z, m = load("tiny2_grad")
    full, _ = run(z, m)
    for s in range(len(z["in_src_x"])):
        zz = {k: z[k] for k in z.files}
        zz["in_src_x"], zz["in_src_z"] = z["in_src_x"][s:s + 1], z["in_src_z"][s:s + 1]
        zz["in_wavelet"] = z["in_wavelet"][s:s + 1]
        one, _ = run(zz, m)
        for k in ("p", "vx", "vz"):
            np.testing.assert_array_equal(np_(one[k])[0], np_(full[k])[s])


def test_batched_gradient_equals_sum_of_single_shot_gradients():
    """The config-3 batch at reduced size (8 shots along the surface of the
    350 x 1700 model, 1000 steps, long enough for the shots' wavefields to
    overlap): the batched vp gradient against the sum of the 8 single-shot
    gradients (FP32 sums in another order, so relative L2), and each shot's
    wavelet gradient bit for bit."""
    import bench
    print("[case] batched gradient vs single shots, 8 shots 350x1700 x 1000 steps", flush=True)
    w = bench.workload(n_shots=8, nt=1000)
    S, nt = w["wavelet"].shape
    R = w["rcv_x"].size
    g_p = np.random.default_rng(3).standard_normal((S, nt, R)).astype(np.float32)

    def grad(shots):
        vp = _t(w["vp"], dtype=torch.float32).requires_grad_(True)
        wv = _t(w["wavelet"][shots], dtype=torch.float32).requires_grad_(True)
        out = wave.acoustic_forward_kernel(
            nx=w["nx"], nz=w["nz"], dx=w["dx"], dz=w["dx"], nt=nt, dt=w["dt"], nabc=w["nabc"],
            free_surface=False, src_x=w["src_x"][shots], src_z=w["src_z"][shots],
            src_n=len(shots), src_v=wv, rcv_x=w["rcv_x"], rcv_z=w["rcv_z"], rcv_n=R,
            damp=_t(w["damp"], dtype=torch.float32), vp=vp,
            rho=_t(w["rho"], dtype=torch.float32), checkpoint_segments=8)
        (_t(g_p[shots]) * out["p"]).sum().backward()
        return np_(vp.grad).astype(np.float64), np_(wv.grad)

    g_all, w_all = grad(np.arange(S))
    g_sum = np.zeros_like(g_all)
    for s in range(S):
        g_one, w_one = grad(np.array([s]))
        g_sum += g_one
        np.testing.assert_array_equal(w_one[0], w_all[s], err_msg=f"wavelet gradient, shot {s}")
    e = rel_l2(g_all, g_sum)
    print(f"  vp gradient: rel-L2 batched vs sum of single shots {e:.2e}")
    assert np.abs(g_all).max() > 0 and e <= 1e-6


@pytest.mark.parametrize("fs", [False, True])
def test_propagator_end_to_end(fs):
    """AcousticPropagator as examples 07/09 drive it (free surface, Gardner
    density, integrated Ricker), against the FP32 oracle."""
    print(f"[case] propagator fs={fs}", flush=True)
    from scipy.integrate import cumulative_trapezoid
    nz, nx, nt, dx, dt = 40, 90, 600, 10.0, 1e-3
    vp = oracle_vp = np.asarray(
        np.linspace(1500, 3000, nz)[:, None] * np.ones((1, nx)), np.float32)
    grid = wave.GridConfig(nx=nx, nz=nz, dx=dx, dz=dx)
    bnd = wave.BoundaryConfig(type="pml", n_layers=12, free_surface=fs)
    model = wave.AcousticModel(grid, bnd, vp=vp, rho=np.ones_like(vp), device="cuda")
    w, _ = wave.RickerWavelet()(f0=10.0, dt=dt)
    w = np.pad(w.numpy(), (0, nt - len(w)))
    w = cumulative_trapezoid(w, initial=0)
    src = wave.Source(nt=nt, dt=dt, f0=10.0)
    for x in (5, 45, 80):
        src.add_source(x=x, z=1, wavelet=w)
    rcv = wave.Receiver(nt=nt, dt=dt)
    rcv.add_receivers(x=np.arange(nx), z=np.ones(nx, int), rcv_type="pr")
    prop = wave.AcousticPropagator(model, wave.Survey(src, rcv), device="cuda")
    res = prop.forward(checkpoint_segments=4)
    rho = synthetic_rho = (310.0 * np.power(oracle_vp, 0.25)).astype(np.float32)
    damp = oracle.bc_pml(nx, nz, dx, dx, 12, float(vp.max()), False).astype(np.float32)
    ref = oracle.forward(nz=nz, nx=nx, nabc=12, free_surface=fs, dt=dt, dz=dx,
                         src_z=[1, 1, 1], src_x=[5, 45, 80],
                         src_v=np.stack([w] * 3).astype(np.float32), rcv_z=np.ones(nx, int),
                         rcv_x=np.arange(nx), vp=vp, rho=synthetic_rho, damp=damp,
                         chunk=chunk_len(nt, 4), precision="f32")
    np.testing.assert_array_equal(model.rho.detach().cpu().numpy(), rho)
    for k in ("p", "vx", "vz", "forward_wavefield_p"):
        e = rel_l2(np_(res[k]), ref[k])
        print(f"  {k}: rel-L2 vs FP32 oracle {e:.2e}")
        assert e <= TOL
    snap = prop.forward(shot_index=1, snapshot_every=10)
    assert snap.snapshots.shape == (60, nz, nx) and snap.snapshots.device.type == "cpu"
    np.testing.assert_array_equal(np_(snap.p)[0], np_(res.p)[1])


def test_propagator_reads_moved_positions():
    """Moving the shots by reassigning ``prop.src_x`` (device tensor, as the
    reference's propagator holds it) changes the next forward: the result
    equals a propagator built with the new positions; moving them back
    restores the first result bit for bit."""
    print("[case] propagator with reassigned source positions", flush=True)
    nz, nx, nt, dx, dt = 30, 80, 300, 10.0, 1e-3
    vp = np.asarray(np.linspace(1500, 2500, nz)[:, None] * np.ones((1, nx)), np.float32)
    grid = wave.GridConfig(nx=nx, nz=nz, dx=dx, dz=dx)
    bnd = wave.BoundaryConfig(type="pml", n_layers=10, free_surface=False)
    w, _ = wave.RickerWavelet()(f0=15.0, dt=dt)
    w = np.pad(w.numpy(), (0, nt - len(w)))

    def make(xs):
        model = wave.AcousticModel(grid, bnd, vp=vp, rho=np.ones_like(vp), device="cuda")
        src = wave.Source(nt=nt, dt=dt, f0=15.0)
        for x in xs:
            src.add_source(x=x, z=2, wavelet=w)
        rcv = wave.Receiver(nt=nt, dt=dt)
        rcv.add_receivers(x=np.arange(nx), z=np.full(nx, 2), rcv_type="pr")
        return wave.AcousticPropagator(model, wave.Survey(src, rcv), device="cuda")

    prop = make([10, 40])
    first = np_(prop.forward().p)
    prop.src_x = torch.tensor([25, 70], device="cuda")
    moved = np_(prop.forward().p)
    np.testing.assert_array_equal(moved, np_(make([25, 70]).forward().p))
    assert not np.array_equal(moved, first)
    prop.src_x[:] = torch.tensor([10, 40], device="cuda")  # in place: a new version
    np.testing.assert_array_equal(np_(prop.forward().p), first)


def _c2r_device_args():
    z, m = load("c2r_grad")
    T = lambda a, **kw: torch.as_tensor(np.asarray(a), device=DEV, **kw)  # noqa: E731
    args = dict(nx=m["nx"], nz=m["nz"], dx=m["dx"], dz=m["dx"], nt=m["nt"], dt=m["dt"],
                nabc=m["nabc"], free_surface=m["free_surface"], src_x=T(z["in_src_x"]),
                src_z=T(z["in_src_z"]), src_n=len(z["in_src_x"]),
                src_v=T(z["in_wavelet"], dtype=torch.float32), rcv_x=T(z["in_rcv_x"]),
                rcv_z=T(z["in_rcv_z"]), rcv_n=len(z["in_rcv_x"]),
                damp=T(z["in_damp"], dtype=torch.float32), rho=T(z["in_rho"], dtype=torch.float32),
                checkpoint_segments=m["segments"])
    return z, m, args


def test_device_positions_copied_once():
    """Positions given as device tensors, as the reference's propagator holds
    them: the first call copies them to the host once and prepares the
    geometry on the device; a repeated call copies nothing and gives the same
    seismogram and gradient bit for bit."""
    print("[case] device positions: one host copy, then none", flush=True)
    z, m, args = _c2r_device_args()
    with torch.no_grad():
        d_obs = wave.acoustic_forward_kernel(vp=_t(z["in_vp_obs"], dtype=torch.float32), **args)["p"]
    vp = _t(z["in_vp"], dtype=torch.float32).requires_grad_(True)

    def step():
        p = wave.acoustic_forward_kernel(vp=vp, **args)["p"]
        (g,) = torch.autograd.grad(p, vp, grad_outputs=2.0 * (p - d_obs))
        return p, g

    p0, g0 = step()
    torch.cuda.synchronize()
    copies = K.host_index_copies
    p1, g1 = step()
    torch.cuda.synchronize()
    assert K.host_index_copies == copies, "a repeated call copied positions to the host"
    assert torch.equal(p1, p0) and torch.equal(g1, g0)
    e = rel_l2(np_(g0), z["grad_vp"])
    print(f"  gradient rel-L2 vs reference autograd {e:.2e}")
    assert e <= TOL


def test_forward_and_gradient_captured_through_the_c_abi():
    """With a prepared geometry (gbw_acoustic2d_geometry_setup), the forward with
    checkpoints, the adjoint sources 2(p - d_obs) (torch ops) and the backward
    only enqueue kernels: one HIP graph holds the whole gradient.  Three
    replays, with eager work and synchronisation between them, equal the eager
    calls bit for bit, and the gradient is the reference's (rel-L2 <= 1e-5)."""
    print("[case] C-ABI forward + gradient captured in one HIP graph", flush=True)
    lib = _native.load()
    z, m = load("c2r_grad")
    S, R, nt = len(z["in_src_x"]), len(z["in_rcv_x"]), m["nt"]
    prob = _native.Problem()
    prob.nx, prob.nz, prob.dx, prob.dz = m["nx"], m["nz"], m["dx"], m["dx"]
    prob.nt, prob.dt, prob.nabc, prob.free_surface = nt, m["dt"], m["nabc"], int(m["free_surface"])
    prob.src_n, prob.rcv_n, prob.energy_chunk = S, R, nt
    prob.checkpoint_every = chunk_len(nt, m["segments"])
    host = {k: np.ascontiguousarray(z["in_" + k], np.int64) for k in ("src_x", "src_z", "rcv_x", "rcv_z")}
    # (row-major copies: the fixture's damping array is stored column-major)
    dev = {k: _t(z["in_" + k], dtype=torch.float32).contiguous()
           for k in ("vp", "rho", "damp", "wavelet")}
    stream = torch.cuda.current_stream(DEV).cuda_stream
    geo = torch.empty(lib.gbw_acoustic2d_geometry_bytes(ctypes.byref(prob)), dtype=torch.uint8,
                      device=DEV)
    ins = _native.Inputs()
    ins.src_x, ins.src_z = host["src_x"].ctypes.data, host["src_z"].ctypes.data
    ins.rcv_x, ins.rcv_z = host["rcv_x"].ctypes.data, host["rcv_z"].ctypes.data
    _native.check(lib.gbw_acoustic2d_geometry_setup(ctypes.byref(prob), ctypes.byref(ins),
                                                    geo.data_ptr(), geo.numel(), stream))
    ins = _native.Inputs()  # positions are not read with a prepared geometry
    ins.geometry = geo.data_ptr()
    ins.src_v, ins.damp = dev["wavelet"].data_ptr(), dev["damp"].data_ptr()
    ins.vp, ins.rho = dev["vp"].data_ptr(), dev["rho"].data_ptr()
    f32 = dict(dtype=torch.float32, device=DEV)
    rec = {k: torch.empty((S, nt, R), **f32) for k in ("p", "vx", "vz")}
    ck = torch.empty(lib.gbw_acoustic2d_checkpoint_bytes(ctypes.byref(prob)), dtype=torch.uint8,
                     device=DEV)
    ws0 = torch.empty(lib.gbw_acoustic2d_workspace_bytes(ctypes.byref(prob), 0), dtype=torch.uint8,
                      device=DEV)
    ws1 = torch.empty(lib.gbw_acoustic2d_workspace_bytes(ctypes.byref(prob), 1), dtype=torch.uint8,
                      device=DEV)
    g_p = torch.empty((S, nt, R), **f32)
    grad = torch.empty_like(dev["vp"])
    out = _native.Outputs()
    out.rec_p, out.rec_vx, out.rec_vz = (rec[k].data_ptr() for k in ("p", "vx", "vz"))
    out.checkpoints = ck.data_ptr()
    adj = _native.Adjoint()
    adj.g_p, adj.checkpoints, adj.grad_vp = g_p.data_ptr(), ck.data_ptr(), grad.data_ptr()
    with torch.no_grad():
        obs = run(z, m, vp=z["in_vp_obs"], segments=1)[0]["p"]

    def step():
        _native.check(lib.gbw_acoustic2d_forward(ctypes.byref(prob), ctypes.byref(ins),
                                                 ctypes.byref(out), ws0.data_ptr(), ws0.numel(),
                                                 torch.cuda.current_stream(DEV).cuda_stream))
        torch.sub(rec["p"], obs, out=g_p)
        g_p.mul_(2.0)
        _native.check(lib.gbw_acoustic2d_backward(ctypes.byref(prob), ctypes.byref(ins),
                                                  ctypes.byref(adj), ws1.data_ptr(), ws1.numel(),
                                                  torch.cuda.current_stream(DEV).cuda_stream))

    step()
    torch.cuda.synchronize()
    p0, g0 = rec["p"].clone(), grad.clone()
    ref_out, _ = run(z, m)
    assert torch.equal(p0, ref_out["p"]), "forward differs from acoustic_forward_kernel"
    e = rel_l2(np_(g0), z["grad_vp"])
    print(f"  eager gradient rel-L2 vs reference autograd {e:.2e}")
    assert e <= TOL
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for rep in range(3):
        rec["p"].fill_(float("nan"))
        grad.fill_(float("nan"))
        junk = torch.randn(1 << 20, device=DEV).sum()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(rec["p"], p0) and torch.equal(grad, g0), f"replay {rep}"
        print(f"  replay {rep}: seismogram and gradient equal the eager call ({float(junk):.1f})")


@pytest.mark.skip(reason="capturing the autograd backward of acoustic_forward_kernel inside "
                         "torch.cuda.graph ended in a host segmentation fault in capture_end on "
                         "the MI355X box; the cause was not identified, so the case stays off the "
                         "GPU (the same kernels captured through the C ABI are tested above)")
def test_autograd_forward_and_gradient_in_one_graph():
    z, m, args = _c2r_device_args()
    vp = _t(z["in_vp"], dtype=torch.float32).requires_grad_(True)

    def step():
        p = wave.acoustic_forward_kernel(vp=vp, **args)["p"]
        (g,) = torch.autograd.grad(p, vp, grad_outputs=2.0 * p)
        return p, g

    p0, g0 = step()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        pg, gg = step()
    graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(pg, p0) and torch.equal(gg, g0)


def test_prepared_geometry_checks_its_shape():
    """A prepared geometry buffer used for a problem of another shape is refused
    before any kernel reads it."""
    lib = _native.load()
    z, m = load("tiny0_fwd")
    prob = _native.Problem()
    prob.nx, prob.nz, prob.dx, prob.dz = m["nx"], m["nz"], m["dx"], m["dx"]
    prob.nt, prob.dt, prob.nabc = m["nt"], m["dt"], m["nabc"]
    prob.src_n, prob.rcv_n = len(z["in_src_x"]), len(z["in_rcv_x"])
    host = {k: np.ascontiguousarray(z["in_" + k], np.int64) for k in ("src_x", "src_z", "rcv_x", "rcv_z")}
    ins = _native.Inputs()
    ins.src_x, ins.src_z = host["src_x"].ctypes.data, host["src_z"].ctypes.data
    ins.rcv_x, ins.rcv_z = host["rcv_x"].ctypes.data, host["rcv_z"].ctypes.data
    n = lib.gbw_acoustic2d_geometry_bytes(ctypes.byref(prob))
    buf = torch.empty(n, dtype=torch.uint8, device=DEV)
    stream = torch.cuda.current_stream(DEV).cuda_stream
    _native.check(lib.gbw_acoustic2d_geometry_setup(ctypes.byref(prob), ctypes.byref(ins),
                                                    buf.data_ptr(), n, stream))
    other = _native.Problem.from_buffer_copy(prob)
    other.nx += 4  # another grid: the buffer's receiver map does not fit it
    ins2 = _native.Inputs()
    ins2.geometry = buf.data_ptr()
    out = _native.Outputs()
    ws = torch.empty(lib.gbw_acoustic2d_workspace_bytes(ctypes.byref(other), 0), dtype=torch.uint8,
                     device=DEV)
    rc = lib.gbw_acoustic2d_forward(ctypes.byref(other), ctypes.byref(ins2), ctypes.byref(out),
                                    ws.data_ptr(), ws.numel(), stream)
    assert rc == -1 and "geometry buffer" in _native.last_error()
    bad = dict(host, rcv_z=np.full_like(host["rcv_z"], 10_000))
    ins3 = _native.Inputs()
    ins3.src_x, ins3.src_z = bad["src_x"].ctypes.data, bad["src_z"].ctypes.data
    ins3.rcv_x, ins3.rcv_z = bad["rcv_x"].ctypes.data, bad["rcv_z"].ctypes.data
    with pytest.raises(IndexError):
        _native.check(lib.gbw_acoustic2d_geometry_setup(ctypes.byref(prob), ctypes.byref(ins3),
                                                        buf.data_ptr(), n, stream))
    # the failed setup unregistered the buffer
    rc = lib.gbw_acoustic2d_forward(ctypes.byref(prob), ctypes.byref(ins2), ctypes.byref(out),
                                    ws.data_ptr(), ws.numel(), stream)
    assert rc == -1 and "geometry buffer" in _native.last_error()


@pytest.mark.parametrize("where", ["interior", "absorbing"])
def test_nan_velocity_propagates_as_the_oracle(where):
    """A NaN in vp reaches the seismograms and energy maps exactly where the
    reference's arithmetic puts it: the HIP outputs equal the FP32 oracle's
    bit for bit, NaN positions included (the reference itself does not check
    its inputs for NaN)."""
    print(f"[case] NaN in vp ({where})", flush=True)
    z, m = load("c1_fwd")
    vp = z["in_vp"].copy()
    if where == "interior":
        vp[50, 120] = np.nan
    else:
        vp[0, 0] = np.nan  # edge cell: pad_model copies it into the absorbing corner
    out, _ = run(z, m, vp=vp)
    ref = oracle.forward(**dict(oracle_args(z, m), vp=vp), chunk=chunk_len(m["nt"], m["segments"]),
                         precision="f32")
    for k in ("p", "vx", "vz", "forward_wavefield_p", "forward_wavefield_vx",
              "forward_wavefield_vz"):
        got = np_(out[k])
        nan = int(np.isnan(got).sum())
        np.testing.assert_array_equal(got, ref[k], err_msg=k)
        print(f"  {k}: {nan} NaN of {got.size}, bit-equal to the oracle")
    assert np.isnan(np_(out["p"])).any()


def test_zero_wavelet_gives_exact_zeros():
    """A zero wavelet leaves every output exactly zero, and the gradient of any
    seismogram loss is exactly zero too (every term carries a zero field)."""
    print("[case] zero wavelet", flush=True)
    z, m = load("tiny1_grad")
    out, ts = run(z, m, grad=("vp", "rho"), wavelet=np.zeros_like(z["in_wavelet"]))
    for k in ("p", "vx", "vz", "forward_wavefield_p", "forward_wavefield_vx",
              "forward_wavefield_vz"):
        assert not np_(out[k]).any(), k
    g = adjoint_sources(z, m)
    sum((_t(gk) * out[k]).sum() for gk, k in zip(g, ("p", "vx", "vz"))).backward()
    assert not np_(ts["vp"].grad).any() and not np_(ts["rho"].grad).any()


def test_dot_product_fp32():
    """<F w, g> vs <w, F^T g> through the HIP forward and adjoint."""
    print("[case] dot product", flush=True)
    z, m = load("c2r_grad")
    rng = np.random.default_rng(0)
    out, ts = run(z, m, grad=("w",))
    S, nt, R = out["p"].shape
    gs = [_t(rng.standard_normal((S, nt, R)).astype(np.float32)) for _ in range(3)]
    lhs = sum(float((g.double() * out[k].double()).sum()) for g, k in zip(gs, ("p", "vx", "vz")))
    sum((g * out[k]).sum() for g, k in zip(gs, ("p", "vx", "vz"))).backward()
    rhs = float((ts["w"].grad.double() * ts["w"].detach().double()).sum())
    print(f"  lhs {lhs:.9e} rhs {rhs:.9e} rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("name", ["vp", "rho"])
def test_gradient_matches_finite_differences_config1(name):
    """The HIP FP32 vp (rho) gradient of the config-1 problem (100 x 200, 1000
    steps) against FP64 central differences of the oracle: +- 1e-2 m/s (kg/m^3)
    on 16 seeded cells (the FP32 forward cannot resolve so small a step itself)."""
    print(f"[case] config-1 {name} gradient vs finite differences, 16 cells", flush=True)
    z, m, gs, ii, jj, fd = fd_c1(name)
    out, ts = run(z, m, grad=(name,))
    sum((_t(g.astype(np.float32)) * out[k]).sum() for g, k in zip(gs, ("p", "vx", "vz"))).backward()
    g = np_(ts[name].grad)
    err = np.abs(g[ii, jj] - fd) / np.abs(g).max()
    print(f"  max |HIP - FD| / max |grad| = {err.max():.2e}")
    assert err.max() <= 1e-5


def test_invalid_geometry_raises_before_launch():
    z, m = load("tiny0_fwd")
    bad = {k: z[k] for k in z.files}
    bad["in_src_x"] = np.array([0, m["nx"] + m["nabc"]])
    with pytest.raises(IndexError, match="index out of range"):
        run(bad, m)
    bad = {k: z[k] for k in z.files}
    bad["in_rcv_z"] = bad["in_rcv_z"].copy()
    bad["in_rcv_z"][0] = m["nz"] + m["nabc"] - 1  # vz has one row fewer than p
    with pytest.raises(IndexError, match="receiver 0"):
        run(bad, m)
    with pytest.raises(ValueError, match="src_v"):
        run(z, m, wavelet=z["in_wavelet"][:, :5])


def test_zz_full_size_config2_linearity_and_batch():
    """350 x 1700, 4000 steps: doubling the wavelet doubles every output, and a
    4-shot batch reproduces each shot run alone bit for bit.

    Scaling by two is exact for every normal-range operation, but at the
    wavefront's leading edge the particle velocities (about alpha2 = dt/(rho dz)
    ~ 5e-8 times the pressure differences) fall into the subnormal range and
    round with O(1) relative error; the FP32 oracle shows the same ~4e-7
    departure, so the bound is the FP32 parity bar."""
    print("[case] full-size config 2 linearity", flush=True)
    from geobrain_amd import synthetic
    nz, nx, nt = 350, 1700, 4000
    vp = synthetic.grf_model(nz, nx, 5.0, seed=0)
    w = synthetic.ricker_padded(15.0, 5e-4, nt)
    damp = wave.bc_pml(nx, nz, 5.0, 5.0, 20, float(vp.max()), False).astype(np.float32)
    xs = synthetic.surface_shots(nx, 4)
    z = dict(in_vp=vp, in_rho=synthetic.gardner(vp), in_damp=damp,
             in_wavelet=np.stack([w] * 4), in_src_x=xs, in_src_z=np.full(4, 2),
             in_rcv_x=np.arange(nx), in_rcv_z=np.full(nx, 2))
    m = dict(nz=nz, nx=nx, nt=nt, dx=5.0, dt=5e-4, nabc=20, free_surface=False, segments=80,
             snapshot_every=0)
    with torch.no_grad():
        a, _ = run(z, m)
        b, _ = run(z, m, wavelet=2 * z["in_wavelet"])
        for k in ("p", "vx", "vz"):
            e = rel_l2(np_(b[k]), 2 * np_(a[k]))
            same = float((2 * a[k] == b[k]).float().mean())
            print(f"  {k}: rel-L2(F(2w), 2F(w)) = {e:.1e}, bitwise-equal fraction {same:.6f}")
            assert e <= TOL, k
        z1 = dict(z, in_src_x=xs[2:3], in_src_z=np.full(1, 2), in_wavelet=z["in_wavelet"][2:3])
        c, _ = run(z1, m)
        for k in ("p", "vx", "vz"):
            assert torch.equal(c[k][0], a[k][2]), k
    assert float(a["p"].abs().max()) > 0


def test_bf16_checkpoints_match_oracle_emulation():
    """bf16 checkpoint storage: forward outputs unchanged; the gradient equals the
    oracle's gradient with the same bf16 checkpoint rounding (rel-L2 <= 1e-5)
    and its distance from the FP32-checkpoint gradient is reported."""
    print("[case] bf16 checkpoints", flush=True)
    z, m = load("c2r_grad")
    with torch.no_grad():
        obs, _ = run(z, m, vp=z["in_vp_obs"], segments=1)
        obs = obs["p"]
    grads = {}
    for dt_ in (torch.float32, torch.bfloat16):
        vp_t = _t(z["in_vp"], dtype=torch.float32).clone().requires_grad_(True)
        out = wave.acoustic_forward_kernel(
            nx=m["nx"], nz=m["nz"], dx=m["dx"], dz=m["dx"], nt=m["nt"], dt=m["dt"],
            nabc=m["nabc"], free_surface=m["free_surface"], src_x=_t(z["in_src_x"]),
            src_z=_t(z["in_src_z"]), src_n=1, src_v=_t(z["in_wavelet"], dtype=torch.float32),
            rcv_x=_t(z["in_rcv_x"]), rcv_z=_t(z["in_rcv_z"]), rcv_n=len(z["in_rcv_x"]),
            damp=_t(z["in_damp"], dtype=torch.float32), vp=vp_t,
            rho=_t(z["in_rho"], dtype=torch.float32), checkpoint_segments=m["segments"],
            checkpoint_dtype=dt_)
        ((out["p"] - obs) ** 2).sum().backward()
        grads[dt_] = (np_(vp_t.grad), np_(out["p"]))
    np.testing.assert_array_equal(grads[torch.float32][1], grads[torch.bfloat16][1])
    args = oracle_args(z, m)
    fwd = oracle.forward(**args, precision="f32")["p"]
    o_obs = oracle.forward(**dict(args, vp=z["in_vp_obs"]), precision="f32")["p"]
    g_bf, _, _ = oracle.gradient(**args, g_p=2.0 * (fwd - o_obs), chunk=chunk_len(m["nt"], m["segments"]),
                                 precision="f32", checkpoint_bf16=True)
    e_or = rel_l2(grads[torch.bfloat16][0], g_bf)
    e_32 = rel_l2(grads[torch.bfloat16][0], grads[torch.float32][0])
    e_64 = rel_l2(grads[torch.bfloat16][0], z["grad_vp"])
    print(f"  bf16-checkpoint grad: vs oracle bf16 emulation {e_or:.2e}, vs FP32-checkpoint grad "
          f"{e_32:.2e}, vs reference {e_64:.2e}")
    assert e_or <= TOL


@pytest.mark.parametrize("spb", [0, 1, 3])
def test_multitile_multishot_bitwise_vs_oracle(spb):
    """120 x 400 grid (many tiles in both directions), 7 shots grouped several per
    workgroup: the forward is deterministic run to run and bit-identical to the
    FP32 oracle."""
    print(f"[case] multi-tile multi-shot spb={spb}", flush=True)
    from geobrain_amd import synthetic
    nz, nx, nt, S = 120, 400, 500, 7
    vp = synthetic.grf_model(nz, nx, 5.0, seed=3)
    w = synthetic.ricker_padded(15.0, 5e-4, nt)
    damp = wave.bc_pml(nx, nz, 5.0, 5.0, 20, float(vp.max()), False).astype(np.float32)
    z = dict(in_vp=vp, in_rho=synthetic.gardner(vp), in_damp=damp,
             in_wavelet=np.stack([w * (1 + 0.1 * s) for s in range(S)]).astype(np.float32),
             in_src_x=synthetic.surface_shots(nx, S), in_src_z=np.arange(S) % 5,
             in_rcv_x=np.arange(0, nx, 3), in_rcv_z=np.full(len(range(0, nx, 3)), 4))
    m = dict(nz=nz, nx=nx, nt=nt, dx=5.0, dt=5e-4, nabc=20, free_surface=False, segments=1,
             snapshot_every=0)
    launch = K.LaunchChoice(0, spb) if spb else None
    with torch.no_grad():
        a, _ = run(z, m, launch=launch)
        b, _ = run(z, m, launch=launch)
    ref = oracle.forward(**oracle_args(z, m), precision="f32")
    for k in ("p", "vx", "vz"):
        same_run = float((a[k] == b[k]).float().mean())
        same_ref = float((a[k].cpu() == torch.as_tensor(ref[k])).float().mean())
        print(f"  {k}: repeat-run equal fraction {same_run:.6f}, oracle-equal fraction {same_ref:.6f}, "
              f"rel-L2 vs oracle {rel_l2(np_(a[k]), ref[k]):.2e}")
    for k in ("p", "vx", "vz"):
        np.testing.assert_array_equal(np_(a[k]), np_(b[k]))
        np.testing.assert_array_equal(np_(a[k]), ref[k])


@pytest.mark.parametrize("name", ["tiny1_grad", "tiny2_grad", "c2r_grad"])
def test_marching_recompute_equals_tiled(name):
    """The z-marching recompute stores the same brackets as the tiled one, so the
    gradients agree bit for bit (and both match the reference)."""
    print(f"[case] march recompute {name}", flush=True)
    z, m = load(name)
    want = ["vp"] + (["rho"] if "grad_rho" in z else []) + (["w"] if "grad_w" in z else [])
    g = adjoint_sources(z, m) if "g_seed" in m else None
    res = {}
    for march in (1, 2):
        out, ts = run(z, m, grad=want, launch=K.LaunchChoice(0, 0, march))
        if g is not None:
            loss = sum((_t(gk) * out[k]).sum() for gk, k in zip(g, ("p", "vx", "vz")))
        else:
            loss = (out["p"] ** 2).sum()
        loss.backward()
        res[march] = {k: np_(ts[k].grad) for k in want}
    for k in want:
        np.testing.assert_array_equal(res[1][k], res[2][k])
        if "g_seed" in m:
            assert rel_l2(res[2][k], z[{"vp": "grad_vp", "rho": "grad_rho", "w": "grad_w"}[k]]) <= TOL


def test_marching_recompute_multishot_segments():
    """7 shots on a 120 x 400 grid with the z range split into several marching
    segments: gradient identical to the tiled recompute."""
    print("[case] march recompute multishot", flush=True)
    from geobrain_amd import synthetic
    nz, nx, nt, S = 120, 400, 300, 7
    vp = synthetic.grf_model(nz, nx, 5.0, seed=4)
    w = synthetic.ricker_padded(15.0, 5e-4, nt)
    damp = wave.bc_pml(nx, nz, 5.0, 5.0, 20, float(vp.max()), False).astype(np.float32)
    z = dict(in_vp=vp, in_rho=synthetic.gardner(vp), in_damp=damp,
             in_wavelet=np.stack([w] * S).astype(np.float32),
             in_src_x=synthetic.surface_shots(nx, S), in_src_z=np.arange(S) % 4,
             in_rcv_x=np.arange(nx), in_rcv_z=np.full(nx, 3))
    res = {}
    for fs in (False, True):
        m = dict(nz=nz, nx=nx, nt=nt, dx=5.0, dt=5e-4, nabc=20, free_surface=fs, segments=6,
                 snapshot_every=0)
        for march in (1, 2):
            out, ts = run(z, m, grad=("vp", "rho"), launch=K.LaunchChoice(0, 0, march))
            (out["p"] ** 2).sum().backward()
            res[(fs, march)] = (np_(ts["vp"].grad), np_(ts["rho"].grad))
        for i in range(2):
            np.testing.assert_array_equal(res[(fs, 1)][i], res[(fs, 2)][i])


@pytest.mark.parametrize("march", [1, 2])
def test_free_surface_receivers_on_mirror_rows(march):
    """Free surface with receivers on the mirrored rows: unpadded rcv_z = -1 is
    padded row f-1, which the reference's indexing accepts; its vz sample is
    VZ[f-1] after the mirror VZ[f-1] = VZ[f] (acoustic_2d.py:267-274).  Source
    and receivers span tile boundaries.  Seismograms bit-identical to the FP32
    oracle; gradients (vp, rho, wavelet) from random adjoint sources on all
    three seismograms within TOL of the FP32 and FP64 oracles."""
    print(f"[case] free-surface receivers on row f-1, march={march}", flush=True)
    from geobrain_amd import synthetic
    nz, nx, nt, S, nabc = 60, 300, 240, 3, 10
    vp = synthetic.grf_model(nz, nx, 5.0, seed=5)
    w = synthetic.ricker_padded(15.0, 5e-4, nt)
    damp = wave.bc_pml(nx, nz, 5.0, 5.0, nabc, float(vp.max()), False).astype(np.float32)
    rx = np.arange(0, nx, 2)
    rz = np.fromiter(([-1, 0, 1, 2, 6][i % 5] for i in range(len(rx))), dtype=np.int64)
    z = dict(in_vp=vp, in_rho=synthetic.gardner(vp), in_damp=damp,
             in_wavelet=np.stack([w * (1 + 0.2 * s) for s in range(S)]).astype(np.float32),
             in_src_x=np.array([5, 150, 290]), in_src_z=np.array([-1, 1, 3]),
             in_rcv_x=rx, in_rcv_z=rz)
    m = dict(nz=nz, nx=nx, nt=nt, dx=5.0, dt=5e-4, nabc=nabc, free_surface=True, segments=4,
             snapshot_every=0)
    args = oracle_args(z, m)
    ref = oracle.forward(**args, chunk=chunk_len(nt, 4), precision="f32")
    out, ts = run(z, m, grad=("vp", "rho", "w"), launch=K.LaunchChoice(0, 0, march))
    for k in ("p", "vx", "vz"):
        bad = np.argwhere(np_(out[k]) != ref[k])
        print(f"  {k}: {len(bad)} samples differ from the FP32 oracle"
              + (f", first at (shot, step, receiver) {tuple(bad[0])}" if len(bad) else ""))
    for k in ("p", "vx", "vz"):
        np.testing.assert_array_equal(np_(out[k]), ref[k])
    g = tuple(np.random.default_rng(11).standard_normal((S, nt, len(rx))).astype(np.float32)
              for _ in range(3))
    loss = sum((_t(gk) * out[k]).sum() for gk, k in zip(g, ("p", "vx", "vz")))
    loss.backward()
    for prec in ("f32", "f64"):
        gv, gr, gw = oracle.gradient(**args, g_p=g[0], g_vx=g[1], g_vz=g[2],
                                     chunk=chunk_len(nt, 4), precision=prec)
        for key, want in (("vp", gv), ("rho", gr), ("w", gw)):
            e = rel_l2(np_(ts[key].grad), want)
            print(f"  grad {key}: rel-L2 vs {prec} oracle {e:.2e}")
            assert e <= TOL, (key, prec)


@pytest.mark.parametrize("fs,tile_rows", [(False, 0), (True, 0), (False, 24), (True, 24)])
def test_two_step_launches_equal_one_step(fs, tile_rows):
    """Two time steps per launch (temporal blocking) against one step per
    launch on a multi-tile grid with an odd step count and odd checkpoint /
    energy-chunk lengths, so launches of both kinds mix: seismograms and
    gradients bit-identical, energy maps within TOL (summation order differs),
    and seismograms bit-identical to the FP32 oracle."""
    print(f"[case] two-step vs one-step launches, free_surface={fs}, tile_rows={tile_rows}",
          flush=True)
    from geobrain_amd import synthetic
    nz, nx, nt, S = 70, 300, 241, 5
    vp = synthetic.grf_model(nz, nx, 5.0, seed=6)
    w = synthetic.ricker_padded(15.0, 5e-4, nt)
    damp = wave.bc_pml(nx, nz, 5.0, 5.0, 20, float(vp.max()), False).astype(np.float32)
    rx = np.arange(0, nx, 3)
    z = dict(in_vp=vp, in_rho=synthetic.gardner(vp), in_damp=damp,
             in_wavelet=np.stack([w * (1 + 0.1 * s) for s in range(S)]).astype(np.float32),
             in_src_x=synthetic.surface_shots(nx, S), in_src_z=np.arange(S) % 3,
             in_rcv_x=rx, in_rcv_z=np.full(len(rx), 2))
    m = dict(nz=nz, nx=nx, nt=nt, dx=5.0, dt=5e-4, nabc=20, free_surface=fs, segments=4,
             snapshot_every=0)
    res = {}
    for spl in (1, 2):
        out, ts = run(z, m, grad=("vp", "rho", "w"),
                      launch=K.LaunchChoice(tile_rows, 2, steps_per_launch=spl))
        loss = (out["p"] ** 2).sum() + (out["vz"] ** 2).sum()
        loss.backward()
        res[spl] = ({k: np_(out[k]) for k in out},
                    {k: np_(ts[k].grad) for k in ("vp", "rho", "w")})
    ref = oracle.forward(**oracle_args(z, m), chunk=chunk_len(nt, 4), precision="f32")
    for k in ("p", "vx", "vz"):
        np.testing.assert_array_equal(res[2][0][k], res[1][0][k])
        np.testing.assert_array_equal(res[2][0][k], ref[k])
    for k in ("forward_wavefield_p", "forward_wavefield_vx", "forward_wavefield_vz"):
        e = rel_l2(res[2][0][k], res[1][0][k])
        print(f"  {k}: rel-L2 two-step vs one-step {e:.2e}")
        assert e <= TOL, k
    for k in ("vp", "rho", "w"):
        np.testing.assert_array_equal(res[2][1][k], res[1][1][k])


@pytest.mark.parametrize("case", ["nt1", "nt3_fs", "tiny_grid", "corner_source_nabc1", "no_receivers"])
def test_edge_cases_vs_oracle(case):
    """Edge cases against the FP32 oracle: one and three time steps (single-step
    launches only), the smallest padded grid (6 x 7), a source on the padded
    corner (a cell no update touches, so injections accumulate) with nabc = 1,
    and no receivers at all.  Seismograms bit-identical, energy maps to 1e-6; vp
    gradients (where there are seismograms) within TOL of the FP32 oracle."""
    print(f"[case] edge case {case}", flush=True)
    from geobrain_amd import synthetic
    nz, nx, nt, S, nabc, fs = 30, 70, 40, 2, 5, False
    src_x, src_z = np.array([10, 50]), np.array([2, 3])
    if case == "nt1":
        nt = 1
    elif case == "nt3_fs":
        nt, fs = 3, True
    elif case == "tiny_grid":
        nz, nx, nabc, S = 2, 3, 2, 1
        src_x, src_z = np.array([1]), np.array([0])
    elif case == "corner_source_nabc1":
        nabc = 1
        src_x, src_z = np.array([-1, nx]), np.array([-1, nz])
    vp = synthetic.grf_model(nz, nx, 5.0, seed=8) if min(nz, nx) > 4 else \
        np.full((nz, nx), 2000.0, np.float32)
    w = synthetic.ricker_padded(25.0, 5e-4, nt) + np.float32(1e-3)
    rx = np.arange(0, nx, 2) if case != "no_receivers" else np.zeros(0, np.int64)
    # bc_pml divides by (nabc - 1) dx, as the reference does: nabc = 1 gets no damping
    damp = (wave.bc_pml(nx, nz, 5.0, 5.0, nabc, float(vp.max()), False).astype(np.float32)
            if nabc > 1 else np.zeros((nz + 2, nx + 2), np.float32))
    z = dict(in_vp=vp, in_rho=synthetic.gardner(vp), in_damp=damp,
             in_wavelet=np.stack([w * (1 + s) for s in range(S)]).astype(np.float32),
             in_src_x=src_x, in_src_z=src_z, in_rcv_x=rx,
             in_rcv_z=np.full(len(rx), 1, np.int64))
    m = dict(nz=nz, nx=nx, nt=nt, dx=5.0, dt=5e-4, nabc=nabc, free_surface=fs, segments=1,
             snapshot_every=0)
    args = oracle_args(z, m)
    ref = oracle.forward(**args, chunk=nt, precision="f32")
    grad = () if case == "no_receivers" else ("vp",)
    out, ts = run(z, m, grad=grad)
    for k in ("p", "vx", "vz"):
        got = np_(out[k])
        assert got.shape == ref[k].shape, (k, got.shape, ref[k].shape)
        np.testing.assert_array_equal(got, ref[k], err_msg=k)
    # energy maps: a two-step launch adds two steps' squares before adding them
    # to the map, so they agree to rounding (bitwise for one-step launches)
    for k in ("forward_wavefield_p", "forward_wavefield_vx", "forward_wavefield_vz"):
        got = np_(out[k])
        assert got.shape == ref[k].shape, (k, got.shape, ref[k].shape)
        assert rel_l2(got, ref[k]) <= 1e-6, k
    if grad:
        g = np.random.default_rng(3).standard_normal(ref["p"].shape).astype(np.float32)
        (_t(g) * out["p"]).sum().backward()
        gv, _, _ = oracle.gradient(**args, g_p=g, chunk=nt, precision="f32")
        e = rel_l2(np_(ts["vp"].grad), gv)
        print(f"  grad vp rel-L2 vs FP32 oracle {e:.2e}")
        assert e <= TOL
