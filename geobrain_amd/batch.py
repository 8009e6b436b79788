"""Batched forward modelling and gradients of many independent problems.

``BatchedForward`` runs the forward pass of ``n`` independent
``acoustic_forward_kernel`` problems -- different grids, step counts, shot
counts and wavelets -- as one launch sequence on the GPU: each launch advances
every problem that is still running by two time steps (the last step of a
problem with an odd step count is a single one).  The reference models one
problem per call (``geobrain/physics/wave/kernels/acoustic_2d.py:294-501``);
the seismograms here are those n calls' ``'p'``, ``'vx'``, ``'vz'`` outputs,
bit for bit.

The launch sequence allocates, copies and synchronises nothing, so
``capture()`` records it once into a HIP graph (``torch.cuda.CUDAGraph``) and
``replay()`` re-runs the whole batch with one graph launch.  Inputs are read
by ``setup()`` (coefficients, receiver maps); call it again after changing
them.  Per-problem latency comes from device wall-clock stamps the last
workgroup of each problem writes when the problem's final launch completes.

``BatchedGradient`` does the same for the gradient: the forward with
checkpoints, then the recompute and adjoint, of all problems as one launch
sequence each.  Problems are aligned at their last step, so their checkpoint
intervals coincide and the backward walks them together.  Each problem's
seismograms, vp gradient and wavelet gradient equal those of its own
``acoustic_forward_kernel`` call and autograd backward through
``torch.utils.checkpoint`` (``acoustic_2d.py:448-460``) for the same adjoint
sources.  Forward, the caller's misfit and backward capture into one graph.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native
from .physics.wave import kernels as K


@dataclass
class BatchLaunch:
    """``shots_per_block``: shots one workgroup steps through (0 = all shots of
    its problem); ``steps_per_launch``: 2 (default) or 1; ``tile_rows``: 24 for
    24 x 96 tiles, else 16 x 128."""
    shots_per_block: int = 0
    steps_per_launch: int = 0
    tile_rows: int = 0


class BatchedForward:
    """Forward passes of ``problems`` (each a dict of the keyword arguments of
    ``acoustic_forward_kernel``: ``nx, nz, dx, dz, nt, dt, nabc, free_surface,
    src_x, src_z, src_n, src_v, rcv_x, rcv_z, rcv_n, damp, vp, rho``) run
    together on one GPU."""

    def __init__(self, problems: Sequence[dict], launch: Optional[BatchLaunch] = None):
        if len(problems) == 0:
            raise ValueError("BatchedForward needs at least one problem")
        self.lib = _native.load()
        self.launch = launch or BatchLaunch()
        n = len(problems)
        self.n = n
        self.device = problems[0]["vp"].device
        self._probs = (_native.Problem * n)()
        self._ins = (_native.Inputs * n)()
        self._outs = (_native.Outputs * n)()
        self._keep: List[object] = []
        self.outputs: List[Dict[str, torch.Tensor]] = []
        f32 = dict(dtype=torch.float32, device=self.device)
        for i, a in enumerate(problems):
            vp = K._on_gpu(a["vp"], "vp", self.device).to(torch.float32).contiguous()
            rho = K._on_gpu(a["rho"], "rho", self.device).to(torch.float32).contiguous()
            src_v = K._on_gpu(a["src_v"], "src_v", self.device).to(torch.float32).contiguous()
            damp = K._on_gpu(a["damp"], "damp", self.device).to(torch.float32).contiguous()
            if a.get("dtype", torch.float32) != torch.float32:
                raise NotImplementedError("only float32 is supported")
            sx, sz = K._host_index(a["src_x"], "src_x"), K._host_index(a["src_z"], "src_z")
            rx, rz = K._host_index(a["rcv_x"], "rcv_x"), K._host_index(a["rcv_z"], "rcv_z")
            S, nt, R = int(a["src_n"]), int(a["nt"]), int(a["rcv_n"])
            # every shape check acoustic_forward_kernel makes, before any pointer is taken
            K.validate_problem(int(a["nx"]), int(a["nz"]), nt, int(a["nabc"]), S, R, sx, sz,
                               rx, rz, vp, rho, damp, src_v, where=f"problem {i}: ")
            p = self._probs[i]
            p.nx, p.nz = int(a["nx"]), int(a["nz"])
            p.dx, p.dz = float(a["dx"]), float(a["dz"])
            p.nt, p.dt = nt, float(a["dt"])
            p.nabc, p.free_surface = int(a["nabc"]), int(bool(a["free_surface"]))
            p.src_n, p.rcv_n = S, R
            p.wavelet_shared = int(src_v.dim() == 1)
            p.energy_chunk = nt
            p.shots_per_block = int(self.launch.shots_per_block)
            p.tile_rows = int(self.launch.tile_rows)
            p.steps_per_launch = int(self.launch.steps_per_launch)
            ins = self._ins[i]
            ins.src_x, ins.src_z = sx.ctypes.data, sz.ctypes.data
            ins.rcv_x = rx.ctypes.data if rx.size else None
            ins.rcv_z = rz.ctypes.data if rz.size else None
            ins.src_v, ins.damp = src_v.data_ptr(), damp.data_ptr()
            ins.vp, ins.rho = vp.data_ptr(), rho.data_ptr()
            out = {k: torch.zeros((S, nt, R), **f32) for k in ("p", "vx", "vz")}
            o = self._outs[i]
            o.rec_p, o.rec_vx, o.rec_vz = (out[k].data_ptr() for k in ("p", "vx", "vz"))
            self.outputs.append(out)
            self._keep += [vp, rho, src_v, damp, sx, sz, rx, rz]
        nbytes = self.lib.gbw_acoustic2d_batch_workspace_bytes(self._probs, n)
        if nbytes == 0:
            _native.check(-1)
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.ticks = torch.zeros(n + 1, dtype=torch.int64, device=self.device)
        # the library reports completion in nt order, longest first (stable)
        self.sorted_index = sorted(range(n), key=lambda i: -int(self._probs[i].nt))
        self.cells = sum(int(p.src_n) * (p.nz + 2 * p.nabc) * (p.nx + 2 * p.nabc) * int(p.nt)
                         for p in self._probs)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.setup()

    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def setup(self) -> None:
        """Validate geometry, build coefficients and receiver maps (synchronises)."""
        rc = self.lib.gbw_acoustic2d_batch_setup(self._probs, self._ins, self._outs, self.n,
                                                 self.ws.data_ptr(), self.ws.numel(),
                                                 self._stream())
        _native.check(rc)

    def run(self) -> List[Dict[str, torch.Tensor]]:
        """Enqueue the whole batch on the current stream; returns the outputs
        (``'p'``, ``'vx'``, ``'vz'`` per problem, in the caller's order)."""
        rc = self.lib.gbw_acoustic2d_batch_forward(self._probs, self.n, self.ws.data_ptr(),
                                                   self.ws.numel(), self.ticks.data_ptr(),
                                                   self._stream())
        _native.check(rc)
        return self.outputs

    def capture(self) -> torch.cuda.CUDAGraph:
        """Record ``run()`` into a HIP graph (``replay()`` launches it)."""
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def replay(self) -> List[Dict[str, torch.Tensor]]:
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.outputs

    def latencies_ms(self) -> np.ndarray:
        """Per-problem latency of the last run, in the caller's order: device
        wall clock from the start of the run to the end of the problem's last
        launch (synchronises)."""
        torch.cuda.synchronize(self.device)
        t = self.ticks.cpu().numpy().astype(np.int64)
        khz = self.lib.gbw_device_wall_clock_khz()
        if khz <= 0:
            raise RuntimeError("device wall-clock rate unavailable")
        lat_sorted = (t[1:] - t[0]) / float(khz)
        out = np.empty(self.n)
        for pos, i in enumerate(self.sorted_index):
            out[i] = lat_sorted[pos]
        return out


def problem_inputs(spec: dict, device: torch.device) -> dict:
    """The ``acoustic_forward_kernel`` arguments of one ``synthetic.uneven_batch``
    entry: layered Gaussian-random velocity, Gardner density, PML from the
    model's maximum velocity, ``n_shots`` surface shots with a Ricker wavelet at
    ``f0`` (amplitude varied per shot), receivers on every column at depth 2."""
    from . import synthetic
    from .physics.wave import bc_pml
    nz, nx, nt, S = spec["nz"], spec["nx"], spec["nt"], spec["n_shots"]
    dx, dt, nabc = spec["dx"], spec["dt"], spec.get("nabc", 20)
    vp = synthetic.grf_model(nz, nx, dx, seed=spec["seed"])
    w = synthetic.ricker_padded(spec["f0"], dt, nt)
    wav = np.stack([w * (1.0 + 0.05 * s) for s in range(S)]).astype(np.float32)
    damp = bc_pml(nx, nz, dx, dx, nabc, float(vp.max()), False).astype(np.float32)
    T = lambda a: torch.as_tensor(a, device=device)  # noqa: E731
    return dict(nx=nx, nz=nz, dx=dx, dz=dx, nt=nt, dt=dt, nabc=nabc,
                free_surface=bool(spec.get("free_surface", False)),
                src_x=synthetic.surface_shots(nx, S), src_z=np.full(S, 2, np.int64), src_n=S,
                src_v=T(wav), rcv_x=np.arange(nx, dtype=np.int64),
                rcv_z=np.full(nx, 2, np.int64), rcv_n=nx, damp=T(damp), vp=T(vp),
                rho=T(synthetic.gardner(vp)))


class BatchedGradient:
    """Gradients of ``n`` independent problems (each a dict of the keyword
    arguments of ``acoustic_forward_kernel``) w.r.t. their velocity models, run
    together on one GPU: ``forward()`` models every problem's seismograms (and
    keeps checkpoints in the workspace), the caller writes each problem's
    adjoint sources -- dLoss/d(seismogram) -- into ``adjoint[i]['p']`` (and
    ``'vx'`` / ``'vz'`` when requested), and ``backward()`` writes
    ``grad_vp[i]`` (and ``grad_src_v[i]`` with ``wavelet_grad``).

    Each problem's seismograms equal its own ``acoustic_forward_kernel`` call
    bit for bit, and its gradients equal what autograd through that call gives
    for the same adjoint sources (the reference computes them with autograd
    through ``torch.utils.checkpoint``, ``acoustic_2d.py:448-460``), up to the
    order of FP32 sums across shot groups.  ``forward`` and ``backward`` only
    enqueue kernels, so ``capture(fn)`` can record a whole misfit-gradient
    step (for example ``misfit_step``) into one HIP graph.

    Problems run aligned at their last time step, so checkpoint intervals
    (``checkpoint_every``, even, 0 = automatic) coincide across problems."""

    def __init__(self, problems: Sequence[dict], launch: Optional[BatchLaunch] = None,
                 checkpoint_every: int = 0, adjoint_fields: Sequence[str] = ("p",),
                 wavelet_grad: bool = False):
        if len(problems) == 0:
            raise ValueError("BatchedGradient needs at least one problem")
        if any(k not in ("p", "vx", "vz") for k in adjoint_fields) or not adjoint_fields:
            raise ValueError("adjoint_fields must name some of 'p', 'vx', 'vz'")
        self.lib = _native.load()
        self.launch = launch or BatchLaunch()
        n = len(problems)
        self.n = n
        self.device = problems[0]["vp"].device
        self._probs = (_native.Problem * n)()
        self._ins = (_native.Inputs * n)()
        self._outs = (_native.Outputs * n)()
        self._adjs = (_native.Adjoint * n)()
        self._keep: List[object] = []
        self.outputs: List[Dict[str, torch.Tensor]] = []
        self.adjoint: List[Dict[str, torch.Tensor]] = []
        self.grad_vp: List[torch.Tensor] = []
        self.grad_src_v: List[Optional[torch.Tensor]] = []
        f32 = dict(dtype=torch.float32, device=self.device)
        for i, a in enumerate(problems):
            vp = K._on_gpu(a["vp"], "vp", self.device).detach().to(torch.float32).contiguous()
            rho = K._on_gpu(a["rho"], "rho", self.device).detach().to(torch.float32).contiguous()
            src_v = K._on_gpu(a["src_v"], "src_v", self.device).detach().to(torch.float32).contiguous()
            damp = K._on_gpu(a["damp"], "damp", self.device).detach().to(torch.float32).contiguous()
            if a.get("dtype", torch.float32) != torch.float32:
                raise NotImplementedError("only float32 is supported")
            sx, sz = K._host_index(a["src_x"], "src_x"), K._host_index(a["src_z"], "src_z")
            rx, rz = K._host_index(a["rcv_x"], "rcv_x"), K._host_index(a["rcv_z"], "rcv_z")
            S, nt, R = int(a["src_n"]), int(a["nt"]), int(a["rcv_n"])
            K.validate_problem(int(a["nx"]), int(a["nz"]), nt, int(a["nabc"]), S, R, sx, sz,
                               rx, rz, vp, rho, damp, src_v, where=f"problem {i}: ")
            p = self._probs[i]
            p.nx, p.nz = int(a["nx"]), int(a["nz"])
            p.dx, p.dz = float(a["dx"]), float(a["dz"])
            p.nt, p.dt = nt, float(a["dt"])
            p.nabc, p.free_surface = int(a["nabc"]), int(bool(a["free_surface"]))
            p.src_n, p.rcv_n = S, R
            p.wavelet_shared = int(src_v.dim() == 1)
            p.energy_chunk = nt
            p.checkpoint_every = int(checkpoint_every)
            p.shots_per_block = int(self.launch.shots_per_block)
            p.tile_rows = int(self.launch.tile_rows)
            p.steps_per_launch = 0
            ins = self._ins[i]
            ins.src_x, ins.src_z = sx.ctypes.data, sz.ctypes.data
            ins.rcv_x = rx.ctypes.data if rx.size else None
            ins.rcv_z = rz.ctypes.data if rz.size else None
            ins.src_v, ins.damp = src_v.data_ptr(), damp.data_ptr()
            ins.vp, ins.rho = vp.data_ptr(), rho.data_ptr()
            out = {k: torch.zeros((S, nt, R), **f32) for k in ("p", "vx", "vz")}
            o = self._outs[i]
            o.rec_p, o.rec_vx, o.rec_vz = (out[k].data_ptr() for k in ("p", "vx", "vz"))
            adj = {k: torch.zeros((S, nt, R), **f32) for k in adjoint_fields}
            gv = torch.zeros_like(vp)
            gw = torch.zeros_like(src_v) if wavelet_grad else None
            d = self._adjs[i]
            d.g_p = adj["p"].data_ptr() if "p" in adj else None
            d.g_vx = adj["vx"].data_ptr() if "vx" in adj else None
            d.g_vz = adj["vz"].data_ptr() if "vz" in adj else None
            d.grad_vp = gv.data_ptr()
            d.grad_src_v = gw.data_ptr() if gw is not None else None
            self.outputs.append(out)
            self.adjoint.append(adj)
            self.grad_vp.append(gv)
            self.grad_src_v.append(gw)
            self._keep += [vp, rho, src_v, damp, sx, sz, rx, rz]
        nbytes = self.lib.gbw_acoustic2d_batch_gradient_workspace_bytes(self._probs, n)
        if nbytes == 0:
            _native.check(-1)
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.ticks = torch.zeros(n + 1, dtype=torch.int64, device=self.device)
        self.sorted_index = sorted(range(n), key=lambda i: -int(self._probs[i].nt))
        self.cells = sum(int(p.src_n) * (p.nz + 2 * p.nabc) * (p.nx + 2 * p.nabc) * int(p.nt)
                         for p in self._probs)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.setup()

    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def setup(self) -> None:
        """Validate geometry, build coefficients and receiver maps (synchronises)."""
        rc = self.lib.gbw_acoustic2d_batch_gradient_setup(
            self._probs, self._ins, self._outs, self._adjs, self.n, self.ws.data_ptr(),
            self.ws.numel(), self._stream())
        _native.check(rc)

    def forward(self) -> List[Dict[str, torch.Tensor]]:
        """Enqueue every problem's forward pass; returns the seismograms
        (``'p'``, ``'vx'``, ``'vz'`` per problem, in the caller's order)."""
        rc = self.lib.gbw_acoustic2d_batch_gradient_forward(
            self._probs, self.n, self.ws.data_ptr(), self.ws.numel(), self.ticks.data_ptr(),
            self._stream())
        _native.check(rc)
        return self.outputs

    def backward(self) -> List[torch.Tensor]:
        """Enqueue the recompute and adjoint of every problem from the adjoint
        sources in ``adjoint``; returns ``grad_vp`` (caller's order)."""
        rc = self.lib.gbw_acoustic2d_batch_gradient_backward(
            self._probs, self.n, self.ws.data_ptr(), self.ws.numel(), self.ticks.data_ptr(),
            self._stream())
        _native.check(rc)
        return self.grad_vp

    def misfit_step(self, d_obs: Sequence[torch.Tensor]) -> List[torch.Tensor]:
        """Gradient of ``sum_i sum((p_i - d_obs_i)^2)``: forward, adjoint
        sources ``2 (p_i - d_obs_i)``, backward (nothing synchronises)."""
        self.forward()
        for out, adj, d in zip(self.outputs, self.adjoint, d_obs):
            torch.sub(out["p"], d, out=adj["p"])
            adj["p"].mul_(2.0)
        return self.backward()

    def capture(self, fn) -> torch.cuda.CUDAGraph:
        """Record ``fn()`` (e.g. ``lambda: self.misfit_step(d_obs)``) into a HIP
        graph; ``replay()`` launches it."""
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        torch.cuda.synchronize(self.device)
        self.graph = g
        return g

    def replay(self) -> List[torch.Tensor]:
        if self.graph is None:
            raise RuntimeError("capture() a step first")
        self.graph.replay()
        return self.grad_vp

    def latencies_ms(self) -> np.ndarray:
        """Per-problem latency of the last forward + backward, in the caller's
        order: device wall clock from the start of the forward to the end of
        the problem's gradient (synchronises)."""
        torch.cuda.synchronize(self.device)
        t = self.ticks.cpu().numpy().astype(np.int64)
        khz = self.lib.gbw_device_wall_clock_khz()
        if khz <= 0:
            raise RuntimeError("device wall-clock rate unavailable")
        lat_sorted = (t[1:] - t[0]) / float(khz)
        out = np.empty(self.n)
        for pos, i in enumerate(self.sorted_index):
            out[i] = lat_sorted[pos]
        return out
