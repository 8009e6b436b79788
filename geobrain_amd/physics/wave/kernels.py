"""2D acoustic FDTD operator on MI355X, drop-in for ``geobrain.physics.wave.kernels``.

``acoustic_forward_kernel`` keeps the reference's signature, argument meaning,
defaults and output keys (``geobrain/physics/wave/kernels/acoustic_2d.py:294-501``),
but every time step, the receiver sampling, the energy maps and the gradient
run in the hand-written HIP library ``libgbw_acoustic2d.so``:

* forward: ``gbw_acoustic2d_forward`` -- one fused stencil launch per step;
* gradient: a ``torch.autograd.Function`` whose backward calls
  ``gbw_acoustic2d_backward`` -- the exact discrete adjoint with forward
  checkpoints every ``k`` steps and recomputation in between, replacing
  autograd through ``torch.utils.checkpoint`` (``acoustic_2d.py:448-460``).

``checkpoint_segments`` keeps its meaning: the checkpoint interval is the
``torch.chunk`` chunk length ``ceil(nt / checkpoint_segments)`` and the energy
maps are summed per chunk as the reference sums them.  With
``checkpoint_segments == 1`` the interval is chosen to balance checkpoint and
recompute memory (``~sqrt(3 nt)``); results do not depend on it.

Inputs must live on a ROCm GPU.  There is no CPU path: a CPU tensor, a
missing library or a missing GPU raises ``RuntimeError``.
"""

from __future__ import annotations

import collections
import ctypes
import math
import weakref
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np
import torch
from torch import Tensor

from ... import _native


def pad_model(v: Tensor, pml: int, nz: int, nx: int,
              device: torch.device = torch.device("cpu")) -> Tensor:
    """Edge-replicating pad of ``(nz, nx)`` to ``(nz + 2 pml, nx + 2 pml)``
    (``acoustic_2d.py:18-66``): cell ``(i, j)`` takes ``v[clamp(i - pml), clamp(j - pml)]``.
    The HIP library applies the same map on the fly; this is the tensor form."""
    ii = torch.arange(nz + 2 * pml, device=v.device).sub(pml).clamp_(0, nz - 1)
    jj = torch.arange(nx + 2 * pml, device=v.device).sub(pml).clamp_(0, nx - 1)
    return v.to(torch.float32)[ii][:, jj].to(device)


def chunk_length(nt: int, checkpoint_segments: int) -> int:
    """Chunk size ``torch.chunk(src_v, checkpoint_segments, -1)`` uses."""
    return -(-nt // max(1, int(checkpoint_segments)))


def auto_checkpoint_interval(nt: int) -> int:
    """Interval minimising checkpoints (12 B/cell each, ``nt/k`` of them) plus the
    per-interval bracket store (4 B/cell per step, ``k`` of them)."""
    return max(1, min(nt, int(round(math.sqrt(3.0 * nt)))))


@dataclass
class LaunchChoice:
    """Launch parameters the library accepts: ``tile_rows`` (16 or 32 rows for
    one-step launches; 24 selects 24 x 96 tiles for two-step launches),
    ``shots_per_block`` and ``march`` (gradient: 1 tiles, 2 z-marching
    recompute strips, 3 z-marching two-step adjoint wavefronts),
    ``steps_per_launch`` (1 or 2 time steps per forward / recompute launch),
    ``fwd_march`` (two-step forward and recompute launches: 1 tiles in shot
    groups, 2 z-marching wavefronts, 3 two-phase tiles: the first multiple of
    256 tiles in shot groups and the rest in groups of 8 shots); zeros select
    the library's defaults.
    ``bwd_tile_rows`` and
    ``bwd_shots_per_block`` apply to the gradient call (0 = same as the
    forward's): the forward and adjoint kernels have different register
    budgets, so their best tiles differ.  The autotuner fills these."""
    tile_rows: int = 0
    shots_per_block: int = 0
    march: int = 0
    bwd_tile_rows: int = 0
    bwd_shots_per_block: int = 0
    steps_per_launch: int = 0
    fwd_march: int = 0


_launch_override: Optional[LaunchChoice] = None
# choices the autotuner found, per problem shape (see ``shape_key``)
_tuned: Dict[tuple, LaunchChoice] = {}


def set_launch_choice(choice: Optional[LaunchChoice]) -> None:
    """Pin the launch parameters for subsequent calls (``None`` = library default).
    A pinned choice applies to every problem shape and takes precedence over
    the tuned choices registered with ``register_tuned``."""
    global _launch_override
    _launch_override = choice


def get_launch_choice() -> Optional[LaunchChoice]:
    """The pinned launch choice (``None`` when none is pinned)."""
    return _launch_override


def shape_key(nz: int, nx: int, nabc: int, src_n: int, rcv_n: int, free_surface: bool,
              shared: bool) -> tuple:
    """The problem shape a tuned launch choice is valid for."""
    return (int(nz) + 2 * int(nabc), int(nx) + 2 * int(nabc), int(src_n), int(rcv_n),
            bool(free_surface), bool(shared))


def register_tuned(key: tuple, choice: Optional[LaunchChoice]) -> None:
    """Use ``choice`` for calls whose ``shape_key`` is ``key`` (``None`` forgets it)."""
    if choice is None:
        _tuned.pop(key, None)
    else:
        _tuned[key] = choice


@dataclass
class _Spec:
    prob: _native.Problem
    src_x: np.ndarray
    src_z: np.ndarray
    rcv_x: np.ndarray
    rcv_z: np.ndarray
    n_snap: int
    bwd_launch: tuple = (0, 0)  # (tile_rows, shots_per_block) of the gradient call
    keep: Dict[str, object] = field(default_factory=dict)


def _stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _ptr(t: Optional[Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _inputs(spec: _Spec, src_v: Tensor, damp: Tensor, vp: Tensor, rho: Tensor) -> _native.Inputs:
    ins = _native.Inputs()
    ins.src_x = spec.src_x.ctypes.data
    ins.src_z = spec.src_z.ctypes.data
    ins.rcv_x = spec.rcv_x.ctypes.data if spec.rcv_x.size else None
    ins.rcv_z = spec.rcv_z.ctypes.data if spec.rcv_z.size else None
    ins.src_v = src_v.data_ptr()
    ins.damp = damp.data_ptr()
    ins.vp = vp.data_ptr()
    ins.rho = rho.data_ptr()
    geo = spec.keep.get("geometry")
    ins.geometry = geo.data_ptr() if geo is not None else None
    return ins


# Prepared geometry buffers (source cells and receiver maps on the device,
# gbw_acoustic2d_geometry_setup), keyed by the grid and the position values:
# a call that repeats a geometry does no host work on it, so the forward and
# gradient calls then only enqueue kernels and can be captured in a HIP graph.
_GEOMETRY_CACHE_SIZE = 8
_geometry_cache: "collections.OrderedDict[tuple, Tensor]" = collections.OrderedDict()


def _geometry(prob: _native.Problem, sx, sz, rx, rz, device: torch.device) -> Tensor:
    key = (str(device), prob.nz, prob.nx, prob.nabc, prob.free_surface, prob.src_n, prob.rcv_n,
           prob.src_points, sx.tobytes(), sz.tobytes(), rx.tobytes(), rz.tobytes())
    hit = _geometry_cache.get(key)
    if hit is not None:
        _geometry_cache.move_to_end(key)
        return hit
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError(
            "acoustic_forward_kernel: this source/receiver geometry has not been prepared; "
            "run the call once before capturing it in a graph")
    lib = _native.load()
    nbytes = lib.gbw_acoustic2d_geometry_bytes(ctypes.byref(prob))
    if nbytes == 0:
        _native.check(-1)
    buf = torch.empty(nbytes, dtype=torch.uint8, device=device)
    ins = _native.Inputs()
    ins.src_x, ins.src_z = sx.ctypes.data, sz.ctypes.data
    ins.rcv_x = rx.ctypes.data if rx.size else None
    ins.rcv_z = rz.ctypes.data if rz.size else None
    _native.check(lib.gbw_acoustic2d_geometry_setup(ctypes.byref(prob), ctypes.byref(ins),
                                                    buf.data_ptr(), nbytes,
                                                    _stream_handle(device)))
    _geometry_cache[key] = buf
    while len(_geometry_cache) > _GEOMETRY_CACHE_SIZE:
        _geometry_cache.popitem(last=False)
    return buf


class AcousticFunction(torch.autograd.Function):
    """Seismograms and energy maps as a differentiable function of
    ``(vp, rho, src_v)``; ``damp`` and the geometry are constants."""

    @staticmethod
    def forward(ctx, vp, rho, src_v, damp, spec: _Spec):
        lib = _native.load()
        prob = spec.prob
        dev = vp.device
        S, nt, R, nz, nx = prob.src_n, prob.nt, prob.rcv_n, prob.nz, prob.nx
        need_grad = any(ctx.needs_input_grad[:3])
        if not need_grad:
            prob.checkpoint_every = 0
        ws_bytes = lib.gbw_acoustic2d_workspace_bytes(ctypes.byref(prob), 0)
        if ws_bytes == 0:
            _native.check(-1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        ck = None
        if need_grad:
            ck = torch.empty(lib.gbw_acoustic2d_checkpoint_bytes(ctypes.byref(prob)),
                             dtype=torch.uint8, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        p = torch.empty((S, nt, R), **f32)
        vx = torch.empty((S, nt, R), **f32)
        vz = torch.empty((S, nt, R), **f32)
        ep = torch.empty((nz, nx), **f32)
        evx = torch.empty((nz, nx), **f32)
        evz = torch.empty((nz, nx), **f32)
        snaps = torch.empty((spec.n_snap, nz, nx), **f32) if spec.n_snap else torch.empty(0, **f32)
        outs = _native.Outputs()
        outs.rec_p, outs.rec_vx, outs.rec_vz = _ptr(p), _ptr(vx), _ptr(vz)
        outs.energy_p, outs.energy_vx, outs.energy_vz = _ptr(ep), _ptr(evx), _ptr(evz)
        outs.snapshots = _ptr(snaps) if spec.n_snap else None
        outs.checkpoints = _ptr(ck)
        ins = _inputs(spec, src_v, damp, vp, rho)
        rc = lib.gbw_acoustic2d_forward(ctypes.byref(prob), ctypes.byref(ins), ctypes.byref(outs),
                                        ws.data_ptr(), ws_bytes, _stream_handle(dev))
        _native.check(rc)
        ctx.mark_non_differentiable(ep, evx, evz, snaps)
        ctx.spec = spec
        ctx.ck = ck
        ctx.save_for_backward(vp, rho, src_v, damp)
        return p, vx, vz, ep, evx, evz, snaps

    @staticmethod
    def backward(ctx, g_p, g_vx, g_vz, *unused):
        lib = _native.load()
        vp, rho, src_v, damp = ctx.saved_tensors
        spec = ctx.spec
        prob = spec.prob
        dev = vp.device

        def prep(g):
            return None if g is None else g.to(torch.float32).contiguous()

        g_p, g_vx, g_vz = prep(g_p), prep(g_vx), prep(g_vz)
        prob = _native.Problem.from_buffer_copy(prob)
        prob.tile_rows, prob.shots_per_block = spec.bwd_launch
        want_vp, want_rho, want_w = ctx.needs_input_grad[:3]
        grad_vp = torch.empty_like(vp) if want_vp else None
        grad_rho = torch.empty_like(rho) if want_rho else None
        grad_w = torch.empty_like(src_v) if want_w else None
        adj = _native.Adjoint()
        adj.g_p, adj.g_vx, adj.g_vz = _ptr(g_p), _ptr(g_vx), _ptr(g_vz)
        adj.checkpoints = _ptr(ctx.ck)
        adj.grad_vp, adj.grad_rho, adj.grad_src_v = _ptr(grad_vp), _ptr(grad_rho), _ptr(grad_w)
        ws_bytes = lib.gbw_acoustic2d_workspace_bytes(ctypes.byref(prob), 2 if want_rho else 1)
        if ws_bytes == 0:
            _native.check(-1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        ins = _inputs(spec, src_v, damp, vp, rho)
        rc = lib.gbw_acoustic2d_backward(ctypes.byref(prob), ctypes.byref(ins), ctypes.byref(adj),
                                         ws.data_ptr(), ws_bytes, _stream_handle(dev))
        _native.check(rc)
        ctx.ck = None
        return grad_vp, grad_rho, grad_w, None, None


# host copies of device position tensors, keyed by the tensor (weakly) and
# checked against its version counter and storage, so a caller that passes the
# same device positions every call (the propagator, an FWI loop) pays one
# device-to-host copy, not one per call
# (tensors compare elementwise, so the cache is keyed by id() and each entry
# holds a weak reference that both confirms identity and drops the entry when
# the tensor is freed)
_index_cache: Dict[int, tuple] = {}
# device-to-host copies of position tensors made so far (the cache's misses)
host_index_copies = 0


def _host_index(t, name: str, points: bool = False) -> np.ndarray:
    """Host int64 copy of a position array (1-D; with ``points``, also the
    (src_n, K) form of several source points per shot)."""
    if torch.is_tensor(t):
        if t.device.type != "cpu":
            stamp = (t._version, t.data_ptr(), tuple(t.shape), tuple(t.stride()), t.dtype)
            hit = _index_cache.get(id(t))
            if hit is not None and hit[0]() is t and hit[1] == stamp:
                return hit[2]
            global host_index_copies
            host_index_copies += 1
            a = t.detach().to("cpu").numpy()
        else:
            a = t.detach().numpy()
    else:
        a = np.asarray(t)
    if a.ndim != 1 and not (points and a.ndim == 2):
        raise NotImplementedError(
            f"{name}: only one point per shot (1-D positions) is supported here; got shape "
            f"{a.shape}")
    a = np.ascontiguousarray(a, dtype=np.int64)
    if torch.is_tensor(t) and t.device.type != "cpu":
        a.setflags(write=False)
        key = id(t)
        _index_cache[key] = (weakref.ref(t, lambda _r, k=key: _index_cache.pop(k, None)), stamp, a)
    return a


def validate_problem(nx, nz, nt, nabc, src_n, rcv_n, sx, sz, rx, rz, vp, rho, damp,
                     src_v, where: str = "") -> int:
    """Shape and length checks of one 2D problem, made on the host before any
    device pointer is handed to the library (a short wavelet or an undersized
    model would otherwise be read past its end on the device).  ``sx .. rz``
    are the host int64 position arrays; source positions of shape
    ``(src_n, K)`` put K points in each shot (``acoustic_2d.py:234-237``).
    Returns ``wavelet_shared`` (1 for a ``(nt,)`` wavelet, 0 for
    ``(src_n, nt)``, 2 for ``(src_n, K, nt)``, one wavelet per point).  Raises
    ``ValueError`` or ``NotImplementedError``; ``where`` prefixes the
    message."""
    if sx.shape != sz.shape or sx.shape[0] != src_n:
        raise ValueError(f"{where}src_x/src_z must both have src_n={src_n} entries")
    if sx.ndim == 2:
        k = sx.shape[1]
        if k < 1:
            raise ValueError(f"{where}src_x/src_z of shape (src_n, K) need K >= 1")
        if src_v.dim() == 2 and tuple(src_v.shape) == (src_n, nt):
            return 0
        if src_v.dim() == 3 and tuple(src_v.shape) == (src_n, k, nt):
            return 2
        raise ValueError(f"{where}with (src_n, K) source positions src_v must have shape "
                         f"({src_n}, {nt}) or ({src_n}, {k}, {nt})")
    if rx.shape != rz.shape or rx.shape[0] != rcv_n:
        raise ValueError(f"{where}rcv_x/rcv_z must both have rcv_n={rcv_n} entries")
    if tuple(vp.shape) != (nz, nx) or tuple(rho.shape) != (nz, nx):
        raise ValueError(f"{where}vp and rho must have shape ({nz}, {nx})")
    if tuple(damp.shape) != (nz + 2 * nabc, nx + 2 * nabc):
        raise ValueError(f"{where}damp must have shape ({nz + 2 * nabc}, {nx + 2 * nabc})")
    if src_v.dim() == 1:
        if src_v.shape[0] != nt:
            raise ValueError(f"{where}src_v has {src_v.shape[0]} samples, nt={nt}")
        return 1
    if src_v.dim() == 2:
        if tuple(src_v.shape) != (src_n, nt):
            raise ValueError(f"{where}src_v must have shape ({src_n}, {nt}) or ({nt},)")
        return 0
    raise NotImplementedError(f"{where}src_v must be (src_n, nt) or (nt,)")


def _on_gpu(t: Tensor, name: str, device: torch.device) -> Tensor:
    if not torch.is_tensor(t):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(
            f"geobrain_amd runs the acoustic propagator on a ROCm GPU; {name} is on {t.device}")
    if t.device != device:
        raise RuntimeError(f"{name} is on {t.device}, expected {device}")
    return t


def acoustic_forward_kernel(
    nx: int,
    nz: int,
    dx: float,
    dz: float,
    nt: int,
    dt: float,
    nabc: int,
    free_surface: bool,
    src_x: Tensor,
    src_z: Tensor,
    src_n: int,
    src_v: Tensor,
    rcv_x: Tensor,
    rcv_z: Tensor,
    rcv_n: int,
    damp: Tensor,
    vp: Tensor,
    rho: Tensor,
    checkpoint_segments: int = 1,
    device: torch.device = torch.device('cpu'),
    dtype: torch.dtype = torch.float32,
    snapshot_every: int = 0,
    *,
    checkpoint_dtype: torch.dtype = torch.float32,
) -> Dict[str, Tensor]:
    """Forward simulation (signature of ``acoustic_2d.py:294-317``).

    Returns ``'p', 'vx', 'vz'`` ``(src_n, nt, rcv_n)``, the energy maps
    ``'forward_wavefield_p/vx/vz'`` ``(nz, nx)`` (detached) and, when
    ``snapshot_every > 0``, ``'snapshots'`` ``(n_snaps, nz, nx)`` on the CPU.
    Gradients flow to ``vp``, ``rho`` and ``src_v`` when they require grad.

    ``checkpoint_dtype`` (keyword-only; not in the reference) stores the
    forward checkpoints the gradient recomputes from as ``torch.bfloat16``
    (per-field, per-shot power-of-two FP32 scale) instead of FP32; the time
    stepping stays FP32.  It changes the gradient slightly, never the forward
    outputs.
    """
    if dtype != torch.float32:
        raise NotImplementedError("geobrain_amd computes in float32 only (the reference default)")
    if checkpoint_dtype not in (torch.float32, torch.bfloat16):
        raise ValueError("checkpoint_dtype must be torch.float32 or torch.bfloat16")
    dev = vp.device
    for name, t in (("vp", vp), ("rho", rho), ("src_v", src_v), ("damp", damp)):
        _on_gpu(t, name, dev)
    if not torch.cuda.is_available():
        raise RuntimeError("geobrain_amd: no ROCm GPU is visible")
    sx, sz = _host_index(src_x, "src_x", True), _host_index(src_z, "src_z", True)
    rx, rz = _host_index(rcv_x, "rcv_x"), _host_index(rcv_z, "rcv_z")
    shared = validate_problem(nx, nz, nt, nabc, src_n, rcv_n, sx, sz, rx, rz, vp, rho, damp,
                              src_v)
    points = sx.shape[1] if sx.ndim == 2 else 1

    prob = _native.Problem()
    prob.nx, prob.nz = int(nx), int(nz)
    prob.dx, prob.dz = float(dx), float(dz)
    prob.nt, prob.dt = int(nt), float(dt)
    prob.nabc, prob.free_surface = int(nabc), int(bool(free_surface))
    prob.src_n, prob.rcv_n = int(src_n), int(rcv_n)
    prob.wavelet_shared = shared
    prob.src_points = int(points)
    if snapshot_every > 0:
        prob.energy_chunk = int(snapshot_every)
        prob.snapshot_every = int(snapshot_every)
        n_snap = -(-nt // snapshot_every)
    else:
        prob.energy_chunk = chunk_length(nt, checkpoint_segments)
        prob.snapshot_every = 0
        n_snap = 0
    if snapshot_every <= 0 and checkpoint_segments > 1:
        prob.checkpoint_every = chunk_length(nt, checkpoint_segments)
    else:
        prob.checkpoint_every = auto_checkpoint_interval(nt)
    prob.checkpoint_bf16 = int(checkpoint_dtype == torch.bfloat16)
    choice = _launch_override
    if choice is None and _tuned:
        choice = _tuned.get(shape_key(nz, nx, nabc, src_n, rcv_n, free_surface, shared))
    prob.tile_rows = choice.tile_rows if choice else 0
    prob.shots_per_block = choice.shots_per_block if choice else 0
    prob.march = choice.march if choice else 0
    prob.steps_per_launch = choice.steps_per_launch if choice else 0
    prob.fwd_march = choice.fwd_march if choice else 0
    bwd_launch = (prob.tile_rows, prob.shots_per_block)
    if choice and (choice.bwd_tile_rows or choice.bwd_shots_per_block):
        bwd_launch = (choice.bwd_tile_rows, choice.bwd_shots_per_block)
    spec = _Spec(prob=prob, src_x=sx, src_z=sz, rcv_x=rx, rcv_z=rz, n_snap=n_snap,
                 bwd_launch=bwd_launch)
    spec.keep["geometry"] = _geometry(prob, sx, sz, rx, rz, dev)

    f32 = torch.float32
    p, vx, vz, ep, evx, evz, snaps = AcousticFunction.apply(
        vp.to(f32).contiguous(), rho.to(f32).contiguous(), src_v.to(f32).contiguous(),
        damp.to(f32).contiguous(), spec)
    out = {
        'p': p, 'vx': vx, 'vz': vz,
        'forward_wavefield_p': ep,
        'forward_wavefield_vx': evx,
        'forward_wavefield_vz': evz,
    }
    if snapshot_every > 0:
        out['snapshots'] = snaps.cpu()
    return out


# ---------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------

def pad_model_3d(v: Tensor, pml: int, nz: int, ny: int, nx: int,
                 device: torch.device = torch.device("cpu")) -> Tensor:
    """Edge-replicating pad of ``(nz, ny, nx)`` to ``(nz + 2 pml, ny + 2 pml, nx + 2 pml)``
    (``acoustic_2d.py:69-120``): cell ``(i, j, k)`` takes
    ``v[clamp(i - pml), clamp(j - pml), clamp(k - pml)]``.  The HIP library
    applies the same map on the fly; this is the tensor form."""
    ii = torch.arange(nz + 2 * pml, device=v.device).sub(pml).clamp_(0, nz - 1)
    jj = torch.arange(ny + 2 * pml, device=v.device).sub(pml).clamp_(0, ny - 1)
    kk = torch.arange(nx + 2 * pml, device=v.device).sub(pml).clamp_(0, nx - 1)
    return v.to(torch.float32)[ii][:, jj][:, :, kk].to(device)


@dataclass
class LaunchChoice3D:
    """Launch parameters of the 3D library calls: ``step_kernel`` 0 (separate
    p and v kernels, default), 1 (fused z-marching step, tiles in workgroups)
    or 2 (fused step, rows in wavefronts; forward pass only);
    ``shots_per_block`` (any with the p / v kernels; 1, 2 or 4 with the
    z-marching kernel; 1 to 8 with the fused step); ``z_chunk`` (z
    planes per z-marching workgroup or fused-step segment).  Zeros select the
    library's defaults."""
    shots_per_block: int = 0
    step_kernel: int = 0
    z_chunk: int = 0


_launch3: LaunchChoice3D = LaunchChoice3D()


def set_launch_choice_3d(choice: Optional[LaunchChoice3D]) -> None:
    """Pin the 3D launch parameters for subsequent calls (``None`` = defaults)."""
    global _launch3
    _launch3 = choice if choice is not None else LaunchChoice3D()


def set_shots_per_block_3d(n: int) -> None:
    """Shots per workgroup of the 3D kernels (0 = library default)."""
    global _launch3
    _launch3 = LaunchChoice3D(int(n), _launch3.step_kernel, _launch3.z_chunk)


def auto_checkpoint_interval_3d(nt: int, rho_grad: bool) -> int:
    """Interval minimising checkpoints (16 B/cell each, ``nt/k`` of them) plus the
    per-interval store the gradient recomputes (4 B/cell per step, 8 with the
    density gradient, ``k`` of them)."""
    per_step = 8.0 if rho_grad else 4.0
    return max(1, min(nt, int(round(math.sqrt(16.0 * nt / per_step)))))


@dataclass
class _Spec3:
    prob: _native.Problem3
    idx: tuple  # host int64 (src_x, src_y, src_z, rcv_x, rcv_y, rcv_z)


def _inputs3(spec: _Spec3, src_v: Tensor, damp: Tensor, vp: Tensor, rho: Tensor) -> _native.Inputs3:
    sx, sy, sz, rx, ry, rz = spec.idx
    ins = _native.Inputs3()
    ins.src_x, ins.src_y, ins.src_z = sx.ctypes.data, sy.ctypes.data, sz.ctypes.data
    if rx.size:
        ins.rcv_x, ins.rcv_y, ins.rcv_z = rx.ctypes.data, ry.ctypes.data, rz.ctypes.data
    ins.src_v, ins.damp = src_v.data_ptr(), damp.data_ptr()
    ins.vp, ins.rho = vp.data_ptr(), rho.data_ptr()
    return ins


class Acoustic3DFunction(torch.autograd.Function):
    """3D seismograms and energy maps as a differentiable function of
    ``(vp, rho, src_v)``; ``damp`` and the geometry are constants."""

    @staticmethod
    def forward(ctx, vp, rho, src_v, damp, spec: _Spec3):
        lib = _native.load()
        prob = spec.prob
        dev = vp.device
        S, nt, R = prob.src_n, prob.nt, prob.rcv_n
        nz, ny, nx = prob.nz, prob.ny, prob.nx
        need_grad = any(ctx.needs_input_grad[:3])
        if not need_grad:
            prob.checkpoint_every = 0
        ws_bytes = lib.gbw_acoustic3d_workspace_bytes(ctypes.byref(prob), 0)
        if ws_bytes == 0:
            _native.check(-1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        ck = None
        if need_grad:
            ck = torch.empty(lib.gbw_acoustic3d_checkpoint_bytes(ctypes.byref(prob)),
                             dtype=torch.uint8, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        rec = [torch.empty((S, nt, R), **f32) for _ in range(4)]
        eng = [torch.empty((nz, ny, nx), **f32) for _ in range(4)]
        outs = _native.Outputs3()
        outs.rec_p, outs.rec_vx, outs.rec_vy, outs.rec_vz = (t.data_ptr() for t in rec)
        outs.energy_p, outs.energy_vx, outs.energy_vy, outs.energy_vz = (t.data_ptr() for t in eng)
        outs.checkpoints = _ptr(ck)
        ins = _inputs3(spec, src_v, damp, vp, rho)
        rc = lib.gbw_acoustic3d_forward(ctypes.byref(prob), ctypes.byref(ins), ctypes.byref(outs),
                                        ws.data_ptr(), ws_bytes, _stream_handle(dev))
        _native.check(rc)
        ctx.mark_non_differentiable(*eng)
        ctx.spec = spec
        ctx.ck = ck
        ctx.save_for_backward(vp, rho, src_v, damp)
        return (*rec, *eng)

    @staticmethod
    def backward(ctx, g_p, g_vx, g_vy, g_vz, *unused):
        lib = _native.load()
        vp, rho, src_v, damp = ctx.saved_tensors
        spec = ctx.spec
        prob = spec.prob
        dev = vp.device

        def prep(g):
            return None if g is None else g.to(torch.float32).contiguous()

        gs = [prep(g) for g in (g_p, g_vx, g_vy, g_vz)]
        want_vp, want_rho, want_w = ctx.needs_input_grad[:3]
        grad_vp = torch.empty_like(vp) if want_vp else None
        grad_rho = torch.empty_like(rho) if want_rho else None
        grad_w = torch.empty_like(src_v) if want_w else None
        adj = _native.Adjoint3()
        adj.g_p, adj.g_vx, adj.g_vy, adj.g_vz = (_ptr(g) for g in gs)
        adj.checkpoints = _ptr(ctx.ck)
        adj.grad_vp, adj.grad_rho, adj.grad_src_v = _ptr(grad_vp), _ptr(grad_rho), _ptr(grad_w)
        ws_bytes = lib.gbw_acoustic3d_workspace_bytes(ctypes.byref(prob), 2 if want_rho else 1)
        if ws_bytes == 0:
            _native.check(-1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        ins = _inputs3(spec, src_v, damp, vp, rho)
        rc = lib.gbw_acoustic3d_backward(ctypes.byref(prob), ctypes.byref(ins), ctypes.byref(adj),
                                         ws.data_ptr(), ws_bytes, _stream_handle(dev))
        _native.check(rc)
        ctx.ck = None
        return grad_vp, grad_rho, grad_w, None, None


def acoustic_forward_kernel_3d(
    nx: int,
    ny: int,
    nz: int,
    dx: float,
    dy: float,
    dz: float,
    nt: int,
    dt: float,
    nabc: int,
    free_surface: bool,
    src_x: Tensor,
    src_y: Tensor,
    src_z: Tensor,
    src_n: int,
    src_v: Tensor,
    rcv_x: Tensor,
    rcv_y: Tensor,
    rcv_z: Tensor,
    rcv_n: int,
    damp: Tensor,
    vp: Tensor,
    rho: Tensor,
    checkpoint_segments: int = 1,
    device: torch.device = torch.device('cpu'),
    dtype: torch.dtype = torch.float32,
) -> Dict[str, Tensor]:
    """3D forward simulation (signature of ``acoustic_3d.py:275-301``).

    Returns ``'p', 'vx', 'vy', 'vz'`` ``(src_n, nt, rcv_n)`` and the energy
    maps ``'forward_wavefield_p/vx/vy/vz'`` ``(nz, ny, nx)`` (detached), summed
    per ``torch.chunk`` chunk of ``checkpoint_segments`` as the reference sums
    them (``acoustic_3d.py:415-476``).  All ``nt`` steps run in the HIP library
    (``gbw_acoustic3d_forward``).  Gradients flow to ``vp``, ``rho`` and
    ``src_v`` when they require grad: ``gbw_acoustic3d_backward``, the
    hand-written adjoint of the step with forward checkpoints every
    ``ceil(nt / checkpoint_segments)`` steps (or an interval balancing
    checkpoint and recompute memory when ``checkpoint_segments == 1``).
    """
    if dtype != torch.float32:
        raise NotImplementedError("geobrain_amd computes in float32 only (the reference default)")
    dev = vp.device
    for name, t in (("vp", vp), ("rho", rho), ("src_v", src_v), ("damp", damp)):
        _on_gpu(t, name, dev)
    if not torch.cuda.is_available():
        raise RuntimeError("geobrain_amd: no ROCm GPU is visible")
    sx, sy, sz = (_host_index(a, n, True) for a, n in ((src_x, "src_x"), (src_y, "src_y"),
                                                       (src_z, "src_z")))
    rx, ry, rz = (_host_index(a, n) for a, n in ((rcv_x, "rcv_x"), (rcv_y, "rcv_y"),
                                                 (rcv_z, "rcv_z")))
    if not (sx.shape == sy.shape == sz.shape) or sx.shape[0] != src_n:
        raise ValueError(f"src_x/src_y/src_z must each have src_n={src_n} entries")
    if not (rx.shape == ry.shape == rz.shape) or rx.shape[0] != rcv_n:
        raise ValueError(f"rcv_x/rcv_y/rcv_z must each have rcv_n={rcv_n} entries")
    if tuple(vp.shape) != (nz, ny, nx) or tuple(rho.shape) != (nz, ny, nx):
        raise ValueError(f"vp and rho must have shape ({nz}, {ny}, {nx})")
    padded = (nz + 2 * nabc, ny + 2 * nabc, nx + 2 * nabc)
    if tuple(damp.shape) != padded:
        raise ValueError(f"damp must have shape {padded}")
    # (src_n, K) positions: K point sources per shot (acoustic_3d.py:171-176)
    points = sx.shape[1] if sx.ndim == 2 else 1
    if sx.ndim == 2:
        if src_v.dim() == 2 and tuple(src_v.shape) == (src_n, nt):
            shared = 0
        elif src_v.dim() == 3 and tuple(src_v.shape) == (src_n, points, nt):
            shared = 2
        else:
            raise ValueError(f"with (src_n, K) source positions src_v must have shape "
                             f"({src_n}, {nt}) or ({src_n}, {points}, {nt})")
    elif src_v.dim() == 1:
        shared = 1
        if src_v.shape[0] != nt:
            raise ValueError(f"src_v has {src_v.shape[0]} samples, nt={nt}")
    elif src_v.dim() == 2:
        shared = 0
        if tuple(src_v.shape) != (src_n, nt):
            raise ValueError(f"src_v must have shape ({src_n}, {nt}) or ({nt},)")
    else:
        raise NotImplementedError("src_v must be (src_n, nt) or (nt,)")
    prob = _native.Problem3()
    prob.nx, prob.ny, prob.nz = int(nx), int(ny), int(nz)
    prob.dx, prob.dy, prob.dz = float(dx), float(dy), float(dz)
    prob.nt, prob.dt = int(nt), float(dt)
    prob.nabc, prob.free_surface = int(nabc), int(bool(free_surface))
    prob.src_n, prob.rcv_n = int(src_n), int(rcv_n)
    prob.wavelet_shared = shared
    prob.src_points = int(points)
    prob.energy_chunk = chunk_length(nt, checkpoint_segments)
    prob.shots_per_block = _launch3.shots_per_block
    prob.step_kernel = _launch3.step_kernel
    prob.z_chunk = _launch3.z_chunk
    if checkpoint_segments > 1:
        prob.checkpoint_every = chunk_length(nt, checkpoint_segments)
    else:
        prob.checkpoint_every = auto_checkpoint_interval_3d(nt, bool(rho.requires_grad))
    spec = _Spec3(prob=prob, idx=(sx, sy, sz, rx, ry, rz))
    f32 = torch.float32
    res = Acoustic3DFunction.apply(vp.to(f32).contiguous(), rho.to(f32).contiguous(),
                                   src_v.to(f32).contiguous(), damp.to(f32).contiguous(), spec)
    keys = ('p', 'vx', 'vy', 'vz', 'forward_wavefield_p', 'forward_wavefield_vx',
            'forward_wavefield_vy', 'forward_wavefield_vz')
    return dict(zip(keys, res))
