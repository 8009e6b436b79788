"""Propagator front-end, mirroring ``geobrain/physics/wave/core/base.py:296-357``
(``AbstractPropagator``) and ``geobrain/physics/wave/propagators/acoustic.py:23-301``
(``AcousticPropagator``, 2-D and 3-D).

Behaviour kept from the reference, on purpose:
* the PML ``vmax`` is read from the model once, at construction
  (``acoustic.py:128-129``), and the PML is built with ``free_surface=False``
  whatever the model says (``acoustic.py:140``);
* ``forward()`` first calls ``model.forward()`` (Gardner refresh + clipping);
* checkpointing is switched off when no model parameter requires grad
  (``acoustic.py:199-202``);
* ``survey.receiver_masks`` is stored but not applied (``acoustic.py:119``);
* for a 3-D model, locations are ``(x, y, z)`` columns (``acoustic.py:93-116``)
  and ``snapshot_every`` is not passed to the 3-D kernel (``acoustic.py:222-247``).
"""

from __future__ import annotations

from typing import Optional

import torch
import torch.nn as nn

from .boundary import create_boundary
from .kernels import _host_index, acoustic_forward_kernel, acoustic_forward_kernel_3d
from .models import AbstractModel, AcousticModel
from .result import WaveResult
from .survey import Survey
from .utils import numpy2tensor


class AbstractPropagator(nn.Module):
    def __init__(self, model: AbstractModel, device: str = 'cpu',
                 dtype: torch.dtype = torch.float32) -> None:
        super().__init__()
        self.model = model
        self.device = device
        self.dtype = dtype
        self.ox, self.oz = model.ox, model.oz
        self.dx, self.dz = model.dx, model.dz
        self.nx, self.nz = model.nx, model.nz
        self.ny, self.dy, self.oy = model.ny, model.dy, model.oy
        self.abc_type = model.abc_type
        self.nabc = model.nabc
        self.free_surface = model.free_surface

    def forward(self, **kwargs):
        raise NotImplementedError

    def __repr__(self) -> str:
        return (f"{self.__class__.__name__}(\n  model={self.model.__class__.__name__},\n"
                f"  grid=({self.nx}, {self.nz}),\n  device='{self.device}'\n)")


class AcousticPropagator(AbstractPropagator):
    """2-D / 3-D acoustic propagator on the MI355X HIP library.

    ``AcousticPropagator(model, survey, device, dtype)``;
    ``forward(model=None, shot_index=None, checkpoint_segments=1, snapshot_every=0)``
    returns a ``WaveResult`` whose ``p``/``vx``/``vz`` are ``(n_shots, nt, n_rcv)``.

    ``checkpoint_dtype`` (attribute, default ``torch.float32``) selects the
    storage of the gradient's forward checkpoints (``torch.bfloat16`` halves it;
    2-D only: a 3-D model with another dtype raises ``NotImplementedError``).
    """

    checkpoint_dtype: torch.dtype = torch.float32

    def __init__(self, model: AcousticModel, survey: Survey, device: str = 'cpu',
                 dtype: torch.dtype = torch.float32) -> None:
        super().__init__(model, device, dtype)
        if not isinstance(model, AcousticModel):
            raise TypeError(f"Expected AcousticModel, got {type(model).__name__}")
        if not isinstance(survey, Survey):
            raise TypeError(f"Expected Survey, got {type(survey).__name__}")
        self.survey = survey
        self.nt = survey.nt
        self.dt = survey.dt
        self.f0 = survey.f0
        self.damp = None
        self._init_boundary()
        self.source = survey.source
        src_loc = self.source.get_loc()
        self.src_n = self.source.num
        self.wavelet = numpy2tensor(self.source.get_wavelet(), dtype).to(device)
        self.receiver = survey.receiver
        rcv_loc = self.receiver.get_loc()
        self.rcv_n = self.receiver.num
        zc = 2 if model.ndim == 3 else 1  # columns: (x, z) in 2-D, (x, y, z) in 3-D
        self.src_x = numpy2tensor(src_loc[:, 0], torch.long).to(device)
        self.src_z = numpy2tensor(src_loc[:, zc], torch.long).to(device)
        self.rcv_x = numpy2tensor(rcv_loc[:, 0], torch.long).to(device)
        self.rcv_z = numpy2tensor(rcv_loc[:, zc], torch.long).to(device)
        if model.ndim == 3:
            self.src_y = numpy2tensor(src_loc[:, 1], torch.long).to(device)
            self.rcv_y = numpy2tensor(rcv_loc[:, 1], torch.long).to(device)
        self.receiver_masks = survey.receiver_masks

    def _positions(self, *names: str):
        """Host int64 copies of the position attributes, read from the current
        ``src_x`` / ``rcv_x`` / ... tensors on every call (a caller may move
        shots between calls, as the reference's forward re-reads them); the
        copies are cached per tensor and version, so an unchanged device
        tensor is copied to the host once."""
        return [_host_index(getattr(self, n), n) for n in names]

    def _init_boundary(self, vmax: Optional[float] = None) -> None:
        if vmax is None:
            vmax = float(self.model.vp.detach().cpu().max())
        kw = dict(ny=self.ny, dy=self.dy) if self.model.ndim == 3 else {}
        damp = create_boundary(bc_type=self.abc_type, nx=self.nx, nz=self.nz, dx=self.dx,
                               dz=self.dz, n_layers=self.nabc, vmax=vmax,
                               alpha=self.model.abc_alpha, free_surface=False, **kw)
        self.damp = numpy2tensor(damp, self.dtype).to(self.device)

    def forward(self, model: Optional[AcousticModel] = None, shot_index: Optional[int] = None,
                checkpoint_segments: int = 1, snapshot_every: int = 0) -> WaveResult:
        if self.model.ndim == 3 and self.checkpoint_dtype != torch.float32:
            raise NotImplementedError("3-D gradients store float32 checkpoints only")
        model = model if model is not None else self.model
        model.forward()
        if checkpoint_segments > 1 and not any(p.requires_grad for p in model.parameters()):
            checkpoint_segments = 1
        sl = slice(shot_index, shot_index + 1) if shot_index is not None else slice(None)
        sx, sz, rx, rz = self._positions("src_x", "src_z", "rcv_x", "rcv_z")
        src_x, src_z = sx[sl], sz[sl]
        wavelet = self.wavelet[sl]
        src_n = 1 if shot_index is not None else self.src_n
        if self.model.ndim == 3:
            sy, ry = self._positions("src_y", "rcv_y")
            result = acoustic_forward_kernel_3d(
                nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy, dz=self.dz,
                nt=self.nt, dt=self.dt, nabc=self.nabc, free_surface=self.free_surface,
                src_x=src_x, src_y=sy[sl], src_z=src_z, src_n=src_n,
                src_v=wavelet, rcv_x=rx, rcv_y=ry, rcv_z=rz,
                rcv_n=self.rcv_n, damp=self.damp, vp=model.vp, rho=model.rho,
                checkpoint_segments=checkpoint_segments, device=torch.device(self.device),
                dtype=self.dtype)
            return WaveResult.from_dict(result)
        result = acoustic_forward_kernel(
            nx=self.nx, nz=self.nz, dx=self.dx, dz=self.dz, nt=self.nt, dt=self.dt,
            nabc=self.nabc, free_surface=self.free_surface,
            src_x=src_x, src_z=src_z, src_n=src_n, src_v=wavelet,
            rcv_x=rx, rcv_z=rz, rcv_n=self.rcv_n,
            damp=self.damp, vp=model.vp, rho=model.rho,
            checkpoint_segments=checkpoint_segments, device=torch.device(self.device),
            dtype=self.dtype, snapshot_every=snapshot_every,
            checkpoint_dtype=self.checkpoint_dtype)
        return WaveResult.from_dict(result)

    def __call__(self, model: Optional[AcousticModel] = None, shot_index: Optional[int] = None,
                 checkpoint_segments: int = 1, snapshot_every: int = 0) -> WaveResult:
        return self.forward(model=model, shot_index=shot_index,
                            checkpoint_segments=checkpoint_segments,
                            snapshot_every=snapshot_every)
