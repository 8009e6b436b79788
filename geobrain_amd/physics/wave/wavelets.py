"""Source wavelets, mirroring ``geobrain/physics/wave/wavelets/ricker.py:19-109``."""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch
import torch.nn as nn
from torch import Tensor


class WaveletGenerator(nn.Module):
    """Base class: ``forward(f0, dt, device) -> (wavelet, time_axis)``."""

    def forward(self, f0: float, dt: float, device: Optional[str] = None) -> Tuple[Tensor, Tensor]:
        raise NotImplementedError

    def __call__(self, f0: float, dt: float, device: Optional[str] = None):
        return self.forward(f0, dt, device)

    @property
    def name(self) -> str:
        return getattr(self, "_name", self.__class__.__name__.lower())


class RickerWavelet(WaveletGenerator):
    """``w(t) = (1 - 2 (pi f0 t)^2) exp(-(pi f0 t)^2)`` on a centred, odd-length
    float32 time axis of ``2 int(length_factor / f0 / dt / 2) + 1`` samples."""

    def __init__(self, length_factor: float = 2.2):
        super().__init__()
        self.length_factor = length_factor
        self._name = 'ricker'

    def forward(self, f0: float, dt: float, device: Optional[str] = None):
        if isinstance(f0, Tensor):
            f0 = f0.item()
        if isinstance(dt, Tensor):
            dt = dt.item()
        device = device or 'cpu'
        nw = self.get_length(f0, dt)
        nc = nw // 2
        t = (torch.arange(nw, device=device, dtype=torch.float32) - nc) * dt
        arg = np.pi * f0 * t
        return (1.0 - 2.0 * arg ** 2) * torch.exp(-arg ** 2), t

    def get_length(self, f0: float, dt: float) -> int:
        nw = self.length_factor / f0 / dt
        return 2 * int(nw / 2) + 1
