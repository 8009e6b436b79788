"""Grid and boundary configuration, mirroring ``geobrain/physics/wave/core/config.py:17-164``."""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal, Optional, Tuple

import numpy as np


@dataclass
class GridConfig:
    """Regular grid; ``dx == dz`` is required (``core/config.py:71-88``)."""
    nx: int
    nz: int
    dx: float
    dz: float
    ox: float = 0.0
    oz: float = 0.0
    ny: Optional[int] = None
    dy: Optional[float] = None
    oy: float = 0.0

    def __post_init__(self):
        if self.nx <= 0 or self.nz <= 0:
            raise ValueError(f"Grid dimensions must be positive: nx={self.nx}, nz={self.nz}")
        if self.dx <= 0 or self.dz <= 0:
            raise ValueError(f"Grid spacing must be positive: dx={self.dx}, dz={self.dz}")
        if self.dx != self.dz:
            raise ValueError(f"Grid spacing must be equal: dx={self.dx}, dz={self.dz}")
        if (self.ny is None) != (self.dy is None):
            raise ValueError("ny and dy must both be provided for 3-D grids")
        if self.ny is not None:
            if self.ny <= 0:
                raise ValueError(f"ny must be positive: {self.ny}")
            if self.dy <= 0:
                raise ValueError(f"dy must be positive: {self.dy}")
            if self.dy != self.dx:
                raise ValueError(f"Grid spacing must be equal: dx={self.dx}, dy={self.dy}")

    @property
    def ndim(self) -> int:
        return 3 if self.ny is not None else 2

    @property
    def shape(self) -> tuple:
        return (self.nz, self.ny, self.nx) if self.ndim == 3 else (self.nz, self.nx)

    @property
    def x(self) -> np.ndarray:
        return np.arange(self.nx) * self.dx + self.ox

    @property
    def z(self) -> np.ndarray:
        return np.arange(self.nz) * self.dz + self.oz

    @property
    def y(self) -> Optional[np.ndarray]:
        return None if self.ny is None else np.arange(self.ny) * self.dy + self.oy

    @property
    def extent(self) -> Tuple[float, float, float, float]:
        return (self.ox, self.ox + (self.nx - 1) * self.dx,
                self.oz + (self.nz - 1) * self.dz, self.oz)


@dataclass
class BoundaryConfig:
    """Absorbing-boundary settings; defaults as ``core/config.py:151-154``."""
    type: Literal['pml', 'gerjan', 'sincos'] = 'pml'
    n_layers: int = 20
    free_surface: bool = False
    alpha: float = 0.0053

    def __post_init__(self):
        self.type = self.type.lower()
        if self.type not in ('pml', 'gerjan', 'sincos'):
            raise ValueError(f"Unknown boundary type: '{self.type}'. "
                             f"Choose from: 'pml', 'gerjan', 'sincos'")
        if self.n_layers <= 0:
            raise ValueError(f"Number of layers must be positive: {self.n_layers}")
