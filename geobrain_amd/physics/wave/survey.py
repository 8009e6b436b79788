"""Acquisition geometry, mirroring ``geobrain/physics/wave/survey/{source,receiver,survey}.py``.

Positions are grid indices.  ``get_loc()`` returns columns ``(x, z)`` in 2-D,
as the reference's ``Source.get_loc`` (``source.py:199-212``) and
``Receiver.get_loc`` (``receiver.py:165-178``) do.
"""

from __future__ import annotations

from typing import List, Optional

import numpy as np


class Source:
    """Shots: one point source and one wavelet of length ``nt`` each."""

    def __init__(self, nt: int, dt: float, f0: float) -> None:
        self.nt = nt
        self.dt = dt
        self.f0 = f0
        self.t = np.arange(nt) * dt
        self.loc_x: List[int] = []
        self.loc_z: List[int] = []
        self.loc_y: List[int] = []
        self.type: List[str] = []
        self.wavelet: List[np.ndarray] = []
        self.moment_tensor: List[np.ndarray] = []
        self.num: int = 0

    def __repr__(self) -> str:
        if self.num == 0:
            return "Source: empty"
        x, z = np.array(self.loc_x), np.array(self.loc_z)
        return (f"Source:\n  Number: {self.num}\n  Time: {self.nt} samples @ {self.dt * 1000:.2f} ms\n"
                f"  X range: {x.min()} - {x.max()} (grid)\n  Z range: {z.min()} - {z.max()} (grid)\n"
                f"  Types: {list(set(self.type))}")

    def add_source(self, x: int, z: int, wavelet: np.ndarray, y: Optional[int] = None,
                   source_type: str = 'mt', moment_tensor: Optional[np.ndarray] = None) -> None:
        if wavelet.shape[0] != self.nt:
            raise ValueError(f"Wavelet length {wavelet.shape[0]} != nt {self.nt}")
        mt = np.eye(3) if moment_tensor is None else moment_tensor
        if mt.shape != (3, 3):
            raise ValueError(f"Moment tensor must be 3x3, got {mt.shape}")
        self.loc_x.append(x)
        self.loc_z.append(z)
        if y is not None:
            self.loc_y.append(y)
        self.type.append(source_type)
        self.wavelet.append(wavelet.copy())
        self.moment_tensor.append(mt.copy())
        self.num += 1

    def add_sources(self, x: np.ndarray, z: np.ndarray, wavelet: np.ndarray,
                    y: Optional[np.ndarray] = None, source_type: str = 'mt',
                    moment_tensor: Optional[np.ndarray] = None) -> None:
        if x.shape != z.shape:
            raise ValueError(f"x and z must have same shape: {x.shape} != {z.shape}")
        if y is not None and y.shape != x.shape:
            raise ValueError(f"y must have same shape as x: {y.shape} != {x.shape}")
        if wavelet.shape[0] != self.nt:
            raise ValueError(f"Wavelet length {wavelet.shape[0]} != nt {self.nt}")
        mt = np.eye(3) if moment_tensor is None else moment_tensor
        n = len(x.flatten())
        self.loc_x.extend(x.flatten().tolist())
        self.loc_z.extend(z.flatten().tolist())
        if y is not None:
            self.loc_y.extend(y.flatten().tolist())
        self.type.extend([source_type] * n)
        self.wavelet.extend([wavelet.copy() for _ in range(n)])
        self.moment_tensor.extend([mt.copy() for _ in range(n)])
        self.num += n

    def get_loc(self) -> np.ndarray:
        cols = [np.array(self.loc_x).reshape(-1, 1)]
        if self.loc_y:
            cols.append(np.array(self.loc_y).reshape(-1, 1))
        cols.append(np.array(self.loc_z).reshape(-1, 1))
        return np.hstack(cols)

    def get_wavelet(self) -> np.ndarray:
        return np.array(self.wavelet)

    def get_moment_tensor(self) -> np.ndarray:
        return np.array(self.moment_tensor)


class Receiver:
    """Receivers shared by every shot."""

    VALID_TYPES = ('pr', 'vx', 'vy', 'vz')

    def __init__(self, nt: int, dt: float) -> None:
        self.nt = nt
        self.dt = dt
        self.loc_x: List[int] = []
        self.loc_z: List[int] = []
        self.loc_y: List[int] = []
        self.type: List[str] = []
        self.num: int = 0

    def __repr__(self) -> str:
        if self.num == 0:
            return "Receiver: empty"
        x, z = np.array(self.loc_x), np.array(self.loc_z)
        return (f"Receiver:\n  Number: {self.num}\n  Recording: {self.nt} samples @ "
                f"{self.dt * 1000:.2f} ms\n  X range: {x.min()} - {x.max()} (grid)\n"
                f"  Z range: {z.min()} - {z.max()} (grid)\n  Types: {list(set(self.type))}")

    def _check_type(self, rcv_type: str) -> str:
        if rcv_type.lower() not in self.VALID_TYPES:
            raise ValueError(f"Receiver type must be one of {self.VALID_TYPES}, got '{rcv_type}'")
        return rcv_type.lower()

    def add_receiver(self, x: int, z: int, rcv_type: str, y: Optional[int] = None) -> None:
        t = self._check_type(rcv_type)
        self.loc_x.append(x)
        self.loc_z.append(z)
        if y is not None:
            self.loc_y.append(y)
        self.type.append(t)
        self.num += 1

    def add_receivers(self, x: np.ndarray, z: np.ndarray, rcv_type: str,
                      y: Optional[np.ndarray] = None) -> None:
        if x.shape != z.shape:
            raise ValueError(f"x and z must have same shape: {x.shape} != {z.shape}")
        if y is not None and y.shape != x.shape:
            raise ValueError(f"y must have same shape as x: {y.shape} != {x.shape}")
        t = self._check_type(rcv_type)
        n = len(x.flatten())
        self.loc_x.extend(x.flatten().tolist())
        self.loc_z.extend(z.flatten().tolist())
        if y is not None:
            self.loc_y.extend(y.flatten().tolist())
        self.type.extend([t] * n)
        self.num += n

    def get_loc(self) -> np.ndarray:
        cols = [np.array(self.loc_x).reshape(-1, 1)]
        if self.loc_y:
            cols.append(np.array(self.loc_y).reshape(-1, 1))
        cols.append(np.array(self.loc_z).reshape(-1, 1))
        return np.hstack(cols)

    def get_type(self, unique: bool = False) -> np.ndarray:
        types = np.array(self.type)
        return np.unique(types) if unique else types


class Survey:
    """Source + receivers; optional per-shot receiver masks are stored as the
    reference stores them (``survey.py:58-106``) -- and, as there, not applied."""

    def __init__(self, source: Source, receiver: Receiver,
                 receiver_masks: Optional[np.ndarray] = None) -> None:
        self.source = source
        self.receiver = receiver
        self.receiver_masks = None
        if receiver_masks is not None:
            self.set_receiver_masks(receiver_masks)

    def __repr__(self) -> str:
        return (f"Survey:\n  Shots: {self.n_shots}\n  Receivers: {self.n_receivers}\n"
                f"  Time: {self.nt} samples @ {self.dt * 1000:.2f} ms\n"
                f"  Receiver masks: {'Yes' if self.receiver_masks is not None else 'No'}")

    def set_receiver_masks(self, receiver_masks: np.ndarray) -> None:
        expected = (self.n_shots, self.n_receivers)
        if receiver_masks.shape != expected:
            raise ValueError(f"Receiver mask shape {receiver_masks.shape} doesn't match "
                             f"survey geometry {expected}")
        self.receiver_masks = receiver_masks.astype(bool)

    @property
    def n_shots(self) -> int:
        return self.source.num

    @property
    def n_receivers(self) -> int:
        return self.receiver.num

    @property
    def nt(self) -> int:
        return self.source.nt

    @property
    def dt(self) -> float:
        return self.source.dt

    @property
    def f0(self) -> float:
        return self.source.f0
