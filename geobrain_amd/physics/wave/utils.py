"""Tensor conversion helpers, mirroring ``geobrain/physics/wave/utils/tensor_ops.py:17-58``."""

from __future__ import annotations

import numpy as np
import torch


def numpy2tensor(a, dtype: torch.dtype = torch.float32) -> torch.Tensor:
    if torch.is_tensor(a):
        return a.to(dtype)
    return torch.tensor(a, requires_grad=False, dtype=dtype)


def tensor2numpy(a) -> np.ndarray:
    if not torch.is_tensor(a):
        return np.asarray(a)
    return a.detach().cpu().numpy()
