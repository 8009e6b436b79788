"""Full-waveform-inversion loop with the semantics of GeoBrain's acoustic FWI example.

The loop of ``examples/09_acoustic_fwi.py:259-338``, as a reusable driver:

* misfit ``trace_l2_misfit`` (``:290-319``): every trace is divided by its
  max-abs over time (clamped at 1e-30), the residual's L2 norm over time is
  taken per (shot, receiver), traces silent in both observed and predicted
  data are dropped, the norms are summed;
* ``loss.backward()`` through the HIP adjoint, then the gradient mask
  (``:325-326``), ``Adam.step`` and ``StepLR.step`` (``:279-282``, ``:328-329``);
* ``model.forward()`` (bound clipping + Gardner density refresh, ``:332``).

The propagator's PML is built from the initial model's vmax, once, as in the
example (``propagators/acoustic.py:128-129``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import torch


def normalize_traces(data: torch.Tensor) -> torch.Tensor:
    """Divide each trace by its max |amplitude| over time (dim -2), floor 1e-30."""
    max_vals = torch.max(torch.abs(data), dim=-2, keepdim=True).values
    return data / max_vals.clamp(min=1e-30)


def trace_l2_misfit(d_pred: torch.Tensor, d_obs: torch.Tensor) -> torch.Tensor:
    """Sum over active traces of ``||norm(d_obs) - norm(d_pred)||_2`` (over time).

    The example takes the square root of every trace's residual energy and then
    drops the traces silent in both data sets (``examples/09_acoustic_fwi.py:
    305-318``).  The square root's derivative at a silent trace is 0/0, so the
    example's gradient is NaN everywhere as soon as one trace is silent (far
    offsets of long lines), and likewise when a kept trace's residual is exactly
    zero (e.g. two far traces whose only nonzero sample is the same last one:
    both normalise to the same spike).  Here the root is taken of the kept
    traces only, and a zero residual contributes a zero gradient (the
    subgradient of the norm at 0): the loss is the same sum of the same values
    in the same order, and wherever the example's gradient is finite this one
    is identical."""
    residual = normalize_traces(d_obs) - normalize_traces(d_pred)
    energy = torch.sum(residual ** 2, dim=-2)
    active = ~((torch.sum(torch.abs(d_obs), dim=-2) == 0) &
               (torch.sum(torch.abs(d_pred), dim=-2) == 0))
    kept = energy[active]
    nonzero = kept > 0
    norms = torch.where(nonzero, torch.sqrt(torch.where(nonzero, kept, torch.ones_like(kept))),
                        torch.zeros_like(kept))
    return torch.sum(norms)


@dataclass
class FWI:
    """Adam-based FWI of ``vp`` for an ``AcousticPropagator`` and its model."""
    propagator: object
    model: object
    d_obs: torch.Tensor
    lr: float = 10.0
    step_size: int = 200
    gamma: float = 0.75
    grad_mask: Optional[torch.Tensor] = None
    checkpoint_segments: int = 10
    history: List[float] = field(default_factory=list)

    def __post_init__(self):
        self.optimizer = torch.optim.Adam(self.model.parameters(), lr=self.lr)
        self.scheduler = torch.optim.lr_scheduler.StepLR(self.optimizer, step_size=self.step_size,
                                                         gamma=self.gamma)

    def step(self) -> float:
        self.optimizer.zero_grad()
        result = self.propagator.forward(checkpoint_segments=self.checkpoint_segments)
        loss = trace_l2_misfit(result['p'], self.d_obs)
        loss.backward()
        if self.grad_mask is not None and self.model.vp.grad is not None:
            self.model.vp.grad.data *= self.grad_mask
        self.optimizer.step()
        self.scheduler.step()
        self.model.forward()
        value = loss.item()
        self.history.append(value)
        return value

    def run(self, n_iter: int, callback=None) -> List[float]:
        for it in range(n_iter):
            value = self.step()
            if callback is not None:
                callback(it, value)
        return self.history
