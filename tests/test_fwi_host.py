"""The FWI driver's misfit follows examples/09_acoustic_fwi.py:290-319 (CPU)."""

import numpy as np
import torch

from geobrain_amd.fwi import normalize_traces, trace_l2_misfit


def test_normalize_traces_and_misfit():
    rng = np.random.default_rng(0)
    obs = rng.standard_normal((3, 50, 7)).astype(np.float32)
    pred = rng.standard_normal((3, 50, 7)).astype(np.float32)
    obs[1, :, 2] = 0.0
    pred[1, :, 2] = 0.0          # silent in both: dropped
    pred[2, :, 4] = 0.0          # silent in prediction only: kept
    no = obs / np.maximum(np.abs(obs).max(axis=1, keepdims=True), 1e-30)
    npd = pred / np.maximum(np.abs(pred).max(axis=1, keepdims=True), 1e-30)
    per = np.sqrt(((no - npd) ** 2).sum(axis=1))
    keep = ~((np.abs(obs).sum(axis=1) == 0) & (np.abs(pred).sum(axis=1) == 0))
    want = per[keep].sum()
    got = trace_l2_misfit(torch.tensor(pred), torch.tensor(obs)).item()
    assert abs(got - want) <= 1e-5 * want
    n = normalize_traces(torch.tensor(obs)).numpy()
    assert np.allclose(np.abs(n).max(axis=1)[keep[:, :]], 1.0)


def test_misfit_gradient_finite_with_silent_traces():
    """A trace silent in both data sets makes the example's form (root of every
    trace, then drop the silent ones) differentiate sqrt at 0: NaN everywhere.
    The driver's misfit has the same value and, on traces where the example's
    gradient is finite, the same gradient; with a silent trace it stays finite."""
    rng = np.random.default_rng(1)
    obs = torch.tensor(rng.standard_normal((2, 40, 5)).astype(np.float32))
    pred = torch.tensor(rng.standard_normal((2, 40, 5)).astype(np.float32))

    def example_form(p, o):
        r = normalize_traces(o) - normalize_traces(p)
        per = torch.sqrt(torch.sum(r ** 2, dim=-2))
        act = ~((torch.sum(torch.abs(o), dim=-2) == 0) & (torch.sum(torch.abs(p), dim=-2) == 0))
        return torch.sum(per[act])

    # no silent trace: same value and gradient, bit for bit
    a = pred.clone().requires_grad_(True)
    b = pred.clone().requires_grad_(True)
    la, lb = trace_l2_misfit(a, obs), example_form(b, obs)
    la.backward()
    lb.backward()
    assert la.item() == lb.item()
    assert torch.equal(a.grad, b.grad)
    # one silent trace: the example's gradient is NaN, the driver's is finite
    obs[0, :, 3] = 0.0
    pred[0, :, 3] = 0.0
    a = pred.clone().requires_grad_(True)
    b = pred.clone().requires_grad_(True)
    la, lb = trace_l2_misfit(a, obs), example_form(b, obs)
    la.backward()
    lb.backward()
    assert la.item() == lb.item()
    assert torch.isnan(b.grad).any()
    assert torch.isfinite(a.grad).all()
    # the silent trace gets zero gradient (it is not in the loss)
    assert torch.count_nonzero(a.grad[0, :, 3]) == 0
    # a kept trace with a zero residual (same single spike in both): the
    # example's gradient is NaN, the driver's is finite, the loss the same
    obs[1, :, 1] = 0.0
    pred[1, :, 1] = 0.0
    obs[1, -1, 1] = 2e-30
    pred[1, -1, 1] = 5e-30
    a = pred.clone().requires_grad_(True)
    b = pred.clone().requires_grad_(True)
    la, lb = trace_l2_misfit(a, obs), example_form(b, obs)
    la.backward()
    lb.backward()
    assert la.item() == lb.item()
    assert torch.isnan(b.grad).any()
    assert torch.isfinite(a.grad).all()
