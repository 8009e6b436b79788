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
