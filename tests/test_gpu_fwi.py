"""End-to-end known answer: the Marmousi2 FWI of GeoBrain's example 09, run on the
HIP propagator, against the loss values its notebook recorded on a CUDA GPU
(examples/notebooks/09_acoustic_fwi.ipynb:420-421).  The iteration-0 loss is
forward modelling only, so it pins the whole forward path (geometry, PML,
Gardner density, integrated Ricker, free surface, trace-normalised misfit);
iteration 10 follows ten Adam steps on the hand-written gradient."""

import json

import numpy as np
import pytest
import torch

from geobrain_amd.fwi import FWI
from geobrain_amd.physics import wave

pytestmark = pytest.mark.gpu


def _marmousi_setup():
    from scipy.integrate import cumulative_trapezoid
    from scipy.ndimage import gaussian_filter
    z = np.load("tests/golden/marmousi2_fwi.npz")
    meta = json.loads(str(z["meta"]))
    vp_true = z["vp_true"]
    NX, NZ, DX, NT, DT, F0 = 200, 88, 40.0, 3000, 0.003, 5.0
    rho_true = 310.0 * np.power(vp_true, 0.25)
    grid = wave.GridConfig(nx=NX, nz=NZ, dx=DX, dz=DX)
    boundary = wave.BoundaryConfig(type="pml", n_layers=30, free_surface=True, alpha=0.007)
    src_x = np.arange(2, NX - 1, 5)
    w, _ = wave.RickerWavelet()(f0=F0, dt=DT)
    w = np.pad(w.numpy(), (0, NT - len(w)))
    w = cumulative_trapezoid(w, initial=0)
    source = wave.Source(nt=NT, dt=DT, f0=F0)
    for x in src_x:
        source.add_source(x=int(x), z=1, wavelet=w, source_type="mt", moment_tensor=np.eye(3))
    receiver = wave.Receiver(nt=NT, dt=DT)
    receiver.add_receivers(x=np.arange(NX), z=np.ones(NX, dtype=int), rcv_type="pr")
    survey = wave.Survey(source=source, receiver=receiver)
    vp_init = vp_true.copy()
    vp_init[12:, :] = gaussian_filter(vp_true[12:, :], sigma=[6, 6], mode="reflect")
    mask = np.ones((NZ, NX), np.float32)
    mask[:12, :] = 0.0
    return meta, grid, boundary, survey, vp_true, rho_true, vp_init, mask


def test_marmousi2_fwi_matches_notebook_losses():
    print("[case] marmousi2 fwi", flush=True)
    meta, grid, boundary, survey, vp_true, rho_true, vp_init, mask = _marmousi_setup()
    model_true = wave.AcousticModel(grid=grid, boundary=boundary, vp=vp_true, rho=rho_true,
                                    vp_grad=False, device="cuda")
    prop = wave.AcousticPropagator(model=model_true, survey=survey, device="cuda")
    with torch.no_grad():
        d_obs = prop.forward(checkpoint_segments=10)["p"].detach().clone()
    assert tuple(d_obs.shape) == (40, 3000, 200)
    rho = model_true.rho.detach().cpu().numpy()
    assert round(float(rho.min())) == 1929 and round(float(rho.max())) == 2567
    model_inv = wave.AcousticModel(grid=grid, boundary=boundary, vp=vp_init.copy(),
                                   rho=rho_true.copy(), vp_grad=True, rho_grad=False,
                                   vp_bound=(float(vp_true.min()), float(vp_true.max())),
                                   auto_update_rho=True, device="cuda")
    prop_inv = wave.AcousticPropagator(model=model_inv, survey=survey, device="cuda")
    fwi = FWI(prop_inv, model_inv, d_obs, lr=10.0,
              grad_mask=torch.tensor(mask, device="cuda"), checkpoint_segments=10)
    losses = fwi.run(11)
    want0 = meta["notebook_loss"]["0"]
    want10 = meta["notebook_loss"]["10"]
    e0 = abs(losses[0] - want0) / want0
    e10 = abs(losses[10] - want10) / want10
    print(f"  iter 0 loss {losses[0]:.7e} (notebook {want0:.6e}, rel {e0:.1e}); "
          f"iter 10 {losses[10]:.7e} (notebook {want10:.6e}, rel {e10:.1e})")
    # the notebook prints 7 significant digits of an FP32 sum of 8000 trace norms
    assert e0 <= 2e-6
    # ten sign-dominated Adam steps amplify last-bit gradient differences a little
    assert e10 <= 1e-3
