"""The per-shot PyTorch baseline that bench.py times beside the HIP path is a
faithful restatement of the reference operator: bit-identical forward outputs
and autograd gradients within 1e-5 of the reference's on the golden inputs (CPU)."""

import numpy as np
import pytest
import torch

import bench
from tests._fixtures import FWD_CASES, GRAD_CASES, adjoint_sources, load, rel_l2


def _run(z, m, grad=False):
    t = lambda a, **kw: torch.as_tensor(np.asarray(a), **kw)  # noqa: E731
    vp = t(z["in_vp"], dtype=torch.float32).clone().requires_grad_(grad)
    out = bench.eager_forward(
        nz=m["nz"], nx=m["nx"], nabc=m["nabc"], free_surface=m["free_surface"], dt=m["dt"],
        dz=m["dx"], src_z=t(z["in_src_z"]), src_x=t(z["in_src_x"]),
        src_v=t(z["in_wavelet"], dtype=torch.float32), rcv_z=t(z["in_rcv_z"]),
        rcv_x=t(z["in_rcv_x"]), damp=t(z["in_damp"], dtype=torch.float32), vp=vp,
        rho=t(z["in_rho"], dtype=torch.float32), checkpoint_segments=m["segments"])
    return out, vp


@pytest.mark.parametrize("name", [n for n in FWD_CASES if n != "tiny_snap"])
def test_eager_forward_bitwise(name):
    z, m = load(name)
    with torch.no_grad():
        out, _ = _run(z, m)
    for k in ("p", "vx", "vz", "forward_wavefield_p", "forward_wavefield_vx",
              "forward_wavefield_vz"):
        np.testing.assert_array_equal(out[k].numpy(), z["out_" + k], err_msg=k)


@pytest.mark.parametrize("name", ["tiny0_grad", "tiny3_grad", "c1_grad"])
def test_eager_gradient(name):
    z, m = load(name)
    out, vp = _run(z, m, grad=True)
    g = adjoint_sources(z, m)
    sum((torch.as_tensor(gk) * out[k]).sum() for gk, k in zip(g, ("p", "vx", "vz"))).backward()
    assert rel_l2(vp.grad.numpy(), z["grad_vp"]) <= 1e-5
