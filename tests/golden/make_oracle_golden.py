"""Write the FP64 oracle's gradient for the full-size single-shot problem.

    python tests/golden/make_oracle_golden.py

The problem is the 350 x 1700 model (+20 absorbing cells per side), dx 5 m,
dt 0.5 ms, 4000 steps, one 15 Hz Ricker shot at (z=2, x=850), 1700 surface
receivers, loss sum((p - d_obs)^2) with d_obs modelled in the 5 %-smoothed
model -- the same recipe as ``make_golden.py``'s ``c2_full_grad`` (the
reference's autograd gradient).  Everything here is FP64 on the CPU oracle:
both forwards, the residual and the adjoint with checkpoints every 50 steps.
The FP64 run takes about 1.5 minutes on 8 cores, too long for every GPU test
run, so its result is stored (as float32, far below the 1e-5 comparison bar)
with the wavelet and digests of the regenerated model inputs.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from geobrain_amd import synthetic  # noqa: E402
from geobrain_amd.physics import wave  # noqa: E402


def c2_full_inputs():
    """The full-size single-shot inputs, regenerated from their seeded recipe."""
    nz, nx, nt, nabc, dx, dt, f0 = 350, 1700, 4000, 20, 5.0, 5e-4, 15.0
    vp = synthetic.grf_model(nz, nx, dx, seed=0)
    return dict(nz=nz, nx=nx, nt=nt, nabc=nabc, dx=dx, dt=dt, vp=vp,
                vp_obs=synthetic.smooth_model(vp, 0.05), rho=synthetic.gardner(vp),
                damp=wave.bc_pml(nx, nz, dx, dx, nabc, float(vp.max()), False).astype(np.float32),
                wavelet=synthetic.ricker_padded(f0, dt, nt)[None, :],
                src_x=np.array([nx // 2], np.int64), src_z=np.array([2], np.int64),
                rcv_x=np.arange(nx, dtype=np.int64), rcv_z=np.full(nx, 2, np.int64), chunk=50)


def digests(w, keys=("vp", "rho", "damp", "wavelet", "vp_obs")):
    return {k: hashlib.sha256(np.ascontiguousarray(w[k]).tobytes()).hexdigest() for k in keys}


MODEL_KEYS = ("vp", "rho", "damp", "vp_obs")


def main():
    w = c2_full_inputs()
    args = dict(nz=w["nz"], nx=w["nx"], nabc=w["nabc"], free_surface=False, dt=w["dt"],
                dz=w["dx"], src_z=w["src_z"], src_x=w["src_x"], src_v=w["wavelet"],
                rcv_z=w["rcv_z"], rcv_x=w["rcv_x"], vp=w["vp"], rho=w["rho"], damp=w["damp"])
    t0 = time.time()
    p = oracle.forward(**args, precision="f64")["p"]
    d = oracle.forward(**dict(args, vp=w["vp_obs"]), precision="f64")["p"]
    g, _, _ = oracle.gradient(**args, g_p=2.0 * (p - d), chunk=w["chunk"], precision="f64")
    meta = dict(name="c2_full_fp64", loss=float(np.sum((p - d) ** 2)), nt=w["nt"],
                checkpoint_every=w["chunk"], seconds=round(time.time() - t0, 1),
                recipe="make_oracle_golden.c2_full_inputs()", in_sha256=digests(w))
    path = os.path.join(HERE, "c2_full_fp64.npz")
    # the wavelet is stored: torch's CPU exp (inside the Ricker) rounds
    # differently on other hosts, so a regenerated wavelet may differ in the last bit
    np.savez_compressed(path, grad_vp=g.astype(np.float32), in_wavelet=w["wavelet"],
                        meta=np.array(json.dumps(meta)))
    print(f"c2_full_fp64: {meta['seconds']} s, {os.path.getsize(path) / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
