"""Receiver sampling of the two-step forward kernel against the FP32 oracle.
The kernel keeps the first receiver of each cell per tile and walks the list
only for cells with several, so the cases put three receivers on one cell (in
shuffled order), receivers on rows either side of tile boundaries (padded rows
15/16, 31/32), on the first and last columns and either side of a 128-column
tile edge, and, with a free surface, on the mirrored row f-1 (its vz sample
comes from row f) and the surface row f; seismograms bit for bit, for shot
groups of 1, 3 and all shots and the default launch."""

import numpy as np
import pytest

import oracle
from geobrain_amd.physics.wave import kernels as K
from tests._fixtures import chunk_len
from tests.test_gpu_adj_march import _problem
from tests.test_gpu_fwd_march import _run

pytestmark = pytest.mark.gpu


def _receivers(nz, nx, fs, seed):
    rz = [0, 0, 0, -5, -4, 11, 12, nz - 2, 3, 3, 7, 7, 20, 20]
    rx = [5, 5, 5, 40, 41, 87, 88, 17, 0, nx - 2, 107, 108, 108, 109]
    if fs:
        rz += [-1, -1, -1, 0]
        rx += [33, 33, 90, 33]
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(rz))
    return np.asarray(rx)[order], np.asarray(rz)[order]


@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("spb", [0, 1, 3, 5])
def test_receivers_match_oracle(fs, spb):
    print(f"[case] receivers vs FP32 oracle, free_surface={fs}, shots per block {spb}", flush=True)
    nz, nx, nt, S = 60, 250, 60, 5
    p = _problem(nz, nx, nt, S, seed=43)
    p["rx"], p["rz"] = _receivers(nz, nx, fs, seed=spb)
    launch = K.LaunchChoice(0, spb, 0, 0, spb, 2, 1) if spb else None
    got = _run(p, launch, fs=fs, segments=2)
    want = oracle.forward(nz=nz, nx=nx, nabc=20, free_surface=fs, dt=5e-4, dz=5.0,
                          src_z=p["sz"], src_x=p["sx"], src_v=p["w"], rcv_z=p["rz"],
                          rcv_x=p["rx"], vp=p["vp"], rho=p["rho"], damp=p["damp"],
                          chunk=chunk_len(nt, 2), precision="f32")
    assert float(np.abs(want["p"]).max()) > 0
    for k in ("p", "vx", "vz"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
