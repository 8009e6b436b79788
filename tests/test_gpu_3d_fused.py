"""The fused 3D step (launch choice step_kernel = 2, fwd3f_kernel) against the
reference's golden vectors, the FP32 oracle and the separate p / v step
kernels: seismograms bit for bit, energy maps bit for bit with one shot group
(relative L2 <= 1e-6 with several).  The cases cover windows that span several
61-column strips and 13-row tiles (the last of each partial), z segments whose
boundaries fall on the mirrored plane f-1, on f and on f+1 with a free
surface, 1-8 shots per group and groups of several shots, sources and
receivers on window and segment edges, and the benchmark's 8-shot 240^3
problem for a shortened time axis."""

import numpy as np
import pytest
import torch

import oracle
from tests._fixtures import FWD3D_CASES, OUT3D_KEYS, chunk_len, load, oracle3d_args
from tests.test_gpu_3d import CASES, _check, random_case, run3

pytestmark = pytest.mark.gpu
FUSED = 2


@pytest.mark.parametrize("name", FWD3D_CASES)
def test_fused_matches_reference(name):
    print(f"[case] fused 3d golden {name}", flush=True)
    z, m = load(name)
    got = run3(oracle3d_args(z, m), segments=m["segments"], kernel=FUSED)
    _check(got, {k: z["out_" + k] for k in OUT3D_KEYS})


@pytest.mark.parametrize("case", CASES)
def test_fused_matches_oracle(case):
    print(f"[case] fused 3d oracle {case}", flush=True)
    nz, ny, nx, nabc, nt, S, fs, seed, shared = case
    a = random_case(nz, ny, nx, nabc, nt, S, fs, seed, shared)
    want = oracle.forward3d(**a, chunk=chunk_len(nt, 3), precision="f32")
    _check(run3(a, segments=3, kernel=FUSED), want)
    for spb, zc in ((1, 0), (2, 3), (4, 5), (8, 2)):
        _check(run3(a, segments=3, spb=spb, kernel=FUSED, zc=zc), want,
               exact_energy=(spb >= S))


@pytest.mark.parametrize("zc", [0, 7, 8, 9, 10])
def test_fused_segments_around_the_surface(zc):
    """z segments of 7, 8, 9 and 10 planes put segment boundaries at the
    mirrored plane f-1, at f (moved to f+1 by the kernel) and at f+1 (f = 8);
    the 90 x 36 grid (106 x 52 padded) spans two strips and four row tiles."""
    print(f"[case] fused 3d oracle 40x36x90, 4 shots, 60 steps, z segment {zc}", flush=True)
    a = random_case(40, 36, 90, 8, 60, 4, True, 11, R=40)
    want = oracle.forward3d(**a, chunk=60, precision="f32")
    _check(run3(a, spb=4, kernel=FUSED, zc=zc), want)
    _check(run3(a, spb=2, kernel=FUSED, zc=zc), want, exact_energy=False)


def test_fused_window_edges():
    """Sources and receivers on the first and last owned column of a strip
    (x = 0, 60, 61, 121, 122) and row of a tile (y = 0, 12, 13, 25, 26), 9
    shots (one group of 8 and one of 1), without a free surface."""
    print("[case] fused 3d window edges 20x40x150, 9 shots", flush=True)
    nabc = 3
    a = random_case(20, 40, 150, nabc, 25, 9, False, 41, R=12)
    xs = np.array([0, 60, 61, 121, 122, 59, 62, 120, 123]) - nabc
    ys = np.array([0, 12, 13, 25, 26, 38, 39, 11, 14]) - nabc
    a["src_x"], a["src_y"] = np.clip(xs, -nabc, 150 + nabc - 1), np.clip(ys, -nabc, 40 + nabc - 1)
    a["src_z"] = np.array([0, 1, 2, 5, 10, 19, 3, 4, 6]) - nabc
    a["rcv_x"][:9], a["rcv_y"][:9] = np.clip(xs, -nabc, 150 + nabc - 2), np.clip(ys, -nabc, 40 + nabc - 2)
    want = oracle.forward3d(**a, chunk=25, precision="f32")
    _check(run3(a, kernel=FUSED), want, exact_energy=False)
    _check(run3(a, spb=8, kernel=FUSED, zc=6), want, exact_energy=False)


def test_fused_equals_split_kernels_at_benchmark_size():
    """The benchmark problem (8 shots, 200^3 grid + 20 absorbing layers) for 60
    steps: the fused step and the separate p / v kernels give the same
    seismograms and energy maps bit for bit."""
    print("[case] fused 3d vs split kernels, 8 shots, 240^3 padded, 60 steps", flush=True)
    from bench import workload_3d
    w = workload_3d(n_shots=8, nt=60)
    a = dict(nz=w["nz"], ny=w["ny"], nx=w["nx"], nabc=w["nabc"], free_surface=False, dt=w["dt"],
             dz=w["dx"], src_z=w["src_z"], src_y=w["src_y"], src_x=w["src_x"], src_v=w["wavelet"],
             rcv_z=w["rcv_z"], rcv_y=w["rcv_y"], rcv_x=w["rcv_x"], vp=w["vp"], rho=w["rho"],
             damp=w["damp"])
    split = run3(a, kernel=0)
    fused = run3(a, kernel=FUSED)
    for k in OUT3D_KEYS:
        np.testing.assert_array_equal(fused[k], split[k], err_msg=k)
    assert np.abs(fused["p"]).max() > 0
    torch.cuda.empty_cache()
