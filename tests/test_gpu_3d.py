"""GPU parity of the 3D HIP propagator (``acoustic_forward_kernel_3d``) against
the reference's golden vectors and the CPU oracle.  Bar: seismograms bit for
bit; energy maps bit for bit when one thread steps every shot (the default),
relative L2 <= 1e-6 when the shots are split over several groups (the sums
are then added in another order)."""

import numpy as np
import pytest
import torch

import oracle
from geobrain_amd.physics import wave
from geobrain_amd.physics.wave import kernels as K
from tests._fixtures import FWD3D_CASES, OUT3D_KEYS, chunk_len, load, oracle3d_args, rel_l2

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
REC = ("p", "vx", "vy", "vz")
ENERGY = tuple(k for k in OUT3D_KEYS if k.startswith("forward_wavefield"))
ETOL = 1e-6


def _t(a, **kw):
    return torch.as_tensor(np.asarray(a), device=DEV, **kw)


def run3(a, segments=1, spb=0):
    """acoustic_forward_kernel_3d on oracle-style arguments ``a``."""
    K.set_shots_per_block_3d(spb)
    try:
        out = wave.acoustic_forward_kernel_3d(
            nx=a["nx"], ny=a["ny"], nz=a["nz"], dx=a["dz"], dy=a["dz"], dz=a["dz"],
            nt=np.asarray(a["src_v"]).shape[-1], dt=a["dt"], nabc=a["nabc"],
            free_surface=a["free_surface"], src_x=_t(a["src_x"]), src_y=_t(a["src_y"]),
            src_z=_t(a["src_z"]), src_n=len(a["src_x"]), src_v=_t(a["src_v"], dtype=torch.float32),
            rcv_x=_t(a["rcv_x"]), rcv_y=_t(a["rcv_y"]), rcv_z=_t(a["rcv_z"]),
            rcv_n=len(a["rcv_x"]), damp=_t(a["damp"], dtype=torch.float32),
            vp=_t(a["vp"], dtype=torch.float32), rho=_t(a["rho"], dtype=torch.float32),
            checkpoint_segments=segments)
    finally:
        K.set_shots_per_block_3d(0)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _check(got, want, exact_energy=True):
    for k in REC:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    for k in ENERGY:
        if exact_energy:
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)
        else:
            assert rel_l2(got[k], want[k]) <= ETOL, k


@pytest.mark.parametrize("name", FWD3D_CASES)
def test_3d_forward_matches_reference(name):
    print(f"[case] 3d golden {name}", flush=True)
    z, m = load(name)
    got = run3(oracle3d_args(z, m), segments=m["segments"])
    _check(got, {k: z["out_" + k] for k in OUT3D_KEYS})


@pytest.mark.parametrize("name", FWD3D_CASES)
def test_3d_shot_groups(name):
    print(f"[case] 3d golden {name} one shot per group", flush=True)
    z, m = load(name)
    got = run3(oracle3d_args(z, m), segments=m["segments"], spb=1)
    _check(got, {k: z["out_" + k] for k in OUT3D_KEYS}, exact_energy=False)


def random_case(nz, ny, nx, nabc, nt, S, fs, seed, shared=False, R=9):
    rng = np.random.default_rng(seed)
    vp = rng.uniform(1700, 3200, (nz, ny, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2500, (nz, ny, nx)).astype(np.float32)
    if nabc > 1:
        damp = wave.bc_pml_3d(nx, ny, nz, 10.0, 10.0, 10.0, nabc, float(vp.max()), False)
    else:  # the quadratic profile divides by pml - 1
        damp = np.zeros((nz + 2, ny + 2, nx + 2))
    w = rng.standard_normal(nt if shared else (S, nt)).astype(np.float32)
    # sources: interior, inside the absorbing layers, on the first padded row
    sz = rng.integers(0, nz, S)
    sy = rng.integers(0, ny, S)
    sx = rng.integers(0, nx, S)
    sz[0], sy[0], sx[0] = 1, ny // 2, nx // 2
    if S > 1:
        sz[1], sy[1], sx[1] = -nabc, -nabc + 1, nx + nabc - 1
    # receivers: random, plus the last index each axis allows and the first padded one
    rz = rng.integers(0, nz, R)
    ry = rng.integers(0, ny, R)
    rx = rng.integers(0, nx, R)
    rz[0], ry[0], rx[0] = nz + nabc - 2, ny + nabc - 2, nx + nabc - 2
    rz[1], ry[1], rx[1] = -nabc, -nabc, -nabc
    return dict(nz=nz, ny=ny, nx=nx, nabc=nabc, free_surface=fs, dt=1e-3, dz=10.0,
                src_z=sz, src_y=sy, src_x=sx, src_v=w, rcv_z=rz, rcv_y=ry, rcv_x=rx,
                vp=vp, rho=rho, damp=damp.astype(np.float32))


CASES = [
    # nz, ny, nx, nabc, nt, S, free surface, seed, shared wavelet
    (14, 11, 17, 4, 30, 3, False, 1, False),
    (14, 11, 17, 4, 30, 3, True, 2, True),
    (9, 12, 56, 4, 24, 2, True, 3, False),     # NX = 64: no padding columns
    (12, 7, 61, 5, 26, 4, False, 4, False),    # NX = 71: a second 64-column block
    (6, 6, 6, 1, 20, 1, False, 5, False),      # smallest absorbing layer
]


@pytest.mark.parametrize("case", CASES)
def test_3d_forward_matches_oracle(case):
    print(f"[case] 3d oracle {case}", flush=True)
    nz, ny, nx, nabc, nt, S, fs, seed, shared = case
    a = random_case(nz, ny, nx, nabc, nt, S, fs, seed, shared)
    want = oracle.forward3d(**a, chunk=chunk_len(nt, 3), precision="f32")
    got = run3(a, segments=3)
    _check(got, want)
    for spb in (1, 2):
        _check(run3(a, segments=3, spb=spb), want, exact_energy=False)


def test_3d_medium_grid_matches_oracle():
    print("[case] 3d oracle 40x36x90, 4 shots, 60 steps", flush=True)
    a = random_case(40, 36, 90, 8, 60, 4, True, 11, R=40)
    want = oracle.forward3d(**a, chunk=60, precision="f32")
    _check(run3(a), want)


def test_3d_linear_in_wavelet():
    """Scaling the wavelet by 2 scales the seismograms by 2 and the energies by
    4.  Exact for normal numbers; the leading edge of the wavefront underflows
    into subnormals, where a product's rounding does not commute with the
    scaling, hence a tolerance far below any physical signal."""
    print("[case] 3d linearity 60x50x70", flush=True)
    a = random_case(60, 50, 70, 10, 80, 2, False, 21, R=30)
    o1 = run3(a)
    a2 = dict(a, src_v=2.0 * np.asarray(a["src_v"]))
    o2 = run3(a2)
    for k in REC:
        assert rel_l2(o2[k], 2.0 * o1[k]) <= 1e-7, k
    for k in ENERGY:
        assert rel_l2(o2[k], 4.0 * o1[k]) <= 1e-7, k
    assert np.abs(o1["p"]).max() > 0


def test_3d_index_errors_and_gradient_refusal():
    print("[case] 3d errors", flush=True)
    a = random_case(10, 9, 11, 3, 5, 1, False, 31)
    bad = dict(a, rcv_z=np.where(np.arange(len(a["rcv_z"])) == 0, 10 + 3 - 1, a["rcv_z"]))
    with pytest.raises(IndexError):
        run3(bad)
    bad = dict(a, src_y=np.array([9 + 3]))
    with pytest.raises(IndexError):
        run3(bad)
    vp = _t(a["vp"], dtype=torch.float32).requires_grad_(True)
    with pytest.raises(NotImplementedError, match="3D acoustic gradient"):
        wave.acoustic_forward_kernel_3d(
            nx=11, ny=9, nz=10, dx=10.0, dy=10.0, dz=10.0, nt=5, dt=1e-3, nabc=3,
            free_surface=False, src_x=a["src_x"], src_y=a["src_y"], src_z=a["src_z"], src_n=1,
            src_v=_t(a["src_v"], dtype=torch.float32), rcv_x=a["rcv_x"], rcv_y=a["rcv_y"],
            rcv_z=a["rcv_z"], rcv_n=len(a["rcv_x"]), damp=_t(a["damp"]), vp=vp,
            rho=_t(a["rho"], dtype=torch.float32))
