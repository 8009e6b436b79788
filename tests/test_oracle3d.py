"""Pins the 3D CPU oracle: against the reference's own golden vectors
(acoustic_forward_kernel_3d forward and its autograd gradient, bc_pml_3d), and
against what the mathematics fixes independently (linearity, zero wavelet,
mirror symmetry in y and x, the y <-> x exchange of a model symmetric under it,
translation invariance before boundary contact, the energy-map / record
identity; for the gradient, the dot-product test of the adjoint and central
differences in FP64).  CPU only."""

import numpy as np
import pytest

import oracle
from tests._fixtures import (FWD3D_CASES, GRAD3D_CASES, MP3D_FWD_CASES, MP3D_GRAD_CASES,
                             OUT3D_KEYS, adjoint_sources_3d, chunk_len, load, oracle3d_args,
                             rel_l2)


@pytest.mark.parametrize("name", FWD3D_CASES + MP3D_FWD_CASES + MP3D_GRAD_CASES)
def test_oracle3d_fp32_forward_is_bitwise_reference(name):
    z, m = load(name)
    out = oracle.forward3d(**oracle3d_args(z, m), chunk=chunk_len(m["nt"], m["segments"]),
                           precision="f32")
    for k in OUT3D_KEYS:
        np.testing.assert_array_equal(out[k], z["out_" + k], err_msg=k)


@pytest.mark.parametrize("name", FWD3D_CASES + MP3D_FWD_CASES)
def test_oracle3d_fp64_forward_close_to_reference(name):
    z, m = load(name)
    out = oracle.forward3d(**oracle3d_args(z, m), chunk=chunk_len(m["nt"], m["segments"]),
                           precision="f64")
    for k in OUT3D_KEYS:
        assert rel_l2(out[k], z["out_" + k]) < 1e-5, k


def test_oracle3d_pml_matches_reference():
    z = np.load("tests/golden/pml.npz")
    for i in range(2):
        nx, ny, nz, dx, n, vmax, fs = z[f"args3d{i}"]
        got = oracle.bc_pml_3d(int(nx), int(ny), int(nz), dx, dx, dx, int(n), vmax, bool(fs))
        np.testing.assert_array_equal(got, z[f"pml3d{i}"])


def _case(nz=12, ny=13, nx=13, nabc=4, nt=30, S=1, seed=0, sym=False):
    rng = np.random.default_rng(seed)
    vp = rng.uniform(1800, 2600, (nz, ny, nx)).astype(np.float32)
    if sym:  # exactly symmetric in y and x about the centre, and under y <-> x
        assert ny == nx and ny % 2 == 1
        h = ny // 2
        for j in range(ny):
            for k in range(nx):
                a, b = abs(j - h), abs(k - h)
                a, b = min(a, b), max(a, b)
                vp[:, j, k] = vp[:, h + a, h + b]
    rho = np.full_like(vp, 2000.0)
    damp = oracle.bc_pml_3d(nx, ny, nz, 10.0, 10.0, 10.0, nabc, float(vp.max()), False)
    w = rng.standard_normal((S, nt)).astype(np.float32)
    return dict(nz=nz, ny=ny, nx=nx, nabc=nabc, free_surface=False, dt=1e-3, dz=10.0,
                src_z=np.full(S, 3), src_y=np.full(S, ny // 2), src_x=np.full(S, nx // 2),
                src_v=w, vp=vp, rho=rho, damp=damp.astype(np.float32))


def test_oracle3d_linear_and_zero_wavelet():
    a = _case()
    rz, ry, rx = np.array([1, 2, 5]), np.array([2, 6, 9]), np.array([3, 6, 10])
    o1 = oracle.forward3d(**a, rcv_z=rz, rcv_y=ry, rcv_x=rx, precision="f32")
    o2 = oracle.forward3d(**dict(a, src_v=2 * a["src_v"]), rcv_z=rz, rcv_y=ry, rcv_x=rx,
                          precision="f32")
    o0 = oracle.forward3d(**dict(a, src_v=0 * a["src_v"]), rcv_z=rz, rcv_y=ry, rcv_x=rx,
                          precision="f32")
    for k in ("p", "vx", "vy", "vz"):
        np.testing.assert_array_equal(o2[k], 2 * o1[k])  # scaling by 2 is exact in FP32
        assert not o0[k].any()
        assert np.abs(o1[k]).max() > 0


def test_oracle3d_mirror_symmetry():
    """A model symmetric in y and x about the source column and under y <-> x:
    p is even in y and x, vy odd in y (staggered half a cell), vx odd in x,
    and exchanging y and x maps vx onto vy.  Pins the direction and placement
    of the y and x staggering."""
    a = _case(sym=True)
    n = 13 + 2 * 4
    out = oracle.forward3d(**a, rcv_z=np.array([1]), rcv_y=np.array([0]), rcv_x=np.array([0]),
                           precision="f64", return_state=True)
    p, vx, vy, vz = (out["state"][f][0] for f in range(4))
    c = n // 2  # padded index of the source column (6 + 4 = 10)
    assert c == 6 + 4
    np.testing.assert_allclose(p, p[:, ::-1, :], rtol=0, atol=1e-12 * np.abs(p).max())
    np.testing.assert_allclose(p, p[:, :, ::-1], rtol=0, atol=1e-12 * np.abs(p).max())
    # vy sits at y+1/2: vy[y] = -vy[2c-1-y]; vx likewise in x
    vy_m = -vy[:, (2 * c - 1 - np.arange(n)) % n, :]
    vx_m = -vx[:, :, (2 * c - 1 - np.arange(n)) % n]
    inner = slice(1, n - 2)
    np.testing.assert_allclose(vy[:, inner, :], vy_m[:, inner, :], rtol=0,
                               atol=1e-12 * np.abs(vy).max())
    np.testing.assert_allclose(vx[:, :, inner], vx_m[:, :, inner], rtol=0,
                               atol=1e-12 * np.abs(vx).max())
    np.testing.assert_allclose(vx, vy.transpose(0, 2, 1), rtol=0, atol=1e-12 * np.abs(vx).max())
    np.testing.assert_allclose(p, p.transpose(0, 2, 1), rtol=0, atol=1e-12 * np.abs(p).max())
    assert np.abs(vz).max() > 0


def test_oracle3d_translation_invariance_is_bitwise():
    """Before the wave reaches the PML (damping exactly zero there), shifting
    source and receivers by one cell in y shifts the records bit for bit."""
    a = _case(nz=14, ny=30, nx=30, nabc=4, nt=12, seed=3)
    a["vp"][:] = 2000.0
    a["damp"] = oracle.bc_pml_3d(30, 30, 14, 10.0, 10.0, 10.0, 4, 2000.0, False).astype(np.float32)
    rz, ry, rx = np.array([3, 5]), np.array([12, 17]), np.array([15, 11])
    o1 = oracle.forward3d(**a, rcv_z=rz, rcv_y=ry, rcv_x=rx, precision="f32")
    b = dict(a, src_y=a["src_y"] + 1)
    o2 = oracle.forward3d(**b, rcv_z=rz, rcv_y=ry + 1, rcv_x=rx, precision="f32")
    for k in ("p", "vx", "vy", "vz"):
        np.testing.assert_array_equal(o1[k], o2[k])


def test_oracle3d_energy_equals_sum_of_squared_records():
    """The energy map at a receiver cell is the sum over shots and steps of
    that cell's squared records."""
    a = _case(nt=20, S=2, seed=5)
    rz, ry, rx = np.array([2, 4, 6]), np.array([3, 7, 9]), np.array([2, 8, 5])
    out = oracle.forward3d(**a, rcv_z=rz, rcv_y=ry, rcv_x=rx, precision="f64")
    for k in ("p", "vx", "vy", "vz"):
        e = out["forward_wavefield_" + k][rz, ry, rx]
        np.testing.assert_allclose(e, (out[k] ** 2).sum(axis=(0, 1)), rtol=1e-12)


@pytest.mark.parametrize("name", GRAD3D_CASES + MP3D_GRAD_CASES)
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_oracle3d_gradient_matches_reference_autograd(name, precision):
    z, m = load(name)
    g = adjoint_sources_3d(z, m)
    gv, gr, gw = oracle.gradient3d(**oracle3d_args(z, m), g_p=g[0], g_vx=g[1], g_vy=g[2],
                                   g_vz=g[3], chunk=chunk_len(m["nt"], m["segments"]),
                                   precision=precision)
    assert rel_l2(gv, z["grad_vp"]) < 1e-5
    if "grad_rho" in z:
        assert rel_l2(gr, z["grad_rho"]) < 1e-5
    assert rel_l2(gw, z["grad_w"]) < 1e-5


def _grad_case(seed, fs, nt=40, S=2, R=6, nz=9, ny=8, nx=10, nabc=3):
    rng = np.random.default_rng(seed)
    vp = rng.uniform(1600, 3000, (nz, ny, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2400, (nz, ny, nx)).astype(np.float32)
    damp = oracle.bc_pml_3d(nx, ny, nz, 10.0, 10.0, 10.0, nabc, float(vp.max()), False)
    rz = rng.integers(-1 if fs else 0, nz, R)
    # receivers on the padded planes nabc - 1 and nabc: with a free surface the
    # mirrored plane f-1 and the surface f, whose transposes fold into f and f+1
    rz[0], rz[1] = -1, 0
    return dict(nz=nz, ny=ny, nx=nx, nabc=nabc, free_surface=fs, dt=1e-3, dz=10.0,
                src_z=rng.integers(0, nz, S), src_y=rng.integers(0, ny, S),
                src_x=rng.integers(0, nx, S),
                src_v=rng.standard_normal((S, nt)).astype(np.float32),
                rcv_z=rz, rcv_y=rng.integers(0, ny, R),
                rcv_x=rng.integers(0, nx, R), vp=vp, rho=rho, damp=damp.astype(np.float32))


@pytest.mark.parametrize("fs", [False, True])
def test_oracle3d_dot_product(fs):
    """<F w, g> == <w, F^T g>: the adjoint propagation is the exact transpose."""
    a = _grad_case(3, fs)
    S, nt = a["src_v"].shape
    R = len(a["rcv_z"])
    rng = np.random.default_rng(5)
    gs = [rng.standard_normal((S, nt, R)) for _ in range(4)]
    out = oracle.forward3d(**a, precision="f64")
    lhs = sum(float(np.sum(out[k] * g)) for k, g in zip(("p", "vx", "vy", "vz"), gs))
    _, _, gw = oracle.gradient3d(**a, g_p=gs[0], g_vx=gs[1], g_vy=gs[2], g_vz=gs[3], chunk=7,
                                 precision="f64")
    rhs = float(np.sum(a["src_v"].astype(np.float64) * gw))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


@pytest.mark.parametrize("fs", [False, True])
def test_oracle3d_gradient_finite_differences(fs):
    """Central differences in FP64 on vp and rho: interior cells, a corner cell
    (its gradient folds the padded cells it is copied to) and an edge cell."""
    a = _grad_case(11, fs, nt=36)
    S, nt = a["src_v"].shape
    R = len(a["rcv_z"])
    rng = np.random.default_rng(2)
    gs = [rng.standard_normal((S, nt, R)) for _ in range(4)]

    def loss(**kw):
        out = oracle.forward3d(**dict(a, **kw), precision="f64")
        return sum(float(np.sum(out[k] * g)) for k, g in zip(("p", "vx", "vy", "vz"), gs))

    gv, gr, _ = oracle.gradient3d(**a, g_p=gs[0], g_vx=gs[1], g_vy=gs[2], g_vz=gs[3], chunk=8,
                                  precision="f64")
    # the oracle reads float32 model values: perturb by exactly representable steps
    for cell in [(0, 0, 0), (4, 3, 5), (8, 7, 9), (2, 0, 6), (6, 5, 1)]:
        for name, grad, h in (("vp", gv, 0.5), ("rho", gr, 0.25)):
            up, dn = a[name].copy(), a[name].copy()
            up[cell] += h
            dn[cell] -= h
            fd = (loss(**{name: up}) - loss(**{name: dn})) / (2 * h)
            assert abs(fd - grad[cell]) <= 1e-5 * abs(grad).max(), (name, cell, fd, grad[cell])
