"""Pins the CPU oracle: against the reference's own golden vectors, and against
what the mathematics fixes independently of the reference (dot-product test,
finite differences in FP64, linearity, translation and mirror symmetry,
the energy-map / seismogram identity).  CPU only."""

import json

import numpy as np
import pytest

import oracle
from tests._fixtures import (FWD_CASES, GRAD_CASES, MP_FWD_CASES, MP_GRAD_CASES, OUT_KEYS,
                             adjoint_sources, chunk_len, fd_c1, load, oracle_args, rel_l2)


@pytest.mark.parametrize("name", FWD_CASES + GRAD_CASES + MP_FWD_CASES + MP_GRAD_CASES)
def test_oracle_fp32_forward_is_bitwise_reference(name):
    z, m = load(name)
    out = oracle.forward(**oracle_args(z, m), chunk=chunk_len(m["nt"], m["segments"]),
                         snapshot_every=m["snapshot_every"], precision="f32")
    checked = 0
    for k in OUT_KEYS:
        if "out_" + k in z:
            np.testing.assert_array_equal(out[k], z["out_" + k], err_msg=k)
            checked += 1
    assert checked >= 3


@pytest.mark.parametrize("name", FWD_CASES + MP_FWD_CASES)
def test_oracle_fp64_forward_close_to_reference(name):
    z, m = load(name)
    out = oracle.forward(**oracle_args(z, m), chunk=chunk_len(m["nt"], m["segments"]),
                         snapshot_every=m["snapshot_every"], precision="f64")
    for k in OUT_KEYS:
        if "out_" + k in z:
            assert rel_l2(out[k], z["out_" + k]) < 1e-5, k


@pytest.mark.parametrize("name", GRAD_CASES + MP_GRAD_CASES)
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_oracle_gradient_matches_reference_autograd(name, precision):
    z, m = load(name)
    args = oracle_args(z, m)
    obs = fwd = None
    if "g_seed" not in m:
        fwd = oracle.forward(**args, precision="f32")["p"]
        obs = oracle.forward(**dict(args, vp=z["in_vp_obs"]), precision="f32")["p"]
    g = adjoint_sources(z, m, fwd, obs)
    gv, gr, gw = oracle.gradient(**args, g_p=g[0], g_vx=g[1], g_vz=g[2],
                                 chunk=chunk_len(m["nt"], m["segments"]), precision=precision,
                                 wavelet_shared=z["in_wavelet"].ndim == 1)
    assert rel_l2(gv, z["grad_vp"]) < 1e-5
    if "grad_rho" in z:
        assert rel_l2(gr, z["grad_rho"]) < 1e-5
    if "grad_w" in z:
        assert rel_l2(gw, z["grad_w"]) < 1e-5


def _small_problem(seed=0, nz=12, nx=15, nabc=4, fs=False, S=2, R=7, nt=60):
    rng = np.random.default_rng(seed)
    vp = rng.uniform(1500, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2400, (nz, nx)).astype(np.float32)
    damp = oracle.bc_pml(nx, nz, 10.0, 10.0, nabc, float(vp.max()), False).astype(np.float32)
    src_z, src_x = rng.integers(0, nz, S), rng.integers(0, nx, S)
    src_v = rng.standard_normal((S, nt)).astype(np.float32)
    rcv_z, rcv_x = rng.integers(0, nz, R), rng.integers(0, nx, R)
    # receivers on the padded rows nabc - 1 and nabc: with a free surface the
    # mirrored row f-1 and the surface f, whose transposes fold into f and f+1
    rcv_z[0], rcv_z[1] = -1, 0
    return dict(nz=nz, nx=nx, nabc=nabc, free_surface=fs, dt=1e-3, dz=10.0,
                src_z=src_z, src_x=src_x, src_v=src_v, rcv_z=rcv_z, rcv_x=rcv_x,
                vp=vp, rho=rho, damp=damp)


@pytest.mark.parametrize("fs", [False, True])
def test_oracle_dot_product(fs):
    """<F w, g> == <w, F^T g>: the adjoint propagation is the exact transpose."""
    args = _small_problem(seed=3, fs=fs)
    rng = np.random.default_rng(5)
    S, nt, R = args["src_v"].shape[0], args["src_v"].shape[1], len(args["rcv_z"])
    gs = [rng.standard_normal((S, nt, R)) for _ in range(3)]
    out = oracle.forward(**args, precision="f64")
    lhs = sum(float(np.sum(out[k] * g)) for k, g in zip(("p", "vx", "vz"), gs))
    _, _, gw = oracle.gradient(**args, g_p=gs[0], g_vx=gs[1], g_vz=gs[2], chunk=7,
                               precision="f64")
    rhs = float(np.sum(args["src_v"].astype(np.float64) * gw))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


@pytest.mark.parametrize("fs", [False, True])
def test_oracle_gradient_finite_differences(fs):
    """Central differences in FP64 on vp and rho (the imaging condition and chain rule)."""
    args = _small_problem(seed=11, fs=fs, nt=50)
    rng = np.random.default_rng(2)
    S, nt, R = args["src_v"].shape[0], args["src_v"].shape[1], len(args["rcv_z"])
    gs = [rng.standard_normal((S, nt, R)) for _ in range(3)]

    def loss(vp=None, rho=None):
        a = dict(args)
        if vp is not None:
            a["vp"] = vp
        if rho is not None:
            a["rho"] = rho
        out = oracle.forward(**a, precision="f64")
        return sum(float(np.sum(out[k] * g)) for k, g in zip(("p", "vx", "vz"), gs))

    gv, gr, _ = oracle.gradient(**args, g_p=gs[0], g_vx=gs[1], g_vz=gs[2], chunk=9,
                                precision="f64")
    # the oracle reads float32 model values, so perturb by an exactly representable step
    for (a, b) in [(0, 0), (5, 7), (11, 14), (3, 0), (6, 9)]:
        for name, grad, h in (("vp", gv, 0.5), ("rho", gr, 0.25)):
            base = args[name].copy()
            up, dn = base.copy(), base.copy()
            up[a, b] += h
            dn[a, b] -= h
            fd = (loss(**{name: up}) - loss(**{name: dn})) / (2 * h)
            assert abs(fd - grad[a, b]) <= 1e-5 * abs(grad).max(), (name, a, b, fd, grad[a, b])


@pytest.mark.parametrize("name", ["vp", "rho"])
def test_oracle_gradient_finite_differences_config1(name):
    """The config-1 problem (100 x 200, 1000 steps): FP64 central differences,
    vp +- 1e-2 m/s (rho +- 1e-2 kg/m^3) on 16 seeded cells, against the FP64
    oracle's gradient."""
    z, m, gs, ii, jj, fd = fd_c1(name)
    gv, gr, _ = oracle.gradient(**oracle_args(z, m), g_p=gs[0], g_vx=gs[1], g_vz=gs[2],
                                chunk=chunk_len(m["nt"], m["segments"]), precision="f64")
    g = gv if name == "vp" else gr
    scale = np.abs(g).max()
    assert scale > 0
    err = np.abs(fd - g[ii, jj]) / scale
    assert err.max() <= 1e-8, (err.max(), list(zip(ii, jj, fd, g[ii, jj])))


@pytest.mark.parametrize("fs", [False, True])
def test_oracle_multipoint_superposition(fs):
    """K point sources in one shot against K one-point shots (FP64): the
    seismograms add (the scheme is linear in the source terms), the energy
    maps do not (they square the summed field), and the per-point wavelet
    gradient equals the one-point shots' wavelet gradients for the same
    adjoint sources.  Distinct cells; a point on the surface row with a free
    surface."""
    rng = np.random.default_rng(17 + fs)
    nz, nx, nabc, nt, K, R = 11, 13, 4, 50, 3, 5
    vp = rng.uniform(1500, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2400, (nz, nx)).astype(np.float32)
    damp = oracle.bc_pml(nx, nz, 10.0, 10.0, nabc, float(vp.max()), False).astype(np.float32)
    sz, sx = np.array([[0, 4, 7]]), np.array([[2, 9, 5]])
    w = rng.standard_normal((1, K, nt)).astype(np.float32)
    rz, rx = rng.integers(0, nz, R), rng.integers(0, nx, R)
    common = dict(nz=nz, nx=nx, nabc=nabc, free_surface=fs, dt=1e-3, dz=10.0, rcv_z=rz, rcv_x=rx,
                  vp=vp, rho=rho, damp=damp)
    multi = oracle.forward(**common, src_z=sz, src_x=sx, src_v=w, precision="f64")
    single = oracle.forward(**common, src_z=sz[0], src_x=sx[0], src_v=w[0], precision="f64")
    for k in ("p", "vx", "vz"):
        np.testing.assert_allclose(multi[k][0], single[k].sum(axis=0), rtol=0,
                                   atol=1e-12 * np.abs(single[k]).max())
    assert rel_l2(multi["forward_wavefield_p"], single["forward_wavefield_p"]) > 1e-3
    g = rng.standard_normal((1, nt, R))
    _, _, gw = oracle.gradient(**common, src_z=sz, src_x=sx, src_v=w, g_p=g, precision="f64")
    _, _, gw1 = oracle.gradient(**common, src_z=sz[0], src_x=sx[0], src_v=w[0],
                                g_p=np.repeat(g, K, axis=0), precision="f64")
    np.testing.assert_allclose(gw[0], gw1, rtol=1e-12, atol=1e-18)
    # one wavelet for the shot: its gradient is the sum over the points
    _, _, gws = oracle.gradient(**common, src_z=sz, src_x=sx, src_v=w[:, 0], g_p=g,
                                precision="f64")
    np.testing.assert_allclose(gws[0], gw1.sum(axis=0), rtol=1e-12, atol=1e-18)


def test_oracle_linear_and_zero_wavelet():
    args = _small_problem(seed=4)
    zero = oracle.forward(**dict(args, src_v=np.zeros_like(args["src_v"])), precision="f32")
    for k in ("p", "vx", "vz", "forward_wavefield_p"):
        assert not np.any(zero[k])
    w2 = np.random.default_rng(9).standard_normal(args["src_v"].shape).astype(np.float32)
    o1 = oracle.forward(**args, precision="f64")
    o2 = oracle.forward(**dict(args, src_v=w2), precision="f64")
    o12 = oracle.forward(**dict(args, src_v=args["src_v"] + w2), precision="f64")
    for k in ("p", "vx", "vz"):
        assert rel_l2(o12[k], o1[k] + o2[k]) < 1e-7


def test_oracle_translation_invariance_is_bitwise():
    """Homogeneous interior, no PML contact: shifting source and receivers shifts
    the traces exactly (the undamped update multiplies by exactly 1)."""
    nz, nx, nabc, nt = 40, 80, 10, 40
    vp = np.full((nz, nx), 2000.0, np.float32)
    rho = np.full((nz, nx), 2100.0, np.float32)
    damp = oracle.bc_pml(nx, nz, 10.0, 10.0, nabc, 2000.0, False).astype(np.float32)
    w = np.random.default_rng(1).standard_normal((1, nt)).astype(np.float32)
    base = dict(nz=nz, nx=nx, nabc=nabc, free_surface=False, dt=1e-3, dz=10.0, src_v=w,
                vp=vp, rho=rho, damp=damp, precision="f32")
    rx = np.arange(30, 36)
    o1 = oracle.forward(**base, src_z=[20], src_x=[33], rcv_z=np.full(6, 18), rcv_x=rx)
    o2 = oracle.forward(**base, src_z=[20], src_x=[45], rcv_z=np.full(6, 18), rcv_x=rx + 12)
    assert np.any(o1["p"])
    for k in ("p", "vx", "vz"):
        np.testing.assert_array_equal(o1[k], o2[k])


def test_oracle_mirror_symmetry():
    """Model symmetric in x, source on the centre column: p is even about the
    source column and vx (staggered half a cell right) is odd."""
    nz, nx, nabc, nt = 20, 31, 6, 80
    rng = np.random.default_rng(3)
    half = rng.uniform(1500, 3000, (nz, nx // 2 + 1)).astype(np.float32)
    vp = np.concatenate([half, half[:, -2::-1]], axis=1)
    rho = np.full((nz, nx), 2000.0, np.float32)
    damp = oracle.bc_pml(nx, nz, 10.0, 10.0, nabc, float(vp.max()), False).astype(np.float32)
    xc = nx // 2
    d = np.arange(1, 8)
    out = oracle.forward(nz=nz, nx=nx, nabc=nabc, free_surface=True, dt=1e-3, dz=10.0,
                         src_z=[3], src_x=[xc],
                         src_v=rng.standard_normal((1, nt)).astype(np.float32),
                         rcv_z=np.full(3 * len(d), 5),
                         rcv_x=np.concatenate([xc + d, xc - d, xc - d - 1]),
                         vp=vp, rho=rho, damp=damp, precision="f64")
    n = len(d)
    p, vx = out["p"][0], out["vx"][0]
    scale = np.abs(p).max()
    np.testing.assert_allclose(p[:, :n], p[:, n:2 * n], atol=1e-12 * scale)
    np.testing.assert_allclose(vx[:, :n], -vx[:, 2 * n:], atol=1e-12 * np.abs(vx).max())


def test_oracle_energy_equals_sum_of_squared_records():
    args = _small_problem(seed=8, S=3, R=5)
    out = oracle.forward(**args, precision="f64")
    for k in ("p", "vx", "vz"):
        e = out["forward_wavefield_" + k]
        for r in range(len(args["rcv_z"])):
            if args["rcv_z"][r] < 0:  # in the absorbing layer: no energy-map cell
                continue
            want = float(np.sum(out[k][:, :, r] ** 2))
            assert abs(e[args["rcv_z"][r], args["rcv_x"][r]] - want) <= 1e-12 * max(want, 1e-300)


def test_oracle_pml_matches_reference():
    z = np.load("tests/golden/pml.npz")
    for i in range(4):
        nx, nz, dx, n, vmax, fs = z[f"args{i}"]
        got = oracle.bc_pml(int(nx), int(nz), dx, dx, int(n), vmax, bool(fs))
        np.testing.assert_array_equal(got, z[f"pml{i}"])


def test_oracle_pml_closed_form():
    d = oracle.bc_pml(30, 20, 5.0, 5.0, 8, 3000.0, False)
    kappa = -3.0 * 3000.0 * np.log(1e-6) / (2.0 * 7 * 5.0)
    assert d.shape == (36, 46)
    assert np.all(d[8:28, 8:38] == 0)                       # interior undamped
    assert np.isclose(d[18, 0], kappa) and np.isclose(d[18, 45], kappa)  # outermost = kappa
    np.testing.assert_allclose(d[18, :8][::-1], kappa * (np.arange(8) / 7.0) ** 2)
    np.testing.assert_array_equal(d, d[::-1, :])             # top/bottom symmetric
    np.testing.assert_array_equal(d, d[:, ::-1])             # left/right symmetric


def test_oracle_pad_model_is_edge_padding():
    """pad_model / pad_model_3d against numpy's edge-mode padding, a library
    routine that replicates edges and corners as the reference's chained
    copies do (acoustic_2d.py:18-66, :69-120), on shapes with every side
    shorter and longer than the pad width."""
    rng = np.random.default_rng(3)
    for shape, n in (((3, 4), 2), ((1, 5), 3), ((6, 2), 4), ((7, 9), 1)):
        v = rng.standard_normal(shape).astype(np.float32)
        np.testing.assert_array_equal(oracle.pad_model(v, n), np.pad(v, n, mode="edge"))
    for shape, n in (((2, 3, 4), 2), ((1, 1, 3), 3), ((5, 4, 2), 1)):
        v = rng.standard_normal(shape).astype(np.float32)
        np.testing.assert_array_equal(oracle.pad_model_3d(v, n), np.pad(v, n, mode="edge"))
    # an x <-> z transposed restatement would fail this non-square case
    v = np.arange(6, dtype=np.float32).reshape(2, 3)
    p = oracle.pad_model(v, 1)
    assert p.shape == (4, 5) and p[0, 4] == 2 and p[3, 0] == 3


def test_oracle_ricker_known_values():
    """Closed-form facts of (1 - 2a^2) exp(-a^2), a = pi f0 t
    (wavelets/ricker.py:82-92): length 2 int(2.2/(f0 dt)/2) + 1, peak 1 at
    t = 0, zeros at t = +-1/(pi f0 sqrt 2), value -1/e at t = +-1/(pi f0).
    A dropped pi, a (1 - a^2) factor or exp(-2a^2) each fails one of them."""
    f0 = 10.0
    w = oracle.ricker(f0, 1e-3)
    assert len(w) == 221 and w[110] == 1.0
    np.testing.assert_allclose(w, w[::-1], rtol=0, atol=0)
    t_zero = 1.0 / (np.pi * f0 * np.sqrt(2.0))
    w = oracle.ricker(f0, t_zero / 5)           # the zeros fall on samples +-5
    c = len(w) // 2
    assert abs(w[c + 5]) < 1e-14 and abs(w[c - 5]) < 1e-14
    assert w[c + 4] > 0 and w[c + 6] < 0
    w = oracle.ricker(f0, 1.0 / (np.pi * f0) / 4)   # a = 1 at samples +-4
    c = len(w) // 2
    np.testing.assert_allclose([w[c + 4], w[c - 4]], -np.exp(-1.0), rtol=1e-14)
    assert len(oracle.ricker(25.0, 1e-3, length_factor=3.0)) == 2 * int(120 / 2) + 1


def test_oracle_ricker_matches_reference_fp32():
    """The reference's own FP32 RickerWavelet outputs (tests/golden/ricker.npz,
    written by make_golden.py from wavelets/ricker.py): the float64 oracle is
    within FP32 rounding of every sample and has the same length."""
    z = np.load("tests/golden/ricker.npz")
    for i in range(json.loads(str(z["meta"]))["n"]):
        f0, dt, lf = z[f"args{i}"]
        w = oracle.ricker(f0, dt, length_factor=lf)
        assert w.shape == z[f"wavelet{i}"].shape
        np.testing.assert_allclose(w, z[f"wavelet{i}"], rtol=0, atol=2e-6)


def test_oracle_gardner_known_values():
    """310 vp^0.25 (models/acoustic.py:169) at fourth powers: 7^4 = 2401 m/s ->
    2170, 6^4 = 1296 -> 1860, 8^4 = 4096 -> 2480, 10^4 -> 3100 (exact in
    float32); a wrong exponent or constant fails."""
    got = oracle.gardner(np.array([2401.0, 1296.0, 4096.0, 10000.0], np.float32))
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, np.array([2170.0, 1860.0, 2480.0, 3100.0], np.float32))
    vp = np.linspace(1500, 4700, 33, dtype=np.float32)
    assert np.all(np.diff(oracle.gardner(vp)) > 0)
