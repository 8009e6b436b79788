"""Helpers shared by the tests: golden-fixture loading and error metrics."""

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

FWD_CASES = ["tiny0_fwd", "tiny1_fwd", "tiny2_fwd", "tiny3_fwd", "tiny_snap", "c1_fwd"]
GRAD_CASES = ["tiny0_grad", "tiny1_grad", "tiny2_grad", "tiny3_grad", "c1_grad", "c2r_grad"]
# several source points per shot ((S, K) positions): per-shot and per-point wavelets
MP_FWD_CASES = ["mp0_fwd", "mp1_fwd"]
MP_GRAD_CASES = ["mp0_grad", "mp1_grad"]
OUT_KEYS = ("p", "vx", "vz", "forward_wavefield_p", "forward_wavefield_vx",
            "forward_wavefield_vz", "snapshots")


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    return z, meta


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def chunk_len(nt, segments):
    return -(-nt // max(1, int(segments)))


def oracle_args(z, m):
    return dict(nz=m["nz"], nx=m["nx"], nabc=m["nabc"], free_surface=m["free_surface"],
                dt=m["dt"], dz=m["dx"], src_z=z["in_src_z"], src_x=z["in_src_x"],
                src_v=z["in_wavelet"], rcv_z=z["in_rcv_z"], rcv_x=z["in_rcv_x"],
                vp=z["in_vp"], rho=z["in_rho"], damp=z["in_damp"])


def adjoint_sources(z, m, fwd_p=None, obs_p=None):
    """The loss behind a gradient fixture, as adjoint sources (g_p, g_vx, g_vz)."""
    S, R, nt = len(z["in_src_z"]), len(z["in_rcv_z"]), m["nt"]
    if "g_seed" in m:
        rng = np.random.default_rng(m["g_seed"])
        return tuple(rng.standard_normal((S, nt, R)).astype(np.float32) for _ in range(3))
    return (2.0 * (fwd_p - obs_p), None, None)


FWD3D_CASES = ["t3d0_fwd", "t3d1_fwd", "t3d2_fwd"]
GRAD3D_CASES = ["t3d0_grad", "t3d1_grad", "t3d2_grad"]
MP3D_FWD_CASES = ["t3dmp0_fwd", "t3dmp1_fwd"]   # (S, K) source positions
MP3D_GRAD_CASES = ["t3dmp0_grad", "t3dmp1_grad"]
OUT3D_KEYS = ("p", "vx", "vy", "vz", "forward_wavefield_p", "forward_wavefield_vx",
              "forward_wavefield_vy", "forward_wavefield_vz")


def fd_c1(name="vp", cells=16, step=1e-2, seed=16):
    """Central finite differences of L = sum(g . seismograms) w.r.t. the model
    parameter ``name`` ("vp" or "rho") on the config-1 problem (100 x 200, one
    shot, 1000 steps; the c1_grad fixture's inputs) with the FP64 oracle:
    ``cells`` seeded cells, parameter +- ``step`` (rounded to the float32 model
    values the operator reads, and divided by the difference of those).
    Returns (fixture, meta, adjoint sources (g_p, g_vx, g_vz), cell rows, cell
    columns, finite differences)."""
    import oracle
    z, m = load("c1_grad")
    a = oracle_args(z, m)
    rng = np.random.default_rng(seed)
    R, nt = len(z["in_rcv_z"]), m["nt"]
    gs = [rng.standard_normal((1, nt, R)) for _ in range(3)]
    nz, nx = a[name].shape
    ii, jj = rng.integers(0, nz, cells), rng.integers(0, nx, cells)

    def loss(model):
        out = oracle.forward(**dict(a, **{name: model}), precision="f64")
        return sum(float(np.sum(out[k] * g)) for k, g in zip(("p", "vx", "vz"), gs))

    fd = np.empty(cells)
    for c, (i, j) in enumerate(zip(ii, jj)):
        up, dn = a[name].copy(), a[name].copy()
        up[i, j] = np.float32(a[name][i, j] + step)
        dn[i, j] = np.float32(a[name][i, j] - step)
        fd[c] = (loss(up) - loss(dn)) / (float(up[i, j]) - float(dn[i, j]))
    return z, m, gs, ii, jj, fd


def oracle3d_args(z, m):
    return dict(nz=m["nz"], ny=m["ny"], nx=m["nx"], nabc=m["nabc"], free_surface=m["free_surface"],
                dt=m["dt"], dz=m["dx"], src_z=z["in_src_z"], src_y=z["in_src_y"],
                src_x=z["in_src_x"], src_v=z["in_wavelet"], rcv_z=z["in_rcv_z"],
                rcv_y=z["in_rcv_y"], rcv_x=z["in_rcv_x"], vp=z["in_vp"], rho=z["in_rho"],
                damp=z["in_damp"])


def adjoint_sources_3d(z, m):
    """The seeded adjoint sources (g_p, g_vx, g_vy, g_vz) of a 3D gradient fixture."""
    S, R, nt = len(z["in_src_z"]), len(z["in_rcv_z"]), m["nt"]
    rng = np.random.default_rng(m["g_seed"])
    return tuple(rng.standard_normal((S, nt, R)).astype(np.float32) for _ in range(4))
