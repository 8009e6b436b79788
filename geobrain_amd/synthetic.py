"""Seeded synthetic inputs for the propagator's workloads.

Every generator is deterministic in its seed (numpy ``default_rng``) and runs
on the host; results are float32 numpy arrays (velocities in m/s).

* ``layered_model``: horizontally layered velocity, interface depths drawn
  without replacement from ``[1, nz)``, layer velocities uniform in a range.
* ``grf_model``: velocity increasing linearly with depth plus a Gaussian random
  field with Gaussian covariance ``exp(-(dx/lh)^2 - (dz/lv)^2)`` (anisotropic,
  lateral length ``lh``, vertical ``lv``) scaled to a given standard
  deviation, clipped to a velocity range.  The field is synthesised by
  filtering white noise in the Fourier domain on a grid padded by three
  correlation lengths, so it does not wrap around.
* ``grf_model_3d``: the same construction on a ``(nz, ny, nx)`` grid, the
  lateral length ``lh`` applying to both horizontal axes.
* ``smooth_model``: Gaussian smoothing with sigma a fraction of ``min(nz, nx)``.
* ``ricker_padded``: the propagator's Ricker wavelet zero-padded to ``nt``.
* ``gardner``: density ``310 vp^0.25`` as ``AcousticModel`` computes it.
* ``uneven_batch``: the seed-7 list of independent problems of uneven size.
"""

from __future__ import annotations

import numpy as np


def layered_model(nz, nx, n_layers=5, vmin=1500.0, vmax=3500.0, seed=0):
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.choice(np.arange(1, nz), size=n_layers - 1, replace=False))
    vels = rng.uniform(vmin, vmax, n_layers)
    vp = np.empty((nz, nx), np.float32)
    edges = [0, *depths.tolist(), nz]
    for k in range(n_layers):
        vp[edges[k]:edges[k + 1], :] = vels[k]
    return vp


def grf_model(nz, nx, dx, seed=0, v_top=1500.0, v_bottom=4500.0, lh=300.0, lv=50.0,
              std=200.0, vclip=(1500.0, 4700.0)):
    rng = np.random.default_rng(seed)
    pz = nz + 2 * int(np.ceil(3 * lv / dx))
    px = nx + 2 * int(np.ceil(3 * lh / dx))
    noise = rng.standard_normal((pz, px))
    kz = np.fft.fftfreq(pz, d=dx)
    kx = np.fft.fftfreq(px, d=dx)
    # spectrum of exp(-(x/lh)^2 - (z/lv)^2) is proportional to exp(-pi^2 (kx^2 lh^2 + kz^2 lv^2))
    spec = np.exp(-np.pi ** 2 * (kz[:, None] ** 2 * lv ** 2 + kx[None, :] ** 2 * lh ** 2))
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(spec)))
    oz, ox = (pz - nz) // 2, (px - nx) // 2
    field = field[oz:oz + nz, ox:ox + nx]
    field = (field - field.mean()) / field.std() * std
    depth = np.linspace(0.0, 1.0, nz)[:, None]
    vp = v_top + (v_bottom - v_top) * depth + field
    return np.clip(vp, vclip[0], vclip[1]).astype(np.float32)


def grf_model_3d(nz, ny, nx, dx, seed=0, v_top=1500.0, v_bottom=4500.0, lh=200.0, lv=50.0,
                 std=200.0, vclip=(1500.0, 4700.0)):
    rng = np.random.default_rng(seed)
    ez, eh = int(np.ceil(3 * lv / dx)), int(np.ceil(3 * lh / dx))
    pz, py, px = nz + 2 * ez, ny + 2 * eh, nx + 2 * eh
    noise = rng.standard_normal((pz, py, px))
    kz = np.fft.fftfreq(pz, d=dx)[:, None, None]
    ky = np.fft.fftfreq(py, d=dx)[None, :, None]
    kx = np.fft.rfftfreq(px, d=dx)[None, None, :]
    spec = np.exp(-np.pi ** 2 * (kz ** 2 * lv ** 2 + (ky ** 2 + kx ** 2) * lh ** 2))
    field = np.fft.irfftn(np.fft.rfftn(noise) * np.sqrt(spec), s=(pz, py, px))
    field = field[ez:ez + nz, eh:eh + ny, eh:eh + nx]
    field = (field - field.mean()) / field.std() * std
    depth = np.linspace(0.0, 1.0, nz)[:, None, None]
    vp = v_top + (v_bottom - v_top) * depth + field
    return np.clip(vp, vclip[0], vclip[1]).astype(np.float32)


def smooth_model(vp, frac=0.05):
    from scipy.ndimage import gaussian_filter
    sigma = frac * min(vp.shape)
    return gaussian_filter(np.asarray(vp, np.float64), sigma=sigma, mode="nearest").astype(np.float32)


def ricker_padded(f0, dt, nt):
    from .physics.wave.wavelets import RickerWavelet
    w, _ = RickerWavelet()(f0=f0, dt=dt)
    w = w.numpy()
    out = np.zeros(nt, np.float32)
    n = min(nt, w.shape[0])
    out[:n] = w[:n]
    return out


def gardner(vp):
    return (310.0 * np.power(np.asarray(vp, np.float32), 0.25)).astype(np.float32)


def surface_shots(nx, n_shots):
    """Shot x positions evenly spaced along the surface: round((i + 1/2) nx / n)."""
    return np.array([int(round((i + 0.5) * nx / n_shots)) for i in range(n_shots)], np.int64)


def uneven_batch(n_problems=96, seed=7, dx=5.0, dt=5e-4):
    """Independent problems: nz in [200, 600], nx in [200, 1200], nt in [1000, 4000],
    1-8 shots at f0 in [8, 20] Hz, layered-Gaussian-random velocity."""
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(n_problems):
        nz = int(rng.integers(200, 601))
        nx = int(rng.integers(200, 1201))
        nt = int(rng.integers(1000, 4001))
        ns = int(rng.integers(1, 9))
        f0 = float(rng.uniform(8.0, 20.0))
        probs.append(dict(nz=nz, nx=nx, nt=nt, n_shots=ns, f0=f0, dx=dx, dt=dt,
                          seed=int(rng.integers(0, 2 ** 31 - 1))))
    return probs
