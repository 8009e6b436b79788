"""Absorbing-boundary damping arrays.

Mirrors ``geobrain/physics/wave/boundary/pml.py:16-98`` (``bc_pml``),
``pml.py:101-195`` (``bc_pml_3d``) and the dispatch of
``boundary/__init__.py:23-116`` (``create_boundary``).  Only the PML type feeds
the acoustic propagators; the reference's Gerjan/sin-cos weight arrays and the
split PML belong to other solvers and are not provided.
"""

from __future__ import annotations

import math

import numpy as np


def bc_pml(nx: int, nz: int, dx: float, dz: float, pml: int, vmax: float,
           free_surface: bool = True) -> np.ndarray:
    """Quadratic PML damping, float64 array of shape ``(NZ, NX)``.

    ``d(i) = kappa (i dx / a)^2`` for ``i = 0..pml-1`` counted outward, with
    ``a = (pml - 1) dx`` and ``kappa = -3 vmax ln(1e-6) / (2 a)``.  The profile
    fills the left/right columns over every row, then the bottom (and, without a
    free surface, the top) rows over a trapezoid that narrows by one column per
    layer; ``NZ = nz + pml`` with a free surface, ``nz + 2 pml`` without.
    """
    NX = nx + 2 * pml
    NZ = nz + pml if free_surface else nz + 2 * pml
    a = (pml - 1) * dx
    kappa = -3.0 * vmax * math.log(1e-6) / (2.0 * a)
    xa = np.arange(pml) * dx / a
    prof = kappa * xa * xa
    d = np.zeros((NZ, NX))
    d[:, pml - 1 - np.arange(pml)] = prof
    d[:, NX - pml + np.arange(pml)] = prof
    for i in range(pml):
        x0, x1 = pml - i, NX - (pml - i)
        if not free_surface:
            d[pml - 1 - i, x0:x1] = prof[i]
        d[NZ - pml + i, x0:x1] = prof[i]
    return d


def _pml_profile(pml: int, dx: float, vmax: float) -> np.ndarray:
    a = (pml - 1) * dx
    kappa = -3.0 * vmax * math.log(1e-6) / (2.0 * a)
    xa = np.arange(pml) * dx / a
    return kappa * xa * xa


def bc_pml_3d(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, pml: int,
              vmax: float, free_surface: bool = True) -> np.ndarray:
    """Quadratic PML damping on the six faces, float64 ``(NZ, NY, NX)``.

    Layer ``i`` (counted outward) of each face carries ``d(i)``, the profile of
    ``bc_pml`` with spacing ``dx`` on every axis; where faces meet, a cell takes
    the largest of its layers' values.  The top face is damped only without a
    free surface; ``NZ = nz + pml`` with one, ``nz + 2 pml`` without.
    """
    NX, NY = nx + 2 * pml, ny + 2 * pml
    NZ = nz + pml if free_surface else nz + 2 * pml
    prof = _pml_profile(pml, dx, vmax)

    def axis(n: int, top: bool) -> np.ndarray:
        d = np.zeros(n)
        if top:
            d[pml - 1 - np.arange(pml)] = prof
        d[n - pml + np.arange(pml)] = prof
        return d

    dz_ = axis(NZ, not free_surface)[:, None, None]
    dy_ = axis(NY, True)[None, :, None]
    dx_ = axis(NX, True)[None, None, :]
    return np.maximum(np.maximum(dx_, dy_), dz_)


def create_boundary(bc_type: str, nx: int, nz: int, dx: float, dz: float, n_layers: int,
                    vmax: float = None, alpha: float = 0.0053, free_surface: bool = False,
                    ny: int = None, dy: float = None) -> np.ndarray:
    """Damping array for the propagator (``boundary/__init__.py:23-116``)."""
    bc_type = bc_type.lower()
    if ny is not None:
        if bc_type != 'pml':
            raise NotImplementedError(
                f"3-D boundary type '{bc_type}' is not yet supported. "
                f"Only 'pml' is available for 3-D.")
        if vmax is None:
            raise ValueError("vmax is required for PML boundary")
        if dy is None:
            raise ValueError("dy is required for 3-D PML boundary")
        return bc_pml_3d(nx, ny, nz, dx, dy, dz, n_layers, vmax, free_surface)
    if bc_type == 'pml':
        if vmax is None:
            raise ValueError("vmax is required for PML boundary")
        return bc_pml(nx, nz, dx, dz, n_layers, vmax, free_surface)
    if bc_type in ('pml_xz', 'gerjan', 'sincos'):
        raise NotImplementedError(
            f"boundary type '{bc_type}' is not part of the 2-D acoustic PML path")
    raise ValueError(f"Unknown boundary type: '{bc_type}'. "
                     f"Choose from: 'pml', 'pml_xz', 'gerjan', 'sincos'")
