"""Scalar waveguide mode solver: the one-time host solve behind the mode source.

Restates src/fdtd/ModeSolver.jl:21-88 (`_mode_operator`, `solve_waveguide_mode`) and the
mode-source construction of src/sim/Experiment.jl:111-165 (`_epsr_mode_plane`,
`_mode_spacings`, the `:mode` branch of `_source_from_config`).  The operator is the
5-point scalar Helmholtz operator k0^2 eps_r + d2/dy2 + d2/dz2 on the cross-section with
the field zero outside it; the fundamental mode is found by shifted inverse iteration
from a Gaussian seed with the shift (k0 n_max)^2 just above the top of the spectrum.  A
solve that fails or returns neff outside [n_min, n_max] falls back to the Gaussian
profile with neff = clamp(guess, n_min, n_max), with a warning, as the reference does.

The result is uploaded once (ModeSource profile); nothing here runs per step.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from .sources import ModeSource, gaussian_mode_profile, source_index


def mode_operator(epsr_plane, dy: float, dz: float, k0: float):
    """_mode_operator (ModeSolver.jl:21-48) as a sparse CSC matrix over the (ny, nz)
    cross-section, column-major cell order (j fastest)."""
    import scipy.sparse as sp
    e = np.asarray(epsr_plane, np.float64)
    ny, nz = e.shape
    N = ny * nz
    idy2, idz2 = 1.0 / (float(dy) ** 2), 1.0 / (float(dz) ** 2)
    p = np.arange(N).reshape((ny, nz), order="F")
    rows = [p.ravel(order="F")]
    cols = [p.ravel(order="F")]
    vals = [(float(k0) ** 2 * e - 2.0 * (idy2 + idz2)).ravel(order="F")]
    for sl_a, sl_b, w in (((slice(1, None), slice(None)), (slice(None, -1), slice(None)), idy2),
                          ((slice(None, -1), slice(None)), (slice(1, None), slice(None)), idy2),
                          ((slice(None), slice(1, None)), (slice(None), slice(None, -1)), idz2),
                          ((slice(None), slice(None, -1)), (slice(None), slice(1, None)), idz2)):
        a, b = p[sl_a].ravel(order="F"), p[sl_b].ravel(order="F")
        rows.append(a)
        cols.append(b)
        vals.append(np.full(a.size, w))
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def inverse_iteration(A, sigma: float, v0, max_iter: int = 20, tol: float = 1e-9):
    """Shifted inverse iteration (ModeSolver.jl:67-77): v <- (A - sigma I)^-1 v, 2-norm
    normalised, Rayleigh quotient beta^2 = v.Av / v.v until it changes by at most
    tol * max(|beta^2|, 1).  Returns (beta^2, v)."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import splu
    lu = splu((A - sigma * sp.identity(A.shape[0], format="csc")).tocsc())
    v = np.asarray(v0, np.float64).copy()
    beta2 = last = float(sigma)
    for _ in range(max(1, int(max_iter))):
        v = lu.solve(v)
        v /= max(float(np.linalg.norm(v)), np.finfo(np.float64).eps)
        beta2 = float(np.dot(v, A @ v) / np.dot(v, v))
        if abs(beta2 - last) <= float(tol) * max(abs(beta2), 1.0):
            break
        last = beta2
    return beta2, v


def _normalize_mode(profile, dy: float, dz: float) -> np.ndarray:
    """_normalize_mode (ModeSolver.jl:15-19): sum(p^2) dy dz = 1."""
    norm = math.sqrt(float(np.sum(np.abs(profile) ** 2)) * float(dy) * float(dz))
    return profile / norm if norm > 0 else profile


def solve_waveguide_mode(epsr_plane, dy: float, dz: float, lambda0: float, neff_guess=None, max_iter: int = 20,
                         tol: float = 1e-9) -> dict:
    """Fundamental scalar mode of an (ny, nz) relative-permittivity cross-section:
    dict(profile=(ny, nz), neff=..., fallback=bool) (solve_waveguide_mode,
    ModeSolver.jl:50-88)."""
    e = np.asarray(epsr_plane, np.float64)
    ny, nz = e.shape
    k0 = 2.0 * math.pi / float(lambda0)
    n_min = math.sqrt(max(float(e.min()), np.finfo(np.float64).eps))
    n_max = math.sqrt(float(e.max()))
    guess = n_max if neff_guess is None else float(neff_guess)
    sigma = (k0 * guess) ** 2
    seed = gaussian_mode_profile(ny, nz, dy, dz).reshape(-1, order="F")

    def fallback():
        return dict(profile=gaussian_mode_profile(ny, nz, dy, dz), neff=min(max(guess, n_min), n_max),
                    fallback=True)

    try:
        beta2, v = inverse_iteration(mode_operator(e, dy, dz, k0), sigma, seed, max_iter, tol)
    except Exception as err:  # singular shift or a failed factorisation
        warnings.warn(f"mode solve failed ({err}); using Gaussian fallback")
        return fallback()
    neff = math.sqrt(max(beta2, 0.0)) / k0
    profile = _normalize_mode(v.reshape((ny, nz), order="F"), dy, dz)
    if not (n_min <= neff <= n_max):
        warnings.warn(f"mode solve returned neff={neff} outside [{n_min}, {n_max}]; using Gaussian fallback")
        return fallback()
    return dict(profile=profile, neff=neff, fallback=False)


def epsr_mode_plane(epsr_volume, grid, axis: str, position) -> np.ndarray:
    """_epsr_mode_plane (Experiment.jl:111-124): the cross-section at the source plane."""
    e = np.asarray(epsr_volume)
    if axis == "x":
        return e[source_index(grid.x, position), :, :]
    if axis == "y":
        return e[:, source_index(grid.y, position), :]
    if axis == "z":
        return e[:, :, source_index(grid.z, position)]
    raise ValueError("mode source axis must be 'x', 'y' or 'z'")


def mode_spacings(grid, axis: str):
    """_mode_spacings (Experiment.jl:126-135): the smallest cell widths of the two
    transverse axes."""
    w = {"x": grid.x.widths(), "y": grid.y.widths(), "z": grid.z.widths()}
    t = {"x": ("y", "z"), "y": ("x", "z"), "z": ("x", "y")}.get(axis)
    if t is None:
        raise ValueError("mode source axis must be 'x', 'y' or 'z'")
    return float(np.min(w[t[0]])), float(np.min(w[t[1]]))


def mode_source(grid, pulse, position, lambda0: float, epsr_volume=None, axis: str = "x", component: str = "Ez",
                neff_guess=None, max_iter: int = 20) -> ModeSource:
    """The `:mode` branch of _source_from_config (Experiment.jl:149-165): the eigenmode of
    the permittivity cross-section at `position`, or, without a permittivity volume, the
    Gaussian profile with neff = neff_guess (1 when absent)."""
    dy, dz = mode_spacings(grid, axis)
    if epsr_volume is None:
        prof = gaussian_mode_profile(grid.y.n, grid.z.n, dy, dz)
        neff = 1.0 if neff_guess is None else float(neff_guess)
    else:
        sol = solve_waveguide_mode(epsr_mode_plane(epsr_volume, grid, axis, position), dy, dz, lambda0,
                                   neff_guess=neff_guess, max_iter=max_iter)
        prof, neff = sol["profile"], sol["neff"]
    return ModeSource(prof, pulse, neff=neff, axis=axis, position=position, component=component)
