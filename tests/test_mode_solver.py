"""Scalar waveguide mode solver (host, one-time) against the reference's test and against
closed forms (CPU).

runtests.jl:427-435 gives the reference's known answer on a 9 x 9 step-index guide.  The
operator is pinned by a loop-built dense matrix; the inverse iteration by the textbook
eigenpair of the 5-point Laplacian with zero boundary (uniform core: beta^2 = k0^2 eps -
4/dy^2 sin^2(pi/(2(ny+1))) - 4/dz^2 sin^2(pi/(2(nz+1))), mode sin x sin) and by dense
symmetric eigensolves of the 9 x 9 case and of a guided 31 x 31 case.
"""
import math
import warnings

import numpy as np
import pytest


def _step_index(n=9, core=slice(3, 6)):
    e = np.full((n, n), 1.44 ** 2)
    e[core, core] = 2.0 ** 2
    return e


def test_reference_known_answer():
    """runtests.jl:427-435.  On this 9 x 9 section the zero-field walls sit 150 nm from the
    core, and the fundamental mode has neff = 1.2399 (dense eigensolve below), under the
    cladding index: the solver takes its Gaussian fallback with neff = clamp(1.8, 1.44, 2),
    and the reference's assertions hold on that path."""
    from mphip.modesolver import solve_waveguide_mode
    dy = dz = 50e-9
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sol = solve_waveguide_mode(_step_index(), dy, dz, 800e-9, neff_guess=1.8, max_iter=8)
    assert 1.44 <= sol["neff"] <= 2.0
    assert sol["profile"].shape == (9, 9)
    assert float(np.sum(np.abs(sol["profile"]) ** 2)) * dy * dz == pytest.approx(1.0, rel=1e-6)
    assert sol["fallback"] and sol["neff"] == 1.8


def test_reference_case_top_eigenvalue_is_below_cladding():
    from mphip.modesolver import inverse_iteration, mode_operator
    from mphip.sources import gaussian_mode_profile
    k0 = 2 * math.pi / 800e-9
    A = mode_operator(_step_index(), 50e-9, 50e-9, k0)
    w = np.linalg.eigvalsh(A.toarray())
    assert math.sqrt(w[-1]) / k0 == pytest.approx(1.2399156869, rel=1e-9)
    b2, _ = inverse_iteration(A, (k0 * 1.8) ** 2, gaussian_mode_profile(9, 9, 50e-9, 50e-9).reshape(-1, order="F"),
                              max_iter=100, tol=1e-15)
    assert b2 == pytest.approx(w[-1], rel=1e-12)


def test_operator_matches_loop_construction():
    from mphip.modesolver import mode_operator
    rng = np.random.default_rng(0)
    ny, nz, dy, dz, k0 = 4, 3, 40e-9, 55e-9, 2 * math.pi / 700e-9
    e = rng.uniform(1, 4, (ny, nz))
    A = mode_operator(e, dy, dz, k0).toarray()
    D = np.zeros((ny * nz, ny * nz))
    idx = lambda j, k: j + k * ny
    for k in range(nz):
        for j in range(ny):
            p = idx(j, k)
            D[p, p] = k0 ** 2 * e[j, k] - 2 * (1 / dy ** 2 + 1 / dz ** 2)
            if j > 0:
                D[p, idx(j - 1, k)] = 1 / dy ** 2
            if j < ny - 1:
                D[p, idx(j + 1, k)] = 1 / dy ** 2
            if k > 0:
                D[p, idx(j, k - 1)] = 1 / dz ** 2
            if k < nz - 1:
                D[p, idx(j, k + 1)] = 1 / dz ** 2
    np.testing.assert_array_equal(A, D)


def test_inverse_iteration_closed_form():
    from mphip.modesolver import inverse_iteration, mode_operator
    ny, nz, dy, dz = 12, 9, 30e-9, 45e-9
    k0 = 2 * math.pi / 1.0e-6
    eps = 2.25
    A = mode_operator(np.full((ny, nz), eps), dy, dz, k0)
    b_exact = (k0 ** 2 * eps - 4 / dy ** 2 * math.sin(math.pi / (2 * (ny + 1))) ** 2
               - 4 / dz ** 2 * math.sin(math.pi / (2 * (nz + 1))) ** 2)
    b2, v = inverse_iteration(A, k0 ** 2 * eps, np.ones(ny * nz), max_iter=200, tol=1e-14)
    assert b2 == pytest.approx(b_exact, rel=1e-12)
    mode = np.outer(np.sin(math.pi * np.arange(1, ny + 1) / (ny + 1)), np.sin(math.pi * np.arange(1, nz + 1) / (nz + 1)))
    mode = mode.reshape(-1, order="F")
    cos = abs(np.dot(v, mode)) / (np.linalg.norm(v) * np.linalg.norm(mode))
    assert cos == pytest.approx(1.0, abs=1e-9)


def test_guided_step_index_matches_dense_eigensolve():
    """A 7 x 7-cell core (n = 2) in a 31 x 31 cladding (n = 1.44) at 800 nm is guided:
    neff = 1.72357, between the two indices."""
    from mphip.modesolver import mode_operator, solve_waveguide_mode
    dy = dz = 50e-9
    lam = 800e-9
    e = _step_index(31, slice(12, 19))
    sol = solve_waveguide_mode(e, dy, dz, lam, max_iter=200, tol=1e-15)
    assert not sol["fallback"] and 1.44 < sol["neff"] < 2.0
    k0 = 2 * math.pi / lam
    w, V = np.linalg.eigh(mode_operator(e, dy, dz, k0).toarray())
    assert sol["neff"] == pytest.approx(math.sqrt(w[-1]) / k0, rel=1e-12)
    p = sol["profile"].reshape(-1, order="F")
    assert abs(np.dot(p, V[:, -1])) / np.linalg.norm(p) == pytest.approx(1.0, abs=1e-9)
    # the fundamental mode is concentrated in the core and symmetric about the centre
    prof = np.abs(sol["profile"])
    assert np.unravel_index(np.argmax(prof), prof.shape) == (15, 15)
    np.testing.assert_allclose(prof, prof[::-1, ::-1], rtol=1e-8, atol=1e-8 * prof.max())
    # default settings (20 iterations, tol 1e-9) agree with the converged mode
    assert solve_waveguide_mode(e, dy, dz, lam)["neff"] == pytest.approx(sol["neff"], rel=1e-9)


def test_fallback_outside_index_bounds():
    """A uniform section has neff < n: outside [n_min, n_max], so the Gaussian is used."""
    from mphip.modesolver import solve_waveguide_mode
    from mphip.sources import gaussian_mode_profile
    e = np.full((8, 6), 2.0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        sol = solve_waveguide_mode(e, 40e-9, 40e-9, 600e-9, neff_guess=1.9)
    assert sol["fallback"] and any("fallback" in str(x.message) for x in w)
    assert sol["neff"] == pytest.approx(math.sqrt(2.0))       # clamp(1.9, sqrt 2, sqrt 2)
    np.testing.assert_array_equal(sol["profile"], gaussian_mode_profile(8, 6, 40e-9, 40e-9))


def test_mode_source_from_permittivity_volume():
    from mphip import grid, sources
    from mphip.modesolver import mode_source, mode_spacings
    g = grid.Grid3D(grid.uniform_axis_n(10, 40e-9), grid.uniform_axis_n(31, 40e-9),
                    grid.uniform_axis_n(27, 50e-9))
    e = np.full((10, 31, 27), 1.44 ** 2)
    e[:, 11:20, 10:17] = 2.0 ** 2
    pulse = sources.GaussianPulse(amplitude=1.0, t0=10e-15, tau=3e-15, omega=2 * math.pi * grid.C0 / 800e-9)
    src = mode_source(g, pulse, 3, 800e-9, epsr_volume=e, component="Ey")
    assert isinstance(src, sources.ModeSource) and src.component == "Ey" and src.axis == "x"
    assert 1.44 < src.neff < 2.0
    assert src.profile.shape == (31, 27)
    j, k = np.unravel_index(np.argmax(np.abs(src.profile)), src.profile.shape)
    assert (j, k) == (15, 13)
    assert mode_spacings(g, "x") == pytest.approx((40e-9, 50e-9), rel=1e-12)
    gsrc = mode_source(g, pulse, 3, 800e-9)
    assert gsrc.neff == 1.0
    dy, dz = mode_spacings(g, "x")      # the grid's cell widths, to rounding
    np.testing.assert_array_equal(gsrc.profile, sources.gaussian_mode_profile(31, 27, dy, dz))
