"""Synthetic workloads of the coupled step (host-side builders).

Each builder returns a StateSpec plus a metadata dict (sizes, step count, monitor
planes).  The workloads:

  small_coupled   64^3 uniform 20 nm, CPML 10, Si3N4 slab (eps 4, 8 cells) under a
                  2-cell Drude+Lorentz film with 4TM + two-sublattice LLB (m = +-0.9),
                  Ex Gaussian pulse (800 nm, 50 fs FWHM), 2000 steps; FP32.
  em_stencil      384^3 uniform, CPML 16, 64 seeded random dielectric boxes (sides 8-48
                  cells, eps 1-12) and a Drude-Lorentz sphere of radius 40 cells, plane
                  source, 5000 steps; EM only.
  coupled_pump    Nx x Ny x Nz with a z-graded axis (2 nm in the 20 nm film, 20 nm
                  elsewhere), Si3N4 waveguide 400 x 800 nm under the film, pump plane
                  source, 4 LLB substeps per EM step, gyrotropic eps_xy driven by m,
                  DFT monitors on 6 planes x 4 frequencies.  256x256x192 and 512x512x256.

The step is the Courant-0.5 CFL step on the smallest primary cell of each axis.
The film's magnetisation easy axis (the LLB x axis) is the film normal; for the
z-normal films here that is the physical z axis, and the gyration couples (Ex, Ey).
"""
from __future__ import annotations

import math

import numpy as np

from . import dispersion, geometry, grid as G, models, sources
from .spec import build_state_spec

CPML_KW = dict(order=3.0, reflection=1e-8, kappa_max=5.0, alpha_max=0.05)
LAMBDA_PUMP = 800e-9
FWHM_50FS_TAU = 50e-15 / math.sqrt(2.0 * math.log(2.0))   # exp(-(t/tau)^2) envelope, intensity FWHM 50 fs


def film_poles(dt: float, eps0: float = G.EPS0):
    """Drude (wp = 1.5e16 rad/s, gamma = 1e14 /s) + Lorentz (d_eps = 2, w0 = 4e15,
    gamma = 5e14) poles via create_pole (src/fdtd/Dispersion.jl:12-19)."""
    return [dispersion.create_pole(0.0, 1e14, (1.5e16) ** 2, dt, eps0),
            dispersion.create_pole(4e15, 5e14, 2.0 * (4e15) ** 2, dt, eps0)]


def plane_wave_gain(dt: float, dz: float) -> float:
    """Amplitude of the plane wave a soft plane source of unit amplitude launches on a
    uniform axis of spacing dz at step dt (1-D Yee, carrier 800 nm), measured on the host
    once so a requested field amplitude maps to a source amplitude."""
    w = 2 * math.pi * G.C0 / LAMBDA_PUMP
    S = G.C0 * dt / dz                          # cells travelled per step
    period = int(math.ceil(2 * math.pi / (w * dt)))
    dist = 10
    steps = int(math.ceil(dist / S)) + 3 * period
    n = int(2 * S * steps) + 2 * dist + 20      # boundaries are not reached during the run
    ez = np.zeros(n)
    hy = np.zeros(n - 1)
    ce = dt / (G.EPS0 * dz)
    ch = dt / (G.MU0 * dz)
    src, probe = n // 2, n // 2 + dist
    peak = 0.0
    for s in range(steps):
        hy += ch * (ez[1:] - ez[:-1])
        ez[1:-1] += ce * (hy[1:] - hy[:-1])
        ez[src] += math.sin(w * s * dt)
        if s >= steps - 2 * period:
            peak = max(peak, abs(ez[probe]))
    return peak


# Soft-source amplitude of the 64^3 case: the film's electrons peak near 1.7e3 K over the
# 2000 steps (1.49e3 K at 6.5e8 in the FP32 oracle).
SMALL_COUPLED_AMPLITUDE = 7.5e8


def small_coupled(dtype: str = "f32", steps: int = 2000, amplitude: float = SMALL_COUPLED_AMPLITUDE, n: int = 64,
                  absorption_model: str = "ade_work", source_kind: str = "plane"):
    nx = ny = nz = n
    d = 20e-9
    grid = G.Grid3D(G.uniform_axis_n(nx, d), G.uniform_axis_n(ny, d), G.uniform_axis_n(nz, d))
    dt = G.cfl_dt_cells(grid, courant=0.5)
    npml = 10 if n >= 32 else max(2, n // 8)
    k_slab = int(round(0.5625 * nz))           # slab top at 36/64
    z = grid.z.edges
    scene = geometry.Scene()
    scene.add(geometry.Box(-1, 1, -1, 1, z[k_slab - 8], z[k_slab]), geometry.Material("Si3N4", epsr=4.0))
    scene.add(geometry.Box(-1, 1, -1, 1, z[k_slab], z[k_slab + 2]),
              geometry.Material("GdFeCo", epsr=1.0, dispersive=True, model=True))
    geo = geometry.rasterize(scene, grid)
    pulse = sources.GaussianPulse(amplitude=amplitude, t0=0.5 * steps * dt, tau=FWHM_50FS_TAU,
                                  omega=2 * math.pi * G.C0 / LAMBDA_PUMP)
    if source_kind == "plane":
        src = sources.PlaneSource(pulse, "Ex", axis="z", position=float(grid.z.centers[k_slab - 14]))
    else:
        src = sources.SoftSource(pulse, "Ex", (max(1, npml + 1), (ny + 1) // 2 - 1, k_slab - 14))
    gd = models.ferrimagnet_parameters()
    spec = build_state_spec(grid, geo, dt, compute_precision=dtype, field_precision=dtype, source=src,
                            n_pml=npml, cpml_kwargs=CPML_KW, model=gd, enable_dispersion=True,
                            diag_poles=film_poles(dt), absorption_model=absorption_model,
                            multiphysics_every=1, subcycles=1)
    nf = spec.n_film
    CT = spec.compute_dtype
    spec.film_init["m_TM_x"][:] = CT.type(0.9)
    spec.film_init["m_RE_x"][:] = CT.type(-0.9)
    spec.film_init["m_TM_y"][:] = 0
    spec.film_init["m_RE_y"][:] = 0
    spec.film_init["m_TM_z"][:] = CT.type(1e-4)
    spec.film_init["m_RE_z"][:] = CT.type(1e-4)
    meta = dict(name=f"small_coupled_{n}^3", shape=grid.shape, steps=steps, n_film=nf, dt=dt)
    return spec, meta


# Storage scales of the reduced-precision form.  BF16 keeps FP32's exponent range, so it
# needs none.  FP16 (max 65504) needs power-of-two scales chosen from the magnitudes of
# the coupled pump at 60 mJ/cm^2: |psi_H| <= 2.4e19, |psi_D| <= 1.6e17, |P| <= 0.61 and
# |J dt/2| <= 1.3e-3 (J is stored in P units), leaving a factor ~2 of headroom.
REDUCED_SCALES = {"bf16": (1.0, 1.0), "f16": (2.0 ** -52, 2.0 ** 13)}


def film_graded_axis(n_cells: int, d_fine: float = 2e-9, n_fine: int = 10, d_max: float = 20e-9,
                     stretch: float = 1.25) -> G.Axis1D:
    """A graded axis of exactly n_cells cells centred on 0: n_fine cells of d_fine in the
    middle, geometric growth by `stretch` up to d_max outside (graded_axis semantics,
    src/grid/Grid.jl:77-95)."""
    if n_cells % 2 or n_fine % 2:
        raise ValueError("cell counts must be even")
    half = n_cells // 2
    e = [0.0]
    curr = 0.0
    for _ in range(n_fine // 2):
        curr += d_fine
        e.append(curr)
    cd = d_fine
    while len(e) - 1 < half:
        cd = min(cd * stretch, d_max)
        curr += cd
        e.append(curr)
    L = 2.0 * e[-1]
    ax = G.graded_axis(L, d_fine, (n_fine // 2) * d_fine, 0.0, d_max, stretch)
    if ax.n != n_cells:
        raise RuntimeError(f"graded axis produced {ax.n} cells, wanted {n_cells}")
    return ax


def coupled_pump(nx: int = 512, ny: int = 512, nz: int = 256, dtype: str = "f32", fluence_mj_cm2: float = 8.0,
                 steps: int = 20000, psi_precision=None, ade_state_precision=None, npml: int = 10,
                 n_planes: int = 6, nbins: int = 4, enable_mo: bool = True, subcycles: int = 4,
                 src_z: float = 300e-9):
    d = 20e-9
    grid = G.Grid3D(G.uniform_axis_n(nx, d, -0.5 * nx * d), G.uniform_axis_n(ny, d, -0.5 * ny * d),
                    film_graded_axis(nz))
    dt = G.cfl_dt_cells(grid, courant=0.5)
    zc = grid.z.centers
    scene = geometry.Scene()
    scene.add(geometry.Box(-1, 1, -400e-9, 400e-9, -410e-9, -10e-9), geometry.Material("Si3N4", epsr=4.0))
    scene.add(geometry.Box(-1, 1, -1, 1, -10e-9, 10e-9),
              geometry.Material("GdFeCo", epsr=1.0, dispersive=True, model=True))
    geo = geometry.rasterize(scene, grid)
    # fluence F = 1/2 c eps0 E0^2 tau sqrt(pi/2) for an exp(-(t/tau)^2) field envelope
    tau = 100e-15 / math.sqrt(2.0 * math.log(2.0))
    F = fluence_mj_cm2 * 10.0          # mJ/cm^2 -> J/m^2
    E0 = math.sqrt(F / (0.5 * G.C0 * G.EPS0 * tau * math.sqrt(math.pi / 2.0)))
    k_src = int(np.searchsorted(zc, src_z))
    gain = plane_wave_gain(dt, float(grid.z.widths()[k_src]))
    pulse = sources.GaussianPulse(amplitude=E0 / gain, t0=0.5 * steps * dt, tau=tau,
                                  omega=2 * math.pi * G.C0 / LAMBDA_PUMP)
    src = sources.PlaneSource(pulse, "Ex", axis="z", position=float(zc[k_src]))
    gd = models.ferrimagnet_parameters()
    psi_scale = REDUCED_SCALES.get(psi_precision, (1.0, 1.0))[0]
    ade_scale = REDUCED_SCALES.get(ade_state_precision, (1.0, 1.0))[1]
    spec = build_state_spec(grid, geo, dt, compute_precision=dtype, field_precision=dtype, source=src,
                            n_pml=npml, cpml_kwargs=CPML_KW, model=gd, enable_dispersion=True,
                            enable_magneto_optic=enable_mo, diag_poles=film_poles(dt),
                            absorption_model="ade_work", multiphysics_every=1, subcycles=subcycles,
                            psi_precision=psi_precision, psi_scale=psi_scale,
                            ade_state_precision=ade_state_precision, ade_state_scale=ade_scale,
                            mo_pair=("x", "y"))
    CT = spec.compute_dtype
    for k, v in (("m_TM_x", 0.9), ("m_RE_x", -0.9), ("m_TM_y", 0.0), ("m_RE_y", 0.0),
                 ("m_TM_z", 1e-4), ("m_RE_z", 1e-4)):
        spec.film_init[k][:] = CT.type(v)
    planes = []
    if n_planes:
        zs = [int(np.searchsorted(zc, 200e-9)), int(np.searchsorted(zc, -600e-9))]
        cand = [(0, int(0.3 * nx)), (0, int(0.7 * nx)), (1, int(0.3 * ny)), (1, int(0.7 * ny)),
                (2, zs[0]), (2, zs[1])]
        planes = cand[:n_planes]
    wp = 2 * math.pi * G.C0 / LAMBDA_PUMP
    bins = [wp * f for f in (0.96, 0.99, 1.01, 1.04)][:nbins]
    meta = dict(name=f"coupled_pump_{nx}x{ny}x{nz}", shape=grid.shape, steps=steps, n_film=spec.n_film,
                dt=dt, planes=planes, omega_bins=bins, E0=E0, fluence_mj_cm2=fluence_mj_cm2)
    return spec, meta


def em_stencil(n: int = 384, npml: int = 16, steps: int = 5000, seed: int = 0, n_boxes: int = 64,
               sphere_radius: int = 40, dtype: str = "f32"):
    d = 20e-9
    grid = G.Grid3D(G.uniform_axis_n(n, d), G.uniform_axis_n(n, d), G.uniform_axis_n(n, d))
    dt = G.cfl_dt_cells(grid, courant=0.5)
    rng = np.random.default_rng(seed)
    e = grid.x.edges
    scene = geometry.Scene()
    lo, hi = npml, n - npml
    for b in range(n_boxes):
        s = rng.integers(8, 49, size=3)
        p0 = [int(rng.integers(lo, max(lo + 1, hi - s[a]))) for a in range(3)]
        p1 = [min(hi, p0[a] + int(s[a])) for a in range(3)]
        # centres of cells p0..p1-1 lie strictly inside (e[p0], e[p1])
        box = geometry.Box(e[p0[0]], e[p1[0]] - 0.5 * d, e[p0[1]], e[p1[1]] - 0.5 * d, e[p0[2]], e[p1[2]])
        scene.add(box, geometry.Material(f"box{b}", epsr=float(rng.uniform(1.0, 12.0))))
    c = 0.5 * n * d
    scene.add(geometry.Sphere(c, c, c, sphere_radius * d),
              geometry.Material("DL", epsr=1.0, dispersive=True))
    geo = geometry.rasterize(scene, grid)
    pulse = sources.GaussianPulse(amplitude=1.0, t0=20e-15, tau=FWHM_50FS_TAU / 4,
                                  omega=2 * math.pi * G.C0 / LAMBDA_PUMP)
    src = sources.PlaneSource(pulse, "Ex", axis="z", position=float(grid.z.centers[npml + 2]))
    spec = build_state_spec(grid, geo, dt, compute_precision=dtype, field_precision=dtype, source=src,
                            n_pml=npml, cpml_kwargs=CPML_KW, model=None, enable_dispersion=True,
                            diag_poles=film_poles(dt), dispersive_cells="poles")
    meta = dict(name=f"em_stencil_{n}^3", shape=grid.shape, steps=steps, n_ade=len(geo.pole_cells), dt=dt)
    return spec, meta
