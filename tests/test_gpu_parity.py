"""GPU parity: libmphip against the C oracle on identical seeded inputs.

FP32 Yee / CPML / ADE / pabs arithmetic is written in the reference's operation order on
both sides, so field volumes must agree bit for bit.  The 4TM + LLB integrator runs FP64
with transcendental calls (exp) whose last-ULP results differ between glibc and the GPU
math library, so film state is compared to a relative tolerance; the bar is the one the
coupled path states: fields within 1e-5 relative L2 and sublattice magnetisation within
1e-4 after N steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a, b = a.astype(np.float64), b.astype(np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def run_pair(spec, steps, monitors=None):
    import driver
    from mphip.state import FDTDState
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    for _ in range(steps):
        gpu.step()
        ref.step()
    gpu.synchronize()
    return gpu, ref


@pytest.mark.parametrize("n,steps", [(24, 150), (40, 300)])
def test_small_coupled_fields_bitwise_f32(n, steps):
    from mphip import configs
    spec, meta = configs.small_coupled(dtype="f32", steps=steps, amplitude=4e8, n=n)
    print(f"case small_coupled n={n} steps={steps}", flush=True)
    gpu, ref = run_pair(spec, steps)
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), f"{name}: max |diff| {np.abs(a - b).max()}"
    fa, fb = gpu.film_state(), ref.film_state()
    assert fb["Te"].max() > 300.5, "the pump must heat the film in this case"
    for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
        assert rel_l2(fa[k], fb[k]) < 1e-6, k
    for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k
    for k in ("U_abs", "window_sum", "peak"):
        assert rel_l2(fa[k], fb[k]) < 1e-12, k
    for c in range(3):
        ga, ra = gpu.ade_state(c), ref.ade_arr[c]
        assert np.array_equal(ga["P"], ra["P"]) and np.array_equal(ga["J"], ra["J"])
    gpu.free()


def test_small_coupled_f64():
    from mphip import configs
    spec, meta = configs.small_coupled(dtype="f64", steps=120, amplitude=4e8, n=24)
    print("case small_coupled f64 n=24", flush=True)
    gpu, ref = run_pair(spec, 120)
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert rel_l2(a, b) < 1e-13, name
    fa, fb = gpu.film_state(), ref.film_state()
    for k in ("Te", "m_TM_x", "m_RE_x"):
        assert rel_l2(fa[k], fb[k]) < 1e-12, k
    gpu.free()


def test_coupled_pump_mo_dft_small():
    """Graded z axis, Ex/Ey gyration driven by m, 4 LLB substeps, DFT planes."""
    import driver
    from mphip import configs
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    spec, meta = configs.coupled_pump(nx=40, ny=36, nz=64, steps=200, fluence_mj_cm2=12.0, npml=6,
                                       src_z=40e-9)
    print("case coupled_pump 40x36x64 steps=200", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    mon = DFTPlanes(gpu, meta["planes"], meta["omega_bins"])
    ref_mon = OracleDFT(ref, meta["planes"], meta["omega_bins"])
    for _ in range(200):
        gpu.step()
        mon.record(gpu)
        ref.step()
        ref_mon.record(ref)
    gpu.synchronize()
    for name in FIELDS:
        assert rel_l2(gpu.field(name), ref.field(name)) < 1e-5, name
    fa, fb = gpu.film_state(), ref.film_state()
    assert fb["Te"].max() > 300.5
    for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k
    for p in range(len(meta["planes"])):
        g = mon.download(p)
        r = ref_mon.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            assert rel_l2(g[k], r[k]) < 1e-5, (p, k)
    gpu.free()


from driver import OracleDFT  # noqa: E402


def test_soft_and_mode_sources_bitwise():
    import math
    import driver
    from mphip import configs, sources
    from mphip.state import FDTDState
    for kind in ("soft", "mode"):
        spec, meta = configs.small_coupled(dtype="f32", steps=80, amplitude=3e8, n=24, source_kind="soft")
        if kind == "mode":
            ny, nz = spec.shape[1], spec.shape[2]
            prof = sources.gaussian_mode_profile(ny, nz, 20e-9, 20e-9)
            spec.source = sources.ModeSource(prof, spec.source.pulse, neff=1.7, axis="x", position=6,
                                             component="Ez")
        print(f"case source={kind}", flush=True)
        gpu = FDTDState(spec)
        ref = driver.OracleState(spec)
        for _ in range(80):
            gpu.step()
            ref.step()
        for name in FIELDS:
            a, b = gpu.field(name), ref.field(name)
            assert np.array_equal(a, b), (kind, name, np.abs(a - b).max())
        gpu.free()


def test_solved_mode_source_bitwise():
    """A mode source whose profile comes from the scalar mode solver on the waveguide's
    cross-section (Experiment.jl:149-165) injects through the same kernel, bit for bit."""
    import math
    import driver
    from mphip import geometry, grid, sources
    from mphip.modesolver import mode_source
    from mphip.spec import build_state_spec
    from mphip.state import FDTDState
    d = 20e-9
    g = grid.Grid3D(grid.uniform_axis_n(48, d), grid.uniform_axis_n(40, d), grid.uniform_axis_n(36, d))
    sc = geometry.Scene()
    sc.add(geometry.Box(0.0, 48 * d, 14 * d, 26 * d, 13 * d, 23 * d), geometry.Material("core", epsr=4.0))
    geo = geometry.rasterize(sc, g, background=1.44 ** 2)
    pulse = sources.GaussianPulse(amplitude=1e3, t0=4e-15, tau=1.5e-15, omega=2 * math.pi * grid.C0 / 800e-9)
    src = mode_source(g, pulse, 8, 800e-9, epsr_volume=geo.epsr, component="Ez")
    assert 1.44 < src.neff < 2.0
    spec = build_state_spec(g, geo, grid.cfl_dt_cells(g, courant=0.5), source=src, n_pml=4)
    print(f"case solved mode source neff={src.neff:.5f}", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    gpu.run(80)
    ref.run(80)
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.array_equal(a, b), (name, np.abs(a - b).max())
    assert np.abs(ref.field("ez")).max() > 0
    gpu.free()


def test_mode_source_dft_record_batched():
    """A mode source also writes H, after the H update.  The DFT record that mp_run_steps
    batches (snapshots, then one pass per batch) must see that H, as the oracle's per-step
    record does; a plane on the source plane itself pins the order."""
    import math
    import driver
    from mphip import geometry, grid, sources
    from mphip.modesolver import mode_source
    from mphip.monitors import DFTPlanes
    from mphip.spec import build_state_spec
    from mphip.state import FDTDState
    d = 20e-9
    g = grid.Grid3D(grid.uniform_axis_n(48, d), grid.uniform_axis_n(40, d), grid.uniform_axis_n(36, d))
    sc = geometry.Scene()
    sc.add(geometry.Box(0.0, 48 * d, 14 * d, 26 * d, 13 * d, 23 * d), geometry.Material("core", epsr=4.0))
    geo = geometry.rasterize(sc, g, background=1.44 ** 2)
    w0 = 2 * math.pi * grid.C0 / 800e-9
    pulse = sources.GaussianPulse(amplitude=1e3, t0=4e-15, tau=1.5e-15, omega=w0)
    src = mode_source(g, pulse, 8, 800e-9, epsr_volume=geo.epsr, component="Ez")
    spec = build_state_spec(g, geo, grid.cfl_dt_cells(g, courant=0.5), source=src, n_pml=4)
    planes = [(0, 8), (0, 30), (1, 20), (2, 18)]
    bins = [w0 * f for f in (0.97, 1.0, 1.03)]
    print("case mode source + batched DFT record 48x40x36", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    mon = DFTPlanes(gpu, planes, bins)
    gpu.attach_dft(mon)
    ref_mon = OracleDFT(ref, planes, bins)
    gpu.run(45)
    for _ in range(45):
        ref.step()
        ref_mon.record(ref)
    gpu.synchronize()
    for name in FIELDS:
        assert np.array_equal(gpu.field(name), ref.field(name)), name
    for p in range(len(planes)):
        gm, rm = mon.download(p), ref_mon.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            # fields are bit-identical; cos/sin of the phase factors differ in the last ULP
            assert rel_l2(gm[k], rm[k]) < 1e-10, (planes[p], k, rel_l2(gm[k], rm[k]))
        assert np.abs(rm["hb"]).max() > 0 or np.abs(rm["ha"]).max() > 0
    gpu.free()


def test_relax_and_reductions():
    import driver
    from mphip import configs
    from mphip.monitors import film_reduce
    from mphip.state import FDTDState
    spec, meta = configs.small_coupled(dtype="f32", steps=100, amplitude=6e8, n=24)
    print("case relax + film reductions", flush=True)
    gpu, ref = run_pair(spec, 100)
    gpu.flush_multiphysics()
    ref.flush_multiphysics()
    for _ in range(20):
        gpu.relax_step(5e-15, subcycles=4)
        ref.relax_step(5e-15, subcycles=4)
    fa, fb = gpu.film_state(), ref.film_state()
    for k in ("Te", "Tl", "m_TM_x", "m_RE_x"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() <= 1e-4 * max(1.0, np.abs(fb[k]).max()), k
    red = film_reduce(gpu)
    assert abs(red["m_TM_x"] - fa["m_TM_x"].astype(np.float64).mean()) < 1e-9
    assert abs(red["Te_max"] - float(fa["Te"].max())) == 0.0
    assert red["switched"] == float(np.mean(fa["m_TM_x"] < 0))
    gpu.free()


def _em_box_spec(nx, ny, nz, npml, seed, dtype="f32", psi=None):
    """EM-only rectangular grid with random dielectric boxes and a soft source."""
    import math
    from mphip import geometry, grid, sources
    from mphip.spec import build_state_spec
    d = 20e-9
    g = grid.Grid3D(grid.uniform_axis_n(nx, d), grid.uniform_axis_n(ny, d), grid.uniform_axis_n(nz, d))
    rng = np.random.default_rng(seed)
    sc = geometry.Scene()
    for b in range(6):
        lo = rng.uniform(0, 0.6, 3)
        hi = lo + rng.uniform(0.1, 0.4, 3)
        L = np.array([nx, ny, nz]) * d
        sc.add(geometry.Box(lo[0] * L[0], hi[0] * L[0], lo[1] * L[1], hi[1] * L[1], lo[2] * L[2], hi[2] * L[2]),
               geometry.Material(f"b{b}", epsr=float(rng.uniform(1, 9))))
    geo = geometry.rasterize(sc, g)
    dt = grid.cfl_dt_cells(g, courant=0.5)
    pulse = sources.GaussianPulse(amplitude=1e3, t0=3e-15, tau=1.5e-15, omega=2 * math.pi * grid.C0 / 800e-9)
    src = sources.SoftSource(pulse, "Ey", (nx // 2, ny // 2, nz // 2))
    return build_state_spec(g, geo, dt, compute_precision=dtype, field_precision=dtype, source=src,
                            n_pml=npml, psi_precision=psi)


@pytest.mark.parametrize("shape", [(264, 12, 16), (26, 18, 14), (36, 20, 22), (512, 8, 8)])
def test_em_boxes_bitwise_f32(shape):
    """Vector kernels (x multiple of 4, rows over several workgroups / idle lanes) and the
    scalar fallback (x not a multiple of 4) against the oracle, bit for bit."""
    import driver
    from mphip.state import FDTDState
    nx, ny, nz = shape
    spec = _em_box_spec(nx, ny, nz, npml=3, seed=nx + ny)
    print(f"case em boxes {shape}", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    gpu.run(60)
    ref.run(60)
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.any(a != 0) or name.startswith("d")
        assert np.array_equal(a, b), (shape, name, np.abs(a - b).max())
    gpu.free()


def test_coupled_pump_batched_run_overlap():
    """Many steps per mp_run_steps call: the film update of step n overlaps step n+1 on a
    side stream; results must match the step-by-step oracle, monitors included."""
    import driver
    from mphip import configs
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    spec, meta = configs.coupled_pump(nx=44, ny=40, nz=64, steps=240, fluence_mj_cm2=14.0, npml=6,
                                       src_z=40e-9)
    print("case coupled_pump batched 44x40x64", flush=True)
    gpu = FDTDState(spec)
    mon = DFTPlanes(gpu, meta["planes"], meta["omega_bins"])
    gpu.attach_dft(mon)
    ref = driver.OracleState(spec)
    ref_mon = OracleDFT(ref, meta["planes"], meta["omega_bins"])
    gpu.run(120)
    gpu.run(120)
    for _ in range(240):
        ref.step()
        ref_mon.record(ref)
    gpu.synchronize()
    assert gpu.n == 240 and abs(gpu.t - ref.t) <= 1e-12 * ref.t
    for name in FIELDS:
        assert rel_l2(gpu.field(name), ref.field(name)) < 1e-5, name
    fa, fb = gpu.film_state(), ref.film_state()
    assert fb["Te"].max() > 300.5
    for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
        assert rel_l2(fa[k], fb[k]) < 1e-6, k
    for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k
    for p in range(len(meta["planes"])):
        g, r = mon.download(p), ref_mon.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            assert rel_l2(g[k], r[k]) < 1e-5, (p, k)
    gpu.free()


@pytest.mark.parametrize("shape", [(40, 36, 64), (44, 40, 78)])
def test_batched_dft_record_is_bitwise_per_step(shape):
    """Inside one mp_run_steps call the DFT record is batched (snapshots, then one pass over
    the stored steps in step order); one step per call records step by step.  The two must
    give the same bits.  The second shape has no axis a multiple of 32 cells."""
    from mphip import configs
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    nx, ny, nz = shape
    spec, meta = configs.coupled_pump(nx=nx, ny=ny, nz=nz, steps=200, fluence_mj_cm2=12.0, npml=6, src_z=40e-9)
    print(f"case batched DFT record vs per step {nx}x{ny}x{nz}", flush=True)
    a, b = FDTDState(spec), FDTDState(spec)
    ma = DFTPlanes(a, meta["planes"], meta["omega_bins"])
    mb = DFTPlanes(b, meta["planes"], meta["omega_bins"])
    a.attach_dft(ma)
    b.attach_dft(mb)
    a.run(37)
    a.run(40)
    for _ in range(77):
        b.step()
    for p in range(len(meta["planes"])):
        ga, gb = ma.download(p), mb.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            assert np.array_equal(ga[k], gb[k]), (p, k)
    for name in FIELDS:
        assert np.array_equal(a.field(name), b.field(name)), name
    a.free()
    b.free()


@pytest.mark.parametrize("poles,dtype", [(2, "f32"), (3, "f32"), (2, "f64")])
def test_fused_film_epilogue_matches_separate_kernels(poles, dtype):
    """The one-pass film list work (ADE x/y/z, gyration, absorbed power) against the five
    separate launches on the same state: every field, list state and film value equal.  With
    the film's two ADE and two MO poles the pass runs its two-pole kernel (every load of a cell
    before its first store); with a third ADE pole, the kernel for any pole count."""
    from mphip import configs, dispersion, grid
    from mphip.state import FDTDState
    spec, meta = configs.coupled_pump(nx=40, ny=36, nz=64, steps=150, fluence_mj_cm2=12.0, npml=6,
                                       src_z=40e-9, dtype=dtype)
    if poles == 3:
        extra = dispersion.create_pole(6e15, 8e14, 0.5 * (6e15) ** 2, spec.dt, grid.EPS0)
        spec.diag_poles = dispersion.poles_array(configs.film_poles(spec.dt) + [extra], spec.compute_dtype)
    print(f"case coupled_pump 40x36x64 fused vs separate, {poles} ADE poles, {dtype}, 150 steps", flush=True)
    fused, sep = FDTDState(spec), FDTDState(spec)
    assert fused.fused_epilogue, "the pump's lists are the film list"
    sep.fused_epilogue = False
    fused.run(150)
    sep.run(150)
    fused.synchronize()
    for name in FIELDS:
        assert np.array_equal(fused.field(name), sep.field(name)), name
    for c in range(3):
        a, b = fused.ade_state(c), sep.ade_state(c)
        for k in a:
            assert np.array_equal(a[k], b[k]), (c, k)
    fa, fb = fused.film_state(), sep.film_state()
    assert fb["Te"].max() > 300.5
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    fused.free()
    sep.free()


@pytest.mark.parametrize("shape,npml", [((96, 68, 64), 6), ((128, 40, 72), 9), ((256, 64, 48), 1)])
def test_split_yee_update_bitwise(shape, npml):
    """Grids whose H and E updates split into the interior-box launch and the CPML-shell
    launch, whose lanes decode their group from the shell layout (box edges on whole cache
    lines, shells of uneven and of one-cell width, rows that are not a multiple of the wave,
    lanes past the end of the shell): every field and psi slab bit-identical to the oracle.
    From the second step on the film update of the step before runs beside the H update, and
    the rest of the H shell then runs split by x (the box's x range without x-slab terms, then
    the x edges), so both H shell forms are covered."""
    import driver
    from mphip import configs
    from mphip.state import FDTDState
    from test_gpu_fullsize import _seed_fields
    nx, ny, nz = shape
    spec, meta = configs.coupled_pump(nx=nx, ny=ny, nz=nz, steps=400, fluence_mj_cm2=12.0, npml=npml,
                                       src_z=40e-9)
    print(f"case split yee {shape} npml={npml}", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    _seed_fields(gpu, ref, spec.shape, 5)
    gpu.run(4)
    ref.run(4)
    gpu.synchronize()
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.array_equal(a, b), (name, int(np.count_nonzero(a != b)))
    for s in range(12):
        assert np.array_equal(gpu.psi("h", s), ref.psi_h[s]), ("psi_h", s)
        assert np.array_equal(gpu.psi("d", s), ref.psi_d[s]), ("psi_d", s)
    assert any(np.any(ref.psi_h[s]) for s in range(12))
    gpu.free()
