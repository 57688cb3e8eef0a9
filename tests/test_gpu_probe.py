"""Probe read-out, polarimetry and transmission on the GPU against the oracle.

The device reductions (mp_probe_shot_reduce, mp_plane_overlap) sum in a fixed partition
that differs from the oracle's column-major order, so they are compared within 1e-12 of
the sum of absolute terms (FP64 reassociation).  ProbeReadout.run records the DFT inside
the step driver (batched) and is compared with the oracle recording after every step;
the EM-only fields are bit-identical, the DFT phases come from device cos/sin, so the
shot values agree to about 1e-12 relative.  The analytic plane wave of the reference's
test (runtests.jl:517-565) is replayed on the device path.
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Z0 = 376.730313668


def _uniform_spec(nx, ny, nz, dx=20e-9, precision="f32", source=None, npml=0, dt=None):
    from mphip import geometry, grid
    from mphip.spec import build_state_spec
    g = grid.Grid3D(grid.uniform_axis_n(nx, dx), grid.uniform_axis_n(ny, dx), grid.uniform_axis_n(nz, dx))
    geo = geometry.rasterize(geometry.Scene(), g)
    dt = dt or grid.cfl_dt_cells(g, courant=0.5)
    return build_state_spec(g, geo, dt, compute_precision=precision, field_precision=precision, source=source,
                            n_pml=npml)


def _close(a, b, scale, rel=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.all(np.abs(a - b) <= rel * np.maximum(scale, 1e-300)), (np.abs(a - b).max(), scale)


@pytest.mark.parametrize("shape,nb", [((6, 37, 29), 5), ((5, 300, 200), 3), ((7, 16, 16), 1), ((6, 12, 10), 13)])
def test_shot_reduce_matches_oracle(shape, nb):
    import driver
    from mphip import _lib
    from mphip.monitors import DFTPlanes, plane_area
    from mphip.state import DeviceBuffer, FDTDState
    print(f"case shot reduce {shape} nb={nb}", flush=True)
    spec = _uniform_spec(*shape)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    nx = shape[0]
    bins = np.linspace(3e15, 4e15, nb)
    planes = [(0, 1), (0, nx - 2)]
    mon = DFTPlanes(gpu, planes, bins)
    odft = driver.OracleDFT(ref, planes, bins)
    rng = np.random.default_rng(shape[1] + nb)
    for p in range(2):
        for k, a in odft.raw(p).items():
            a[...] = rng.standard_normal(a.size) * 1e-3
            mon.bufs[p][k].upload(a)
    npc = shape[1] * shape[2]
    area = plane_area(spec.grid, 0) * rng.uniform(0.5, 1.5, npc)
    mode_w = rng.standard_normal(npc)
    q0 = nb // 2
    dev_out = DeviceBuffer.zeros(8 * (4 * nb + 10))
    a_dev, w_dev = DeviceBuffer.from_array(area), DeviceBuffer.from_array(mode_w)
    _lib.call("probe_shot_reduce", C.byref(mon.arr[0]), C.byref(mon.arr[1]), shape[1], shape[2], nb, q0,
              w_dev.vp, a_dev.vp, dev_out.vp)
    got = dev_out.download(np.float64)
    want = np.zeros(4 * nb + 10)
    rc = ref.L.orc_probe_shot_reduce(C.byref(odft.desc[0]), C.byref(odft.desc[1]), shape[1], shape[2], nb, q0,
                                     mode_w.ctypes.data_as(C.c_void_p), area.ctypes.data_as(C.c_void_p),
                                     want.ctypes.data_as(C.c_void_p))
    assert rc == 0
    # scale of each output: the sum of the absolute values of its terms
    sc = np.zeros_like(want)
    for p in range(2):
        d = odft.download(p)
        A = area.reshape(shape[1:], order="F")
        for q in range(nb):
            s = np.sum((np.abs(d["ea"][..., q]) * np.abs(d["hb"][..., q]) +
                        np.abs(d["eb"][..., q]) * np.abs(d["ha"][..., q])) * A)
            sc[2 * nb * p + 2 * q:2 * nb * p + 2 * q + 2] = s
        sc[4 * nb + p] = np.sum(np.abs(d["energy"]) * A)
    W = np.abs(mode_w.reshape(shape[1:], order="F")) * area.reshape(shape[1:], order="F")
    for i, (p, k) in enumerate(((1, "ea"), (1, "eb"), (0, "ea"), (0, "eb"))):
        sc[4 * nb + 2 + 2 * i:4 * nb + 4 + 2 * i] = np.sum(np.abs(odft.download(p)[k][..., q0]) * W)
    _close(got, want, sc)
    # same partition every call: bit-identical on a repeat
    _lib.call("probe_shot_reduce", C.byref(mon.arr[0]), C.byref(mon.arr[1]), shape[1], shape[2], nb, q0,
              w_dev.vp, a_dev.vp, dev_out.vp)
    assert np.array_equal(dev_out.download(np.float64), got)
    for b in (dev_out, a_dev, w_dev):
        b.free()
    mon.free()
    gpu.free()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_plane_overlap_matches_oracle(precision):
    import driver
    from mphip import _lib
    from mphip.state import DeviceBuffer, FDTDState
    shape = (18, 21, 13)
    print(f"case plane overlap {precision} {shape}", flush=True)
    spec = _uniform_spec(*shape, precision=precision)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    rng = np.random.default_rng(7)
    for k in ("ex", "ey", "ez", "hx", "hy", "hz"):
        v = rng.standard_normal(shape)
        gpu.set_field(k, v)
        ref.set_field(k, v)
    out = DeviceBuffer.zeros(8)
    for axis in range(3):
        n = ((shape[1] * shape[2]), (shape[2] * shape[0]), (shape[0] * shape[1]))[axis]
        w = rng.standard_normal(n)
        w_dev = DeviceBuffer.from_array(w)
        for comp in range(6):
            for idx in (0, shape[axis] // 2, shape[axis] - 1):
                for weight, wd in ((w, w_dev.vp), (None, None)):
                    _lib.call("plane_overlap", C.byref(gpu.yee), comp, axis, idx, wd, out.vp)
                    got = float(out.download(np.float64)[0])
                    want = driver.plane_overlap(ref, comp, axis, idx, weight)
                    assert abs(got - want) <= 1e-12 * n * 4, (axis, comp, idx, got, want)
        w_dev.free()
    out.free()
    gpu.free()


def _x_probe_spec(precision="f32"):
    from mphip import grid, sources
    pulse = sources.GaussianPulse(amplitude=1e3, t0=6e-15, tau=2e-15, omega=2 * math.pi * grid.C0 / 532e-9)
    src = sources.SoftSource(pulse, "Ez", (8, 12, 10))
    return _uniform_spec(48, 24, 20, dx=10e-9, precision=precision, source=src, npml=4)


def test_probe_readout_run_matches_oracle():
    import driver
    from mphip.probe import ProbeReadout
    from mphip.state import FDTDState
    spec = _x_probe_spec()
    steps = 150
    print(f"case ProbeReadout.run {spec.shape} {steps} steps", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    mode = np.outer(np.hanning(24), np.hanning(20)) + 0.1
    ro = ProbeReadout(pre=14, post=34, mode=mode, lambda0=532e-9, trace_stride=10, final_step=steps)
    oro = ProbeReadout(pre=14, post=34, mode=mode, lambda0=532e-9, trace_stride=10, final_step=steps)
    ro.run(gpu, 77)
    ro.run(gpu, steps - 77)       # two driver batches, traces due across the boundary
    op = driver.OracleProbe(ref, oro)
    op.run(steps)
    for k in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert np.array_equal(gpu.field(k), ref.field(k)), k
    shot = ro.probe_shot(state_label="run", dt=spec.dt, steps_probe=steps)
    from mphip.probe import shot_from_reductions, unpack_reductions, trace_moments
    nb = len(ro.omega_bins)
    want = shot_from_reductions(unpack_reductions(op.reductions_raw(), nb), nb, op.mode_norm)
    assert shot["net_energy_pre_J"] != 0.0 and np.max(np.abs(shot["Sx_pre_net_re"])) > 0
    for k in ("Sx_pre_net_re", "Sx_pre_net_im", "Sx_post_net_re", "Sx_post_net_im"):
        scale = np.max(np.abs(np.concatenate([want["Sx_pre_net_re"], want["Sx_post_net_re"]])))
        _close(shot[k], want[k], scale, rel=1e-9)
    for k in ("net_energy_pre_J", "net_energy_post_J"):
        _close(shot[k], want[k], abs(want["net_energy_pre_J"]), rel=1e-9)
    for k in ("T", "R", "A"):
        assert abs(shot[k] - want[k]) <= 1e-9, k
    for k in ("jones_faraday_Ey_re_im", "jones_faraday_Ez_re_im", "jones_kerr_Ey_re_im", "jones_kerr_Ez_re_im"):
        _close(shot[k], want[k], np.max(np.abs(want["jones_kerr_Ez_re_im"])), rel=1e-9)
    for k in ("theta_faraday_deg", "theta_kerr_deg", "ellipticity_faraday_deg", "ellipticity_kerr_deg"):
        assert abs(shot[k] - want[k]) <= 1e-6, k
    # traces: n = 1, every 10th step and the final step
    t_ref = np.array(op.t_trace)
    assert len(shot["trace_time_s"]) == len(t_ref) == 1 + steps // 10
    np.testing.assert_allclose(shot["trace_time_s"], t_ref, rtol=1e-14)
    tr = np.array(op.traces)
    for c, k in enumerate(("Ey_trace_pre", "Ez_trace_pre", "Ey_trace_post", "Ez_trace_post")):
        _close(shot[k], tr[:, c], np.max(np.abs(tr)), rel=1e-12)
    m = trace_moments(t_ref, tr[:, 1])
    assert shot["input_pulse_rms_fs"] == pytest.approx(m["rms_fs"], rel=1e-9)
    ro.free()
    gpu.free()


def test_probe_plane_wave_reference_normalisation_on_device():
    """runtests.jl:517-565 through the device record and reductions."""
    from mphip.probe import ProbeReadout
    from mphip.state import FDTDState
    from mphip import geometry, grid
    from mphip.spec import build_state_spec
    print("case analytic plane wave through the device probe", flush=True)
    g = grid.Grid3D(grid.uniform_axis_n(7, 1e-7), grid.uniform_axis_n(4, 5e-8, -1e-7),
                    grid.uniform_axis_n(4, 5e-8, -1e-7))
    spec = build_state_spec(g, geometry.rasterize(geometry.Scene(), g), 1e-17, compute_precision="f64",
                            field_precision="f64")
    st = FDTDState(spec)
    ro = ProbeReadout(pre=1.0e-7, post=4.0e-7, lambda0=532e-9, trace_stride=1, spectrum_bins=5)
    w0 = 2 * math.pi * 299792458.0 / 532e-9
    for n in range(1, 201):
        t = n * 1e-17
        st.n, st.t = n, t
        st.set_field("ez", np.full(spec.shape, math.cos(w0 * t)))
        st.set_field("hy", np.full(spec.shape, -math.cos(w0 * t) / Z0))
        ro.record(st)
    ref = ro.probe_shot(state_label="reference", is_reference=True, dt=1e-17)
    assert np.max(np.abs(ref["Sx_pre_net_re"])) > 0
    assert all((not np.isfinite(r)) or abs(r) <= 1e-9 for r in ref["R_omega"])
    nonref = ro.probe_shot(state_label="initial", dt=1e-17, incident_energy=2 * ref["incident_energy_J"],
                           incident_sx=2 * ref["Sx_inc"])
    fin = [r for r in nonref["R_omega"] if np.isfinite(r)]
    assert any(r != 0.0 for r in fin) and max(fin) == pytest.approx(0.5, abs=1e-6)
    np.testing.assert_allclose(ref["T_omega"], 1.0, rtol=1e-12)
    t = np.arange(1, 201) * 1e-17
    np.testing.assert_allclose(ref["Ez_trace_pre"], np.cos(w0 * t), rtol=1e-12, atol=1e-15)
    ro.free()
    st.free()


def test_polarimetry_and_transmission_match_oracle():
    import driver
    from mphip.monitors import plane_area
    from mphip.probe import Polarimetry, Reflection, Transmission
    from mphip.state import FDTDState
    spec = _x_probe_spec(precision="f64")
    steps = 60
    print(f"case polarimetry / transmission {spec.shape} {steps} steps", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    g = spec.grid
    mode_x = np.outer(np.hanning(24), np.hanning(20)) + 0.05
    mode_y = np.random.default_rng(3).uniform(0.5, 1.5, (48, 20))
    mons = [Polarimetry(), Polarimetry(at=30, mode=mode_x), Polarimetry(at=12, axis="y", mode=mode_y),
            Transmission(at=34), Reflection(at=14), Transmission(at=34, component="Ez", mode=mode_x),
            Reflection(at=14, component="Ey", mode=mode_x, incident=2.0)]
    ser = {i: [] for i in range(len(mons))}
    wx = (mode_x.reshape(-1, order="F") * plane_area(g, 0))
    wy = (mode_y.T.reshape(-1, order="F") * plane_area(g, 1))
    ax = plane_area(g, 0)
    for _ in range(steps):
        gpu.step()
        ref.step()
        for m in mons:
            m.record(gpu)
        n = 24 * 20
        ser[0].append((driver.plane_overlap(ref, 1, 0, 47) / n, driver.plane_overlap(ref, 2, 0, 47) / n))
        ser[1].append((driver.plane_overlap(ref, 1, 0, 30, wx), driver.plane_overlap(ref, 2, 0, 30, wx)))
        ser[2].append((driver.plane_overlap(ref, 1, 1, 12, wy), driver.plane_overlap(ref, 2, 1, 12, wy)))
        fl = np.zeros(2)
        for s, idx in enumerate((34, 14)):
            assert ref.L.orc_flux_plane(C.byref(ref.yee), 0, idx, ax.ctypes.data_as(C.c_void_p),
                                        fl.ctypes.data_as(C.c_void_p), s) == 0
        ser[3].append(fl[0])
        ser[4].append(-fl[1])
        ser[5].append(driver.plane_overlap(ref, 2, 0, 34, wx))
        ser[6].append(driver.plane_overlap(ref, 1, 0, 14, wx))
    for i in range(3):
        d = mons[i].data()
        want = np.array(ser[i])
        scale = np.max(np.abs(want))
        assert scale > 0
        _close(np.stack([d["ey"], d["ez"]], 1), want, scale * want.shape[0], rel=1e-12)
        assert len(d["t"]) == steps and math.isfinite(d["ellipticity_deg"])
    for i in range(3, 7):
        d = mons[i].data()
        want = np.array(ser[i])
        assert np.max(np.abs(want)) > 0, i
        _close(d["values"], want, np.max(np.abs(want)) * 4, rel=1e-12)
    assert mons[4].data()["reflection"] == pytest.approx(
        float(np.sum(ser[4]) * (mons[4].t[1] - mons[4].t[0])), rel=1e-9)
    assert mons[6].data()["modal"] and mons[6].data()["ratio"] == pytest.approx(
        float(np.sum(np.array(ser[6]) ** 2) * spec.dt) / 2.0, rel=1e-9)
    for m in mons:
        m.free()
    gpu.free()


@pytest.mark.parametrize("how", ["run", "record"])
def test_probe_frames_match_oracle_slices(how):
    """Movie frames staged on the device (mp_probe_frames) against Float32.(abs.(...)) of the
    oracle's field planes at the same steps (_record_probe_frame!, ProbeReadout.jl:193-215):
    the z slice and the pre and post x-planes, stacked frame-last.  `run` takes frames every
    5th step across two driver batches, on the domain's middle z plane; `record` takes one
    every step up to 40 (frame_skip 0) on a given z position, past one 32-frame chunk."""
    import driver
    from mphip.probe import ProbeReadout
    from mphip.state import FDTDState
    spec = _x_probe_spec()
    steps = 70
    print(f"case probe frames {how} {spec.shape}", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    kw = dict(pre=14, post=34, lambda0=532e-9, trace_stride=10)
    if how == "run":
        ro = ProbeReadout(frame_target=12, frame_skip=5, **kw)
        ro.run(gpu, 33)
        ro.run(gpu, steps - 33)
        sched, kc = list(range(5, 61, 5)), (spec.shape[2] + 1) // 2 - 1
    else:
        ro = ProbeReadout(frame_target=40, slice_z_position=float(spec.grid.z.centers[3]), **kw)
        for _ in range(steps):
            gpu.step()
            ro.record(gpu)
        sched, kc = list(range(1, 41)), 3
    assert ro.frame_indices == sched and ro.kc == kc
    want = {k: [] for k in ("Ey_xy", "Ez_xy", "Ey_yz_pre", "Ez_yz_pre", "Ey_yz_post", "Ez_yz_post")}
    times = []
    for n in range(1, steps + 1):
        ref.step()
        if n in sched:
            ey, ez = ref.field("ey"), ref.field("ez")
            for k, f in (("Ey", ey), ("Ez", ez)):
                want[k + "_xy"].append(np.abs(f[:, :, kc]).astype(np.float32))
                want[k + "_yz_pre"].append(np.abs(f[14, :, :]).astype(np.float32))
                want[k + "_yz_post"].append(np.abs(f[34, :, :]).astype(np.float32))
            times.append(ref.t)
    for k in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert np.array_equal(gpu.field(k), ref.field(k)), k
    shot = ro.probe_shot(state_label=how, dt=spec.dt, steps_probe=steps)
    assert shot["probe_frame_count"] == len(sched) and shot["probe_frame_indices"] == sched
    np.testing.assert_allclose(shot["probe_frame_times_s"], times, rtol=1e-14)
    nx, ny, nz = spec.shape
    for k, frames in want.items():
        cube = shot["frames"][k]
        assert cube.dtype == np.float32
        assert cube.shape == ((nx, ny) if k.endswith("_xy") else (ny, nz)) + (len(sched),), (k, cube.shape)
        w = np.stack(frames, axis=2)
        assert np.array_equal(cube, w), (k, int(np.count_nonzero(cube != w)))
        assert np.any(w > 0), k
    ro.free()
    gpu.free()
