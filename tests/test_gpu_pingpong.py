"""Ping-pong steps (the fused interior H + E launch) against the in-place steps on the GPU.

mp_run_steps advances an FP32 state in pairs of ping-pong steps when the state carries a
second E/H buffer set: each step reads E and H from one set and writes the other, with the
interior H and E updates joined in one LDS-tiled launch.  That launch recomputes the row of
H its E update needs from unmodified inputs and runs the per-cell arithmetic of the
separate interior kernels, so every field, psi slab, ADE state and film value must equal
the in-place run's bit for bit.  The cases cover the material-index and inv_eps-volume E
forms, BF16 psi/ADE storage, a plane source on z (two weighted planes), one on x through
the x = 0 plane (cells no E update writes, where the injection alone accumulates), a soft
source, the DFT record and odd step counts (an odd last step runs in place); the fused
launch is checked to have run.  Against the oracle the ping-pong path is held to the
coupled bar of the other pump tests.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d else float(np.linalg.norm(a - b))


def _spec(case):
    from mphip import configs, sources
    kw = dict(nx=96, ny=68, nz=64, steps=200, fluence_mj_cm2=12.0, npml=6, src_z=40e-9)
    if case == "bf16":
        kw.update(psi_precision="bf16", ade_state_precision="bf16")
    spec, meta = configs.coupled_pump(**kw)
    if case == "plane_x":
        xc = spec.grid.x.centers
        spec.source = sources.PlaneSource(spec.source.pulse, "Ey", axis="x", position=0.7 * xc[0] + 0.3 * xc[1])
    elif case == "soft":
        spec.source = sources.SoftSource(spec.source.pulse, "Ex", (40, 30, 36))
    return spec, meta


def _fused_launches(L):
    n, ms = C.c_int64(), C.c_double()
    L.mp_kernel_stats(b"yee_HE", C.byref(n), C.byref(ms))
    return n.value


@pytest.mark.parametrize("case,eps_index", [("plane_z", True), ("plane_z", False), ("bf16", True),
                                            ("plane_x", True), ("soft", True)])
def test_pingpong_matches_in_place(case, eps_index):
    from mphip import _lib
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    spec, meta = _spec(case)
    print(f"case ping-pong vs in place {case} eps_index={eps_index} {spec.shape}", flush=True)
    pp = FDTDState(spec, use_eps_index=eps_index, pingpong=True)
    ip = FDTDState(spec, use_eps_index=eps_index)
    assert pp.yee_alt is not None and ip.yee_alt is None
    mons = []
    if case == "plane_z":
        for st in (pp, ip):
            m = DFTPlanes(st, meta["planes"], meta["omega_bins"])
            st.attach_dft(m)
            mons.append(m)
    L = _lib.lib()
    L.mp_profile_select(None)
    L.mp_profile_enable(1)
    for n in (37, 2, 11):
        pp.run(n)
    fused = _fused_launches(L)
    L.mp_profile_enable(0)
    for n in (37, 2, 11):
        ip.run(n)
    pp.synchronize()
    assert fused == 36 + 2 + 10, fused   # the steps of each call's pairs; odd last steps in place
    for name in FIELDS:
        a, b = pp.field(name), ip.field(name)
        assert np.array_equal(a, b), (name, int(np.count_nonzero(a != b)))
    assert np.any(pp.field("hx")) and np.any(pp.field("ez"))
    for s in range(12):
        assert np.array_equal(pp.psi("h", s), ip.psi("h", s)), ("psi_h", s)
        assert np.array_equal(pp.psi("d", s), ip.psi("d", s)), ("psi_d", s)
    for c in range(3):
        a, b = pp.ade_state(c), ip.ade_state(c)
        for k in ("P", "J", "E_old"):
            assert np.array_equal(a[k], b[k]), (c, k)
    fa, fb = pp.film_state(), ip.film_state()
    assert fb["Te"].max() > 300.2
    for k in fb:
        assert np.array_equal(fa[k], fb[k]), k
    if case == "plane_x":
        # the x = 0 plane is written only by the source: both sets must carry its sum
        assert np.any(pp.field("ey")[0])
    if mons:
        for p in range(len(meta["planes"])):
            ga, gb = mons[0].download(p), mons[1].download(p)
            for k in ga:
                assert np.array_equal(ga[k], gb[k]), (p, k)
        for m in mons:
            m.free()
    pp.free()
    ip.free()


def test_pingpong_against_oracle():
    import driver
    from mphip.state import FDTDState
    spec, _ = _spec("plane_z")
    print("case ping-pong vs oracle pump 96x68x64, 60 steps", flush=True)
    gpu = FDTDState(spec, pingpong=True)
    ref = driver.OracleState(spec)
    gpu.run(60)
    ref.run(60)
    gpu.synchronize()
    for name in FIELDS:
        assert _rel(gpu.field(name), ref.field(name)) < 1e-5, name
    fa, fb = gpu.film_state(), ref.film_state()
    for k in ("Te", "Tl"):
        assert _rel(fa[k], fb[k]) < 1e-6, k
    for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k
    gpu.free()


def test_pingpong_full_size_from_random_fields():
    """512 x 512 x 256 EM volume (CPML 10) from seeded random E, H and D: 4 ping-pong steps
    equal 4 in-place steps bit for bit."""
    from mphip import configs
    from mphip.state import FDTDState
    spec, _ = configs.coupled_pump(nx=512, ny=512, nz=256, steps=20000, n_planes=0)
    print("case ping-pong vs in place 512x512x256 from random fields", flush=True)
    rng = np.random.default_rng(11)
    states = [FDTDState(spec, pingpong=True), FDTDState(spec)]
    for name in FIELDS:
        v = rng.standard_normal(spec.shape).astype(np.float32)
        for st in states:
            st.set_field(name, v)
    for st in states:
        st.run(4)
        st.synchronize()
    for name in FIELDS:
        a, b = states[0].field(name), states[1].field(name)
        assert np.array_equal(a, b), (name, int(np.count_nonzero(a != b)))
    for st in states:
        st.free()
