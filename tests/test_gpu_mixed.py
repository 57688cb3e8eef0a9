"""Reduced-precision storage (BF16 / binary16 psi and ADE P/J, FP32 arithmetic) on the GPU.

Parity: with the gyration off, nothing in the field update depends on the FP64 film
integrator, so fields, raw psi slabs and raw ADE state must equal the oracle's restatement
of the same storage form bit for bit.  With the gyration on, the bar is the coupled
path's: fields within 1e-5 relative L2 and magnetisation within 1e-4.

Accuracy: the coupled pump on a 128 x 128 x 64 copy of the pump geometry (60 mJ/cm^2, 1500
steps, then 600 relax steps of 5 fs) switches about 5 % of the film.  Each storage form is
compared with the FP64 oracle on identical inputs.  Plain FP32 meets the film bar (mean
|dm| <= 1e-3); binary16 storage stays within a few per cent on fields and temperature but
moves the switched fraction; BF16 storage is far outside both (held to its measured envelope).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")
PUMP = dict(nx=128, ny=128, nz=64, steps=1500, fluence_mj_cm2=60.0, npml=8, src_z=60e-9, n_planes=0)
RELAX, RELAX_DT, RELAX_SUB = 600, 5e-15, 10
_REF = {}


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


# 40 x 36 x 64 runs the whole-volume Yee launches; 96 x 68 x 64 the interior-box and
# CPML-shell launches (the shell's psi in the reduced storage)
@pytest.mark.parametrize("kind,shape", [("bf16", (40, 36, 64)), ("f16", (40, 36, 64)),
                                        ("bf16", (96, 68, 64)), ("f16", (96, 68, 64))])
def test_reduced_storage_bitwise(kind, shape):
    import driver
    from mphip import configs
    from mphip.state import FDTDState
    nx, ny, nz = shape
    spec, _ = configs.coupled_pump(nx=nx, ny=ny, nz=nz, steps=200, fluence_mj_cm2=12.0, npml=6,
                                   src_z=40e-9, n_planes=0, enable_mo=False, psi_precision=kind,
                                   ade_state_precision=kind)
    print(f"case reduced storage bitwise {kind} {nx}x{ny}x{nz}", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    gpu.run(200)
    ref.run(200)
    gpu.synchronize()
    for name in FIELDS:
        assert np.array_equal(gpu.field(name), ref.field(name)), name
    for s in range(12):
        assert np.array_equal(gpu.psi("h", s), ref.psi_h[s]), ("psi_h", s)
        assert np.array_equal(gpu.psi("d", s), ref.psi_d[s]), ("psi_d", s)
    assert any(np.any(ref.psi_h[s]) for s in range(12))
    for c in range(3):
        g, r = gpu.ade_state(c), ref.ade_arr[c]
        for k in ("P", "J", "E_old"):
            assert np.array_equal(g[k], r[k]), (c, k)
        assert np.any(r["J"])
    fa, fb = gpu.film_state(), ref.film_state()
    assert fb["Te"].max() > 300.5
    for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
        assert rel_l2(fa[k], fb[k]) < 1e-6, k
    gpu.free()


def test_reduced_storage_with_gyration():
    import driver
    from mphip import configs
    from mphip.state import FDTDState
    spec, _ = configs.coupled_pump(nx=40, ny=36, nz=64, steps=200, fluence_mj_cm2=12.0, npml=6,
                                   src_z=40e-9, n_planes=0, psi_precision="f16", ade_state_precision="f16")
    print("case reduced storage f16 with gyration 40x36x64", flush=True)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    gpu.run(200)
    ref.run(200)
    gpu.synchronize()
    for name in FIELDS:
        assert rel_l2(gpu.field(name), ref.field(name)) < 1e-5, name
    fa, fb = gpu.film_state(), ref.film_state()
    for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k
    gpu.free()


@pytest.mark.parametrize("kind", ["bf16", "f16"])
def test_reduced_storage_fused_pass_matches_five_launches(kind):
    """With BF16 / binary16 ADE state the step's list work still runs as one fused pass; it
    must give what the five separate launches give, bit for bit (gyration on)."""
    from mphip import configs
    from mphip.state import FDTDState
    spec, _ = configs.coupled_pump(nx=96, ny=68, nz=64, steps=150, fluence_mj_cm2=12.0, npml=6,
                                   src_z=40e-9, n_planes=0, psi_precision=kind, ade_state_precision=kind)
    print(f"case reduced storage {kind} fused vs five launches 96x68x64", flush=True)
    fused, sep = FDTDState(spec), FDTDState(spec)
    assert fused.fused_epilogue
    sep.fused_epilogue = False
    fused.run(150)
    sep.run(150)
    fused.synchronize()
    sep.synchronize()
    for name in FIELDS:
        assert np.array_equal(fused.field(name), sep.field(name)), name
    for c in range(3):
        g, r = fused.ade_state(c), sep.ade_state(c)
        for k in ("P", "J", "E_old"):
            assert np.array_equal(g[k], r[k]), (c, k)
        assert np.any(r["J"])
    fa, fb = fused.film_state(), sep.film_state()
    assert fb["Te"].max() > 300.5
    for k in fb:
        assert np.array_equal(fa[k], fb[k]), k
    fused.free()
    sep.free()


def _fp64_reference():
    if "ref" not in _REF:
        import driver
        from mphip import configs
        spec, _ = configs.coupled_pump(dtype="f64", **PUMP)
        ref = driver.OracleState(spec)
        ref.run(PUMP["steps"])
        fields = {k: ref.field(k) for k in ("ex", "ey", "ez", "hx", "hy", "hz")}
        te_peak = float(ref.film_state()["Te"].max())
        ref.flush_multiphysics()
        for _ in range(RELAX):
            ref.relax_step(RELAX_DT, subcycles=RELAX_SUB)
        _REF["ref"] = (fields, te_peak, ref.film_state())
    return _REF["ref"]


def run_form(psi, ade):
    from mphip import configs
    from mphip.state import FDTDState
    spec, _ = configs.coupled_pump(dtype="f32", psi_precision=psi, ade_state_precision=ade, **PUMP)
    st = FDTDState(spec)
    st.run(PUMP["steps"])
    fields = {k: st.field(k) for k in ("ex", "ey", "ez", "hx", "hy", "hz")}
    te_peak = float(st.film_state()["Te"].max())
    st.flush_multiphysics()
    for _ in range(RELAX):
        st.relax_step(RELAX_DT, subcycles=RELAX_SUB)
    film = st.film_state()
    st.free()
    return fields, te_peak, film


def errors(psi, ade):
    ref_f, ref_te, ref_film = _fp64_reference()
    f, te, film = run_form(psi, ade)
    out = {}
    # The cross-polarised components (Ey from the gyration, Hz) are orders of magnitude
    # below Ex, so fields are compared as the E and H vector fields.
    for v, comps in (("E", ("ex", "ey", "ez")), ("H", ("hx", "hy", "hz"))):
        d = sum(np.sum((f[k].astype(np.float64) - ref_f[k]) ** 2) for k in comps)
        out["relL2_" + v] = float(np.sqrt(d / sum(np.sum(ref_f[k] ** 2) for k in comps)))
    out["Te_peak_rel"] = abs(te - ref_te) / ref_te
    for k in ("m_TM_x", "m_RE_x"):
        a, b = film[k].astype(np.float64), ref_film[k]
        out["mean_dm_" + k] = float(abs(a.mean() - b.mean()))
        out["max_dm_" + k] = float(np.abs(a - b).max())
    out["switched_ref"] = float(np.mean(ref_film["m_TM_x"] < 0))
    out["switched"] = float(np.mean(film["m_TM_x"] < 0))
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("kind", ["f32", "f16", "bf16"])
def test_storage_form_vs_fp64_oracle(kind):
    print(f"case storage form {kind} vs FP64 oracle 128x128x64", flush=True)
    e = errors(kind, kind)
    print({k: f"{v:.3e}" for k, v in e.items()}, flush=True)
    assert 0.02 < e["switched_ref"] < 0.2
    if kind == "f32":
        assert e["mean_dm_m_TM_x"] <= 1e-3 and e["mean_dm_m_RE_x"] <= 1e-3
        assert e["Te_peak_rel"] < 1e-3
        assert e["relL2_E"] < 1e-4 and e["relL2_H"] < 1e-4
    elif kind == "f16":
        # binary16 psi + ADE state: measured envelope (fields ~1 %, Te ~5 %, switched
        # fraction within 0.5 % of the film); the film-mean bar of 1e-3 is not met
        assert e["relL2_E"] < 5e-2 and e["relL2_H"] < 2e-2
        assert e["Te_peak_rel"] < 0.1
        assert abs(e["switched"] - e["switched_ref"]) < 5e-3
    else:
        # BF16 psi + ADE state (the form of the 512 x 512 x 256 mixed-storage workload): the
        # 8-bit mantissa of P and J drives the ADE recursion off (measured in the oracle's
        # restatement of this storage: E 0.66, H 0.13, Te peak +123 %, film-mean |dm| 1.23,
        # 68.5 % switched against 6.3 %); held to that envelope, far outside the film bar
        assert e["relL2_E"] < 1.0 and e["relL2_H"] < 0.3
        assert e["Te_peak_rel"] < 2.0
        assert e["mean_dm_m_TM_x"] > 1e-3 and e["switched"] > 0.3
