"""HDF5 output (mphip/h5out.py, ext/MagnetoPhotonicHDF5Ext.jl) on the CPU.

The files are checked with a second, independent HDF5 reader where the image has one: h5py
under the Anaconda interpreter at /opt/conda (its own Python and numpy), in a subprocess.
Those checks pin the on-disk convention the reference's HDF5.jl files have: a Julia array of
size (d1, ..., dn) is an HDF5 dataset of dims (dn, ..., d1) over the same bytes, strings
are variable-length UTF-8 scalars, zero-size arrays have a null dataspace, movies are
chunked one slice per frame.  The states come from the C oracle (small coupled grids).
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from mphip import configs, h5out

pytestmark = pytest.mark.skipif(not h5out.available(), reason="no HDF5 C library in this image")

H5PY_PYTHON = "/opt/conda/bin/python3.9"


def _h5py(script: str, *args):
    """Run `script` under the /opt/conda interpreter with h5py; it prints one JSON value."""
    if not os.path.exists(H5PY_PYTHON):
        pytest.skip("no independent HDF5 reader (/opt/conda python with h5py)")
    env = {k: v for k, v in os.environ.items() if k not in ("PYTHONPATH", "PYTHONHOME")}
    env["PYTHONNOUSERSITE"] = "1"
    r = subprocess.run([H5PY_PYTHON, "-c", script, *args], capture_output=True, text=True, env=env, timeout=120)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("h5py missing from /opt/conda")
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout)


_DUMP = r"""
import json, sys, h5py, numpy as np
out = {}
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        sd = h5py.check_string_dtype(obj.dtype)
        e = {"shape": None if obj.shape is None else list(obj.shape), "chunks": None if obj.chunks is None else list(obj.chunks),
             "kind": obj.dtype.kind, "itemsize": obj.dtype.itemsize,
             "string": None if sd is None else [sd.encoding, sd.length]}
        if sd is not None:
            e["value"] = obj[()].decode("utf-8")
        elif obj.shape is not None and obj.size <= 4096:
            v = obj[()]
            e["value"] = np.asarray(v, dtype=np.float64).tolist()
        out[name] = e
with h5py.File(sys.argv[1], "r") as f:
    f.visititems(visit)
print(json.dumps(out))
"""


def _oracle_state(steps=4, n=24):
    import driver
    spec, meta = configs.small_coupled(dtype="f32", steps=60, amplitude=2e8, n=n)
    ref = driver.OracleState(spec)
    ref.run(steps)
    return ref


def test_values_round_trip_and_layout(tmp_path):
    p = str(tmp_path / "v.h5")
    a3 = np.arange(2 * 3 * 4, dtype=np.float32).reshape((2, 3, 4), order="F")
    a2 = np.asarray([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])          # Julia (3, 2)
    vals = {"i": 7, "f": 2.5, "f32": np.float32(1.25), "flag": np.int8(1), "s": "dt in µm", "tup_i": (1, 2, 3),
            "tup_f": (1.0, 2.5), "none": None, "v": np.linspace(0.0, 1.0, 5), "m": a2, "c": a3,
            "empty": np.zeros(0), "b": True}
    with h5out.H5File(p, "w") as f:
        for k, v in vals.items():
            f.put(f"g/sub/{k}", v)
    with h5out.H5File(p, "r") as f:
        assert f.read("g/sub/i") == 7 and f.read("g/sub/i").dtype == np.int64
        assert f.read("g/sub/f") == 2.5
        assert f.read("g/sub/f32").dtype == np.float32
        assert f.read("g/sub/flag").dtype == np.int8
        assert f.read("g/sub/s") == "dt in µm"
        assert f.read("g/sub/none") == "nothing"
        assert list(f.read("g/sub/tup_i")) == [1, 2, 3]
        np.testing.assert_array_equal(f.read("g/sub/m"), a2)
        np.testing.assert_array_equal(f.read("g/sub/c"), a3)
        assert f.read("g/sub/empty").size == 0
        assert sorted(f.keys("g/sub")) == sorted(vals)
    d = _h5py(_DUMP, p)
    # HDF5 dims are the Julia dims reversed, over the same (column-major) bytes
    assert d["g/sub/c"]["shape"] == [4, 3, 2]
    assert np.array_equal(np.asarray(d["g/sub/c"]["value"]).T, a3)
    assert d["g/sub/m"]["shape"] == [2, 3]
    assert d["g/sub/i"]["kind"] == "i" and d["g/sub/i"]["itemsize"] == 8 and d["g/sub/i"]["shape"] == []
    assert d["g/sub/flag"]["itemsize"] == 1
    assert d["g/sub/f32"]["kind"] == "f" and d["g/sub/f32"]["itemsize"] == 4
    assert d["g/sub/s"]["string"] == ["utf-8", None] and d["g/sub/s"]["value"] == "dt in µm"
    assert d["g/sub/empty"]["shape"] is None       # null dataspace
    assert d["g/sub/tup_f"]["value"] == [1.0, 2.5]


def _payloads(state, n_ts=7):
    Na = state.spec.n_film
    fs = state.film_state()
    t = np.arange(n_ts) * 1e-15
    pump = {"mag_time": t, "Te_avg": np.linspace(300, 900, n_ts), "mx_TM": np.full(n_ts, 0.9)}
    active = {"active_linear_index": np.asarray(state.spec.film_cells, np.int64) + 1,
              "m_TM_x_reduced_active_cells": np.asarray(fs["m_TM_x"], np.float64),
              "final_m_TM_x_reduced_active": -np.asarray(fs["m_TM_x"], np.float64),
              "final_m_RE_x_reduced_active": -np.asarray(fs["m_RE_x"], np.float64),
              "final_switch_fraction": 0.25, "hot_cell_index": 3, "hot_cell_Te_K": np.linspace(300, 400, n_ts)}
    return pump, active, Na


def test_production_file(tmp_path):
    st = _oracle_state()
    nx, ny, nz = st.spec.shape
    pump, active, Na = _payloads(st)
    p = str(tmp_path / "prod.h5")
    h5out.write_production_h5(p, st, pump=pump, relaxation=pump, pump_active=active, shot_active=active,
                              movies={"frame_count": 4}, metadata={"pml_x": 3, "steps_relax": 12})
    sig = {r["path"]: r for r in h5out.h5_schema_signature(p)}
    # groups of the reference's production file
    assert {k.split("/")[0] for k in sig} == {"metadata", "pump", "relaxation", "shot_1", "probe",
                                              "phase2_resume_finalization"}
    assert sig["metadata/Nx"]["dtype"] == "Int64" and sig["metadata/Nx"]["shape"] == ()
    assert sig["metadata/eps_z_xy"]["shape"] == (nx, ny)
    assert sig["metadata/eps_z_xz"]["shape"] == (nx, nz)
    assert sig["metadata/eps_z_yz_30"]["shape"] == (ny, nz)
    assert sig["metadata/save_pump_frames"]["dtype"] == "Int8"
    assert sig["metadata/Q_voigt_TM"]["dtype"] == "Float32"      # the compute-type model parameters
    assert sig["metadata/dt"]["dtype"] == "Float64"
    assert sig["pump/Ez_xy"]["shape"] == (nx, ny, 4) and sig["pump/Ez_xy"]["dtype"] == "Float32"
    assert sig["pump/Ez_yz_10"]["shape"] == (ny, nz, 4)
    assert sig["pump/Te_avg"]["shape"] == (7,)
    assert sig["pump/active_linear_index"]["shape"] == (Na,) and sig["pump/active_linear_index"]["dtype"] == "Int64"
    assert sig["shot_1/final_m_TM_x_reduced_active"]["shape"] == (Na,)
    assert sig["shot_1/hot_cell_Te_K"]["shape"] == (7,)
    assert sig["probe/probe_diag_poles_C1_C2_C3"]["shape"][1] == 3
    assert sig["probe/reference/omega_bins_rad_s"]["shape"] == (5,)
    with h5out.H5File(p, "r") as f:
        np.testing.assert_array_equal(f.read("pump/Te_avg"), pump["Te_avg"])
        assert np.isnan(f.read("pump/Tl_avg")).all()               # absent series: NaN, mag_time's length
        np.testing.assert_allclose(f.read("metadata/x_coords"), st.spec.grid.x.centers * 1e6)
        eps = np.asarray(st.spec.epsr, np.float64).reshape((nx, ny, nz), order="F")
        np.testing.assert_array_equal(f.read("metadata/eps_z_xy"), eps[:, :, (nz + 1) // 2 - 1])
        assert f.read("metadata/absorption_model") == st.spec.absorption_model
        assert f.read("phase2_resume_finalization/steps_relax") == 12
        assert float(f.read("phase2_resume_finalization/final_switch_fraction")) == 0.25
    d = _h5py(_DUMP, p)
    assert d["pump/Ez_xy"]["shape"] == [4, ny, nx] and d["pump/Ez_xy"]["chunks"] == [1, ny, nx]
    assert d["metadata/Nx"]["value"] == nx
    assert d["metadata/absorption_model"]["string"][0] == "utf-8"


def test_production_magnetization_round_trip(tmp_path):
    """load_production_magnetization reads shot_1/final_m_* back onto the film cells by the
    1-based linear index; cells the file lacks keep the metadata's as-grown moments."""
    st = _oracle_state()
    pump, active, Na = _payloads(st)
    idx = np.asarray(active["active_linear_index"])
    keep = np.arange(Na) % 3 != 0                                  # leave every third cell out
    act = dict(active, active_linear_index=idx[keep])
    for k in ("final_m_TM_x_reduced_active", "final_m_RE_x_reduced_active"):
        act[k] = np.asarray(active[k])[keep]
    for k in ("final_m_TM_y_reduced_active", "final_m_TM_z_reduced_active", "final_m_RE_y_reduced_active",
              "final_m_RE_z_reduced_active"):
        act[k] = np.full(int(keep.sum()), 0.125)
    p = str(tmp_path / "prod.h5")
    h5out.write_production_h5(p, st, pump=pump, relaxation=pump, pump_active=act, shot_active=act,
                              metadata={"initial_m_TM_x_reduced": 0.75, "initial_m_RE_x_reduced": -0.5})
    m = h5out.load_production_magnetization(p, st)
    assert m["matched"] == int(keep.sum())
    np.testing.assert_array_equal(m["m_TM_x"][keep], act["final_m_TM_x_reduced_active"])
    np.testing.assert_array_equal(m["m_RE_x"][keep], act["final_m_RE_x_reduced_active"])
    assert (m["m_TM_y"][keep] == 0.125).all() and (m["m_TM_y"][~keep] == 0).all()
    assert (m["m_TM_x"][~keep] == 0.75).all() and (m["m_RE_x"][~keep] == -0.5).all()


def _shot(label, nb=5, nt=9, fc=0, dims=None, rng=None):
    rng = rng or np.random.default_rng(0)
    s = {"A": 0.1, "R": 0.2, "T": 0.7, "T_plus_R_plus_A": 1.0, "theta_faraday_deg": 0.5, "steps_probe": 40,
         "theta_kerr_deg": 0.25, "ellipticity_faraday_deg": 0.1, "ellipticity_kerr_deg": 0.05, "group_delay_fs": 1.5,
         "pulse_broadening_fs": 0.5,
         "probe_frame_count": fc, "state_label": label, "is_reference": 1 if label == "reference" else 0,
         "T_omega": rng.random(nb), "R_omega": rng.random(nb), "omega_bins_rad_s": np.linspace(3e15, 4e15, nb),
         "Ez_trace_pre": rng.random(nt), "trace_time_s": np.arange(nt) * 1e-16,
         "jones_kerr_Ey_re_im": [0.25, -0.5], "x_pre_um": 0.1, "x_post_um": 0.3}
    if fc:
        s["probe_frame_indices"] = list(range(10, 10 + 10 * fc, 10))
        s["probe_frame_times_s"] = [i * 1e-15 for i in range(fc)]
        s["frames"] = {k: rng.random(dims[k] + (fc,)).astype(np.float32) for k in dims}
    return s


def test_goldstd_file_and_appended_shot(tmp_path):
    st = _oracle_state()
    nx, ny, nz = st.spec.shape
    dims = {k: ((nx, ny) if k.endswith("_xy") else (ny, nz)) for k in h5out.PROBE_MOVIE_KEYS}
    shots = {"reference": _shot("reference"), "initial": _shot("initial", fc=3, dims=dims),
             "switched": _shot("switched", fc=3, dims=dims, rng=np.random.default_rng(1))}
    p = str(tmp_path / "gold.h5")
    h5out.write_goldstd_h5(p, shots, state=st)
    sig = {r["path"]: r for r in h5out.h5_schema_signature(p)}
    assert sig["probe/initial/Ey_xy"]["shape"] == (nx, ny, 3)
    assert sig["probe/switched/Ez_yz_post"]["shape"] == (ny, nz, 3)
    assert sig["probe/initial/eps_xy"]["shape"] == (nx, ny)
    assert sig["probe/reference/is_reference"]["dtype"] == "Int8"
    assert "probe/reference/Ey_xy" not in sig                      # no movies for the reference shot
    assert sig["probe/initial/probe_frame_indices"]["dtype"] == "Int64"
    assert sig["probe/contrast/delta_T"]["shape"] == ()
    assert sig["probe/reference/R_omega"]["shape"] == (5,)
    assert sig["probe/reference/Sx_incident_re"]["shape"] == ()     # absent vector: empty (null dataspace)
    with h5out.H5File(p, "r") as f:
        np.testing.assert_array_equal(f.read("probe/switched/Ez_xy"), shots["switched"]["frames"]["Ez_xy"])
        assert float(f.read("probe/contrast/delta_T")) == 0.0
        assert f.read("probe/initial/state_label") == "initial"
        assert np.isnan(float(f.read("probe/reference/throughput_loss")))   # absent number: NaN
    h5out.append_probe_shot_h5(p, "extra", _shot("extra"), state=st)
    with h5out.H5File(p, "r") as f:
        assert f.read("probe/extra/state_label") == "extra"
    with pytest.raises(h5out.H5Error, match="already exists"):
        h5out.append_probe_shot_h5(p, "extra", _shot("extra"))
    d = _h5py(_DUMP, p)
    assert d["probe/initial/Ey_xy"]["shape"] == [3, ny, nx]


def test_frame_writers_on_host_planes(tmp_path):
    """The movie datasets of both writers: (d1, d2, nf) float32 chunked one slice per frame;
    frame f of the file is the f-th recorded plane; unrecorded frames read 0; frames past
    frame_count are dropped."""
    st = _oracle_state()
    nx, ny, nz = st.spec.shape
    p = str(tmp_path / "frames.h5")
    w = h5out.production_frame_writer(p, st, frame_count=3)
    rng = np.random.default_rng(2)
    rec = []
    for _ in range(2):
        planes = {k: rng.random(w.dims[k]).astype(np.float32) for k in w.KEYS}
        rec.append(planes)
        w.write_frame(planes)
    w.close()
    pw = h5out.probe_frame_writer(p, st, group="probe/initial", frame_count=2, pre_x=2, post_x=20)
    prec = []
    for _ in range(3):
        planes = {k: rng.random(pw.dims[k]).astype(np.float32) for k in pw.KEYS}
        prec.append(planes)
        pw.write_frame(planes)
    pw.close()
    with h5out.H5File(p, "r") as f:
        for k in w.KEYS:
            cube = f.read(f"pump/{k}")
            assert cube.shape == w.dims[k] + (3,)
            for i in range(2):
                np.testing.assert_array_equal(cube[:, :, i], rec[i][k])
            assert (cube[:, :, 2] == 0).all()
        for k in pw.KEYS:
            cube = f.read(f"probe/initial/{k}")
            assert cube.shape == pw.dims[k] + (2,)
            for i in range(2):
                np.testing.assert_array_equal(cube[:, :, i], prec[i][k])
    from mphip.sources import source_index
    assert w.kz == (nz + 1) // 2 - 1 and w.jy == source_index(st.spec.grid.y, 0.0)
    d = _h5py(_DUMP, p)
    assert d["pump/Ez_xz"]["shape"] == [3, nz, nx] and d["pump/Ez_xz"]["chunks"] == [1, nz, nx]
    assert d["probe/initial/Ey_yz_pre"]["chunks"] == [1, nz, ny]
    # the production file written afterwards without overwrite keeps the streamed movies
    pump, active, _ = _payloads(st)
    h5out.write_production_h5(p, st, pump=pump, relaxation=pump, active=active, movies={"frame_count": 3},
                              overwrite=False)
    with h5out.H5File(p, "r") as f:
        np.testing.assert_array_equal(f.read("pump/Ez_xy")[:, :, 1], rec[1]["Ez_xy"])


def test_schema_compatibility_check(tmp_path):
    st = _oracle_state()
    pump, active, _ = _payloads(st)
    a, b = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    h5out.write_production_h5(a, st, pump=pump, relaxation=pump, active=active)
    h5out.write_production_h5(b, st, pump=pump, relaxation=pump, active=active)
    assert h5out.assert_h5_schema_compatible(a, b)
    # a missing dataset, a shape and an element type that differ
    c = str(tmp_path / "c.h5")
    h5out.write_production_h5(c, st, pump=dict(pump, mag_time=np.zeros(3)), relaxation=pump, active=active)
    with pytest.raises(h5out.H5Error, match="shape"):
        h5out.assert_h5_schema_compatible(c, a)
    assert h5out.assert_h5_schema_compatible(c, a, require_shapes=False)
    with h5out.H5File(b, "r+") as f:
        f.put("extra/only_here", 1.0)
    with pytest.raises(h5out.H5Error, match="missing datasets"):
        h5out.assert_h5_schema_compatible(a, b)
    x, y = str(tmp_path / "x.h5"), str(tmp_path / "y.h5")
    h5out.write_production_h5(x, st, pump=pump, relaxation=pump, active=active)
    h5out.write_production_h5(y, st, pump=pump, relaxation=pump, active=active)
    with h5out.H5File(y, "r+") as f:
        f.put("extra2/v", np.int64(1))
    with h5out.H5File(x, "r+") as f:
        f.put("extra2/v", 1.5)
    with pytest.raises(h5out.H5Error, match="dtype"):
        h5out.assert_h5_schema_compatible(x, y, require_shapes=False)
    assert h5out.assert_h5_schema_compatible(x, y, require_types=False)


def test_magnetization_and_state_dump(tmp_path):
    st = _oracle_state()
    p = str(tmp_path / "mag.h5")
    h5out.save_magnetization(p, st)
    m = h5out.load_magnetization(p)
    fs = st.film_state()
    np.testing.assert_array_equal(m["m_TM_x"], fs["m_TM_x"])
    assert m["m_TM_x"].dtype == np.float32
    np.testing.assert_array_equal(m["material_cells"], np.asarray(st.spec.film_cells) + 1)   # 1-based
    q = str(tmp_path / "state.h5")
    h5out.save_hdf5_state(q, st)
    s = h5out.load_hdf5_state(q)
    np.testing.assert_array_equal(s["state"]["fields"]["Ez"], st.field("ez"))
    assert s["state"]["fields"]["Ez"].shape == st.spec.shape
    assert s["state"]["compute_T"] == "Float32" and s["state"]["region"]["n_material"] == st.spec.n_film
    np.testing.assert_array_equal(s["state"]["thermal"]["Te"], fs["Te"])
    assert s["provenance"]["package"] == "MagnetoPhotonic"
