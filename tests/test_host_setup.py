"""Host-side preparation of a coupled-step state (CPU)."""
import math

import numpy as np
import pytest


def test_position_map_matches_reference_semantics():
    from mphip.spec import position_map
    cells = np.array([3, 7, 11, 20, 42])
    idx = np.array([42, 3, 11])
    ref = {int(v): n for n, v in enumerate(idx)}
    want = [ref.get(int(c), -1) for c in cells]
    assert position_map(cells, idx).tolist() == want
    assert position_map(cells, np.zeros(0, int)).tolist() == [-1] * 5


def test_rasterize_painting_and_fill():
    from mphip import geometry, grid
    g = grid.Grid3D(grid.uniform_axis_n(8, 1.0), grid.uniform_axis_n(8, 1.0), grid.uniform_axis_n(8, 1.0))
    sc = geometry.Scene()
    sc.add(geometry.Box(0, 8, 0, 8, 2.0, 4.5), geometry.Material("a", epsr=4.0))         # z: cells 2,3 full, 4 half
    sc.add(geometry.Box(0, 8, 0, 8, 4.0, 5.0), geometry.Material("f", epsr=1.0, model=True))
    geo = geometry.rasterize(sc, g)
    assert geo.epsr[0, 0, 2] == 4.0 and geo.epsr[0, 0, 1] == 1.0
    # cell 4: first painted 0.5*4 + 0.5*1 = 2.5, then the film overwrites with f=1 -> 1.0
    assert geo.epsr[0, 0, 4] == 1.0
    assert geo.n_material == 64
    k = geo.material_cells // 64
    assert np.all(k == 4) and np.all(geo.material_fill == 1.0)
    assert np.all(np.diff(geo.material_cells) > 0)


def test_state_spec_small_coupled():
    from mphip import configs
    spec, meta = configs.small_coupled(dtype="f32", steps=10, n=24)
    assert spec.field_dtype == np.float32 and spec.compute_dtype == np.float32
    assert spec.n_film == 24 * 24 * 2
    # parameters carried at Float32 precision, widened for the FP64 multiphysics
    arr = spec.gd.as_array(np.float64)
    assert np.array_equal(arr.astype(np.float32).astype(np.float64), arr)
    assert spec.diag_poles.dtype == np.float32 and len(spec.diag_poles) == 6
    assert all(p.tolist() == list(range(spec.n_film)) for p in spec.pos)
    assert spec.lut[0].dtype == np.float32 and spec.lut[4] == 1800
    # absolute 1/(eps0 eps_r) volumes
    assert spec.inv_eps[0].max() == np.float32(1.0 / 8.8541878128e-12)
    assert spec.has_ade_work()


def test_graded_axis_and_pump_config():
    from mphip import configs
    ax = configs.film_graded_axis(64)
    w = ax.widths()
    assert ax.n == 64 and np.isclose(w.min(), 2e-9, rtol=1e-6, atol=0) and np.isclose(w.max(), 20e-9, rtol=1e-6, atol=0)
    assert np.sum(np.isclose(w, 2e-9, rtol=1e-6, atol=0)) == 10
    spec, meta = configs.coupled_pump(nx=24, ny=20, nz=64, steps=100, npml=4)
    assert spec.mo is not None and spec.mo_pair == ("x", "y")
    assert spec.n_film == 24 * 20 * 10
    assert len(meta["planes"]) == 6 and len(meta["omega_bins"]) == 4
    assert spec.source.pulse.amplitude > 0 and math.isfinite(spec.source.pulse.amplitude)


def test_plane_wave_gain_is_physical():
    """A soft plane source adds v each step; the launched wave grows with 1/(Courant)."""
    from mphip import configs, grid
    dz = 20e-9
    g1 = configs.plane_wave_gain(0.5 * dz / grid.C0, dz)
    g2 = configs.plane_wave_gain(0.25 * dz / grid.C0, dz)
    assert 0.5 < g1 < 10 and g2 > 1.5 * g1


def test_roofline_bulk_bytes():
    from mphip import configs, roofline
    spec, _ = configs.small_coupled(dtype="f32", steps=10, n=24)
    kb = roofline.kernel_bytes(spec)
    n_upd = 23 ** 3
    assert kb["yee_H"] >= 36 * n_upd and kb["yee_E"] >= 60 * n_upd
    assert kb["yee_E"] - 60 * n_upd == kb["yee_H"] - 36 * n_upd   # same psi slab traffic


def test_roofline_pingpong_bytes():
    """The fused interior launch reads E, H, D and the material byte and writes H, D, E once
    per cell: 73 bytes with the index (84 with the volumes), 12 fewer than the two interior
    launches, which the H and E totals lose on the interior box.  The box of the 96 x 68 x 64
    pump with CPML 6: x groups 8..15 (32 cells), rows 6..61, planes 6..57."""
    from mphip import configs, roofline
    spec, _ = configs.coupled_pump(nx=96, ny=68, nz=64, steps=10, npml=6, src_z=40e-9)
    n = roofline.interior_cells(spec)
    assert n == 32 * 56 * 52
    for idx, per in ((True, 73), (False, 84)):
        a = roofline.kernel_bytes(spec, idx)
        b = roofline.kernel_bytes(spec, idx, pingpong=True)
        assert b["yee_HE"] == per * n
        assert b["yee_H"] + b["yee_HE"] + b["yee_E"] == a["yee_H"] + a["yee_E"] - 12 * n
        assert b["pingpong_sync"] == 3 * 2 * 4 * (68 * 64 + 96 * 64 + 96 * 68)


def test_measured_figures_follow_the_kernel_sources(monkeypatch):
    """Counter figures are reported only for the kernel sources they were measured on: an
    entry whose digest is not the current sources' gives (None, None)."""
    from mphip import roofline
    cur_y = roofline.source_digest(roofline.YEE_SOURCES)
    cur_f = roofline.source_digest(roofline.FILM_SOURCES)
    assert len(cur_y) == 16 and cur_y != cur_f
    key = ("w", "plain", "index")
    monkeypatch.setitem(roofline.MEASURED_TRAFFIC, key, {"digest": cur_y, "profile": "p.md", "bytes": {"yee_E": 7}})
    assert roofline.measured_traffic("w", "plain", "yee_E", True) == (7, "p.md")
    assert roofline.measured_traffic("w", "plain", "yee_E", False) == (None, None)
    assert roofline.measured_traffic("w", "plain", "yee_H", True) == (None, None)
    monkeypatch.setitem(roofline.MEASURED_TRAFFIC, key, {"digest": "0" * 16, "profile": "p.md", "bytes": {"yee_E": 7}})
    assert roofline.measured_traffic("w", "plain", "yee_E", True) == (None, None)
    monkeypatch.setitem(roofline.FP64_OPS, "w", {"digest": cur_f, "ops": 3.0, "profile": "q.md"})
    assert roofline.fp64_ops_per_cell_substep("w") == (3.0, "q.md")
    monkeypatch.setitem(roofline.FP64_OPS, "w", {"digest": cur_y, "ops": 3.0, "profile": "q.md"})
    assert roofline.fp64_ops_per_cell_substep("w") == (None, None)


def test_em_stencil_config_small():
    from mphip import configs
    spec, meta = configs.em_stencil(n=64, npml=8, steps=10, n_boxes=6, sphere_radius=6)
    assert spec.ade is not None and not spec.has_film
    assert meta["n_ade"] == len(spec.ade[0].idx) > 0
    assert spec.inv_eps[0].min() < spec.inv_eps[0].max()


def test_eps_index_reproduces_volumes():
    """The compact inv_eps form (uint8 index + triple table) gives back every cell's three
    values exactly, on the pump geometry and on the random-box EM geometry; more than 256
    distinct triples falls back to the volumes."""
    import numpy as np
    from mphip import configs
    from mphip.spec import eps_index
    for spec in (configs.coupled_pump(nx=64, ny=48, nz=64, steps=10)[0],
                 configs.em_stencil(n=64, npml=8, steps=10)[0]):
        idx, tab = eps_index(spec.inv_eps)
        assert idx.dtype == np.uint8 and tab.shape[0] <= 256
        for c in range(3):
            assert np.array_equal(tab[idx, c], spec.inv_eps[c])
    rng = np.random.default_rng(0)
    many = [rng.random(4096).astype(np.float32) for _ in range(3)]
    assert eps_index(many) is None
