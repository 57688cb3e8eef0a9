"""Parity at the full workload sizes (GPU against the oracle, a few steps).

The 384^3 EM stencil (64 random dielectric boxes, a Drude-Lorentz sphere of radius 40
cells, CPML 16, plane source), the 256 x 256 x 192 and the 512 x 512 x 256 coupled pump
(graded z axis, film with 4TM + LLB, gyration, 6 DFT planes x 4 bins).  Both sides start
from the same seeded random E, H and D volumes, so every cell, every CPML slab and every
dispersive list entry does real arithmetic from the first step.  Fields must agree bit for
bit; film temperatures to 1e-6 relative L2 and magnetisation to 1e-4 (FP64 exp differs in
the last ULP between the two math libraries); DFT accumulators to 1e-12 relative L2.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")


def _seed_fields(gpu, ref, shape, seed):
    rng = np.random.default_rng(seed)
    for name, scale in (("ex", 1e3), ("ey", 1e3), ("ez", 1e3), ("hx", 3.0), ("hy", 3.0), ("hz", 3.0)):
        v = (rng.standard_normal(shape, dtype=np.float32) * np.float32(scale))
        gpu.set_field(name, v)
        ref.set_field(name, v)
    for name, e in (("dx", "ex"), ("dy", "ey"), ("dz", "ez")):
        v = ref.field(e).astype(np.float64) * 8.854187817e-12
        gpu.set_field(name, v.astype(np.float32))
        ref.set_field(name, v.astype(np.float32))


def _compare(gpu, ref, film=True):
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.array_equal(a, b), (name, int(np.count_nonzero(a != b)))
    if film:
        fa, fb = gpu.film_state(), ref.film_state()
        for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
            d = np.linalg.norm(fa[k].astype(np.float64) - fb[k]) / np.linalg.norm(fb[k])
            assert d < 1e-6, k
        for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
            assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k


@pytest.mark.timeout(900)
def test_em_stencil_384_full_size():
    import driver
    from mphip import configs
    from mphip.state import FDTDState
    spec, meta = configs.em_stencil(n=384, npml=16, steps=5000)
    print(f"case em_stencil 384^3 ({meta['n_ade']} ADE cells), 3 steps", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    _seed_fields(gpu, ref, spec.shape, 1)
    gpu.run(3)
    ref.run(3)
    gpu.synchronize()
    _compare(gpu, ref, film=False)
    for c in range(3):
        g, r = gpu.ade_state(c), ref.ade_arr[c]
        assert np.array_equal(g["P"], r["P"]) and np.array_equal(g["J"], r["J"]), c
    gpu.free()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("shape", [(256, 256, 192), (512, 512, 256)])
def test_coupled_pump_full_size(shape):
    import driver
    from mphip import configs
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    from test_gpu_parity import OracleDFT, rel_l2
    nx, ny, nz = shape
    spec, meta = configs.coupled_pump(nx=nx, ny=ny, nz=nz, steps=20000)
    print(f"case coupled_pump {nx}x{ny}x{nz} ({spec.n_film} film cells), 3 steps", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    _seed_fields(gpu, ref, spec.shape, 2)
    mon = DFTPlanes(gpu, meta["planes"], meta["omega_bins"])
    gpu.attach_dft(mon)
    ref_mon = OracleDFT(ref, meta["planes"], meta["omega_bins"])
    gpu.run(3)
    for _ in range(3):
        ref.step()
        ref_mon.record(ref)
    gpu.synchronize()
    _compare(gpu, ref)
    for p in range(len(meta["planes"])):
        g, r = mon.download(p), ref_mon.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            assert rel_l2(g[k], r[k]) < 1e-12, (p, k)
    gpu.free()
