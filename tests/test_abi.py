"""The C-ABI library loads, exports every declared entry point, and its struct layouts
match the ctypes mirrors; argument errors are reported before any device work (CPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mphip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|double|void \*|const char \*)\s*(mp_\w+)\(", txt, re.M)))


def test_libmphip_exports_every_declared_symbol():
    from mphip import _lib
    lib = _lib.lib()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_mirrors_kernel_entry_points():
    import driver
    from mphip import abi
    L = driver.lib(abi)
    for name in abi.SIGNATURES:
        assert hasattr(L, "orc_" + name), name


STRUCTS = {"mp_yee": "Yee", "mp_cpml": "Cpml", "mp_ade": "Ade", "mp_mo": "Mo", "mp_film": "Film",
           "mp_lut": "Lut", "mp_plane": "Plane", "mp_pulse": "Pulse", "mp_source": "Source",
           "mp_state": "State", "mp_mlp": "Mlp"}


def test_struct_layouts_match_header(tmp_path):
    from mphip import abi, models
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    expect = []
    for cname, pyname in STRUCTS.items():
        cls = getattr(abi, pyname)
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(cls))
        for fname, _ in cls._fields_:
            if fname.startswith("_pad"):
                continue
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(cls, fname).offset)
    lines.append('printf("%zu\\n", sizeof(mp_ferri_params));')
    expect.append(8 * len(models.FIELDS))
    for macro, value in (("MP_MLP_NPARAM", abi.MLP_NPARAM), ("MP_MLP_NNORM", abi.MLP_NNORM),
                         ("MP_MLP_PACKED_BYTES", abi.MLP_PACKED_BYTES)):
        lines.append(f'printf("%zu\\n", (size_t){macro});')
        expect.append(value)
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect


def test_pulse_values_match_host():
    from mphip import _lib, abi, sources
    lib = _lib.lib()
    cases = [(abi.MP_PULSE_GAUSSIAN, sources.GaussianPulse(3.2e8, 1.9e-14, 4.2e-14, 2.35e15, 0.1)),
             (abi.MP_PULSE_CARRIER_GAUSSIAN, sources.CarrierGaussianPulse(2.0, 5e-14, 1e-14, 3.1e15)),
             (abi.MP_PULSE_CONTINUOUS, sources.ContinuousSource(1.5, 2.2e15, 0.3))]
    for kind, pl in cases:
        p = abi.Pulse()
        p.kind, p.amplitude, p.omega = kind, pl.amplitude, pl.omega
        p.t0 = getattr(pl, "t0", 0.0)
        p.tau = getattr(pl, "tau", getattr(pl, "sigma", 0.0))
        p.phase = getattr(pl, "phase", 0.0)
        for t in np.linspace(0.0, 2e-13, 57):
            assert lib.mp_pulse_value(C.byref(p), float(t)) == pl.value(float(t))


def test_argument_errors_before_device_work():
    """Host-side validation returns MP_ERR_ARG with a message and launches nothing."""
    from mphip import _lib, abi
    lib = _lib.lib()
    y = abi.Yee()
    y.nx, y.ny, y.nz = 8, 8, 8
    cp = abi.Cpml()
    cp.npml_x = 5   # 8 < 2*5+1
    for d in range(3):
        cp.inv_kappa[d] = cp.a[d] = cp.b[d] = 8
    assert lib.mp_update_H_3d(C.byref(y), C.byref(cp), 1e-17, 1.2566e-6) == -1
    assert b"CPML width" in lib.mp_last_error()
    assert lib.mp_inject_soft_3d(C.byref(y), 3, 8, 0, 0, 1.0) == -1
    assert b"out of range" in lib.mp_last_error()
    assert lib.mp_inject_mode_x_3d(C.byref(y), 1, 0, C.c_void_p(8), 1.0, 0.0) == -1
    assert b"Ey or Ez" in lib.mp_last_error()
    f = abi.Film()
    f.n, f.dtype = 4, abi.MP_F32
    lut = abi.Lut()
    lut.n, lut.dtype, lut.tm, lut.re = 10, abi.MP_F32, 8, 8
    prm = np.zeros(97)
    assert lib.mp_multiphysics_step(C.byref(f), prm.ctypes.data_as(C.c_void_p), C.byref(lut), 1e-17, 0, 2, 1) == -1
    assert b"subcycles must be positive" in lib.mp_last_error()
    assert lib.mp_probe_trace_plane(C.byref(y), 3, 0, 0, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8)) == -1
    pl = abi.Plane()
    pl.dft_ea = pl.dft_eb = pl.dft_ha = pl.dft_hb = pl.energy = 8
    v8 = C.c_void_p(8)
    assert lib.mp_probe_shot_reduce(C.byref(pl), C.byref(pl), 4, 4, 0, 0, v8, v8, v8) == -1
    assert b"nb must be" in lib.mp_last_error()
    assert lib.mp_probe_shot_reduce(C.byref(pl), C.byref(pl), 4, 4, 5, 5, v8, v8, v8) == -1
    assert b"center_bin" in lib.mp_last_error()
    assert lib.mp_probe_shot_reduce(C.byref(pl), C.byref(pl), 0, 4, 5, 2, v8, v8, v8) == -1
    pl.energy = None
    assert lib.mp_probe_shot_reduce(C.byref(pl), C.byref(pl), 4, 4, 5, 2, v8, v8, v8) == -1
    assert b"null plane buffer" in lib.mp_last_error()
    assert lib.mp_plane_overlap(C.byref(y), 6, 0, 0, None, v8) == -1
    assert b"component" in lib.mp_last_error()
    assert lib.mp_plane_overlap(C.byref(y), 1, 2, 8, None, v8) == -1
    assert b"out of range" in lib.mp_last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_init_fails_loudly_without_gpu():
    from mphip import _lib
    with pytest.raises(_lib.MPError):
        _lib._initialised = False
        _lib.init(0)
