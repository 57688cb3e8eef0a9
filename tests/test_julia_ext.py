"""The Julia package extension against the C-ABI it binds (no Julia toolchain here).

- Every `struct Mp*` mirror in ext/MagnetoPhotonicHIPExt.jl has the field names and the C
  layout (offsets and size, from the Julia field types under C alignment) of the struct it
  mirrors in include/mphip.h, compiled with gcc.
- Every `ccall` names a function the header declares, with as many argument types as the
  prototype has parameters.
- The extension adds a method for every backend hook of src/core/Backend.jl:19-126 the
  reference calls on device arrays, for every `_ka_*` launcher of src/fdtd/Kernels.jl:
  522-699, for the plane-source injection (src/fdtd/Source.jl:188-206) and for the step
  (`_step_typed!`, Solver.jl:283) and `run!` (Solver.jl:557).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXT = os.path.join(ROOT, "magnetophotonic.jl_amd", "ext", "MagnetoPhotonicHIPExt.jl")
HEADER = os.path.join(ROOT, "include", "mphip.h")

# (size, alignment) of the Julia scalar types the mirrors use, as C lays them out
SCALARS = {"Int64": (8, 8), "Int32": (4, 4), "Cdouble": (8, 8), "Float64": (8, 8), "Cfloat": (4, 4)}


def _ext_text():
    return open(EXT).read()


def julia_structs():
    """{name: [(field, type), ...]} of the `struct Mp*` blocks."""
    out = {}
    for m in re.finditer(r"^struct (Mp\w+)\n(.*?)^end", _ext_text(), re.M | re.S):
        fields = []
        for line in m.group(2).split("\n"):
            for part in line.split(";"):
                part = part.strip()
                if part and "::" in part:
                    name, typ = part.split("::", 1)
                    fields.append((name.strip(), typ.strip()))
        out[m.group(1)] = fields
    return out


def _layout(structs, typ):
    """(size, alignment) of a Julia field type."""
    if typ in SCALARS:
        return SCALARS[typ]
    if typ.startswith("Ptr{"):
        return 8, 8
    m = re.fullmatch(r"NTuple\{(\d+),(.*)\}", typ)
    if m:
        s, a = _layout(structs, m.group(2).strip())
        return s * int(m.group(1)), a
    if typ in structs:
        offs, size, align = _offsets(structs, typ)
        return size, align
    raise KeyError(typ)


def _offsets(structs, name):
    off, align, out = 0, 1, []
    for fname, typ in structs[name]:
        s, a = _layout(structs, typ)
        off = (off + a - 1) // a * a
        out.append((fname, off))
        off += s
        align = max(align, a)
    return out, (off + align - 1) // align * align, align


def _c_name(jname):
    return "mp_" + re.sub(r"(?<!^)(?=[A-Z])", "_", jname[2:]).lower()


def test_struct_mirrors_match_header(tmp_path):
    structs = julia_structs()
    assert {"MpYee", "MpCpml", "MpAde", "MpMo", "MpFilm", "MpLut", "MpMlp", "MpPlane", "MpPulse", "MpSource",
            "MpState"} <= set(structs)
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    expect = []
    for j in sorted(structs):
        c = _c_name(j)
        offs, size, _ = _offsets(structs, j)
        lines.append(f'printf("%zu\\n", sizeof({c}));')
        expect.append((j, "sizeof", size))
        for fname, off in offs:
            if fname.startswith("_pad"):
                continue
            lines.append(f'printf("%zu\\n", offsetof({c}, {fname}));')
            expect.append((j, fname, off))
    lines.append("return 0;}")
    src = tmp_path / "jl_layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "jl_layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)   # a missing field fails to compile
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    bad = [(e, g) for e, g in zip(expect, got) if e[2] != g]
    assert not bad, bad


def _c_prototypes():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"^(?:int|double|void \*|const char \*)\s*(mp_\w+)\((.*?)\);", txt, re.M | re.S):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return protos


def _ccalls():
    """(name, number of argument types) of every ccall into libmphip."""
    txt = _ext_text()
    out = []
    for m in re.finditer(r"ccall\(\(:(mp_\w+), libmphip\),\s*[\w{}]+,\s*\(", txt):
        i, depth, start = m.end(), 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(txt[i], 0)
            i += 1
        inner = txt[start:i - 1]
        level, n, seen = 0, 0, False
        for ch in inner:
            if ch in "({[":
                level += 1
            elif ch in ")}]":
                level -= 1
            elif ch == "," and level == 0:
                n += 1
                continue
            if not ch.isspace() and level >= 0 and ch != ",":
                seen = True
        trailing = inner.strip().endswith(",")
        out.append((m.group(1), (n + (0 if trailing else 1)) if seen else 0))
    return out


def test_ccalls_match_prototypes():
    protos = _c_prototypes()
    calls = _ccalls()
    assert len(calls) >= 25
    for name, nargs in calls:
        assert name in protos, name
        assert nargs == protos[name], (name, nargs, protos[name])


# The backend hooks of src/core/Backend.jl the reference calls with device arrays
# (zeros/adapt/to_host/synchronize/memory hooks :19-126; the reductions and host
# downloads :49-83), and its `_ka_*` launchers (src/fdtd/Kernels.jl:522-699).
HOOKS = ("zeros_backend", "array_type", "adapt_backend", "to_host", "synchronize", "device_memory_info",
         "reclaim_device_memory!", "free_device!", "reduce_sum", "reduce_mean", "reduce_count_negative",
         "reduce_norm_mean", "any_nonfinite", "scalar_to_host", "plane_to_host")
LAUNCHERS = ("_ka_update_H_3d!", "_ka_update_E_3d!", "_ka_inject_soft_3d!", "_ka_inject_mode_x_3d!",
             "_ka_patch_E_dispersive!", "_ka_patch_E_mo_gyration!", "_ka_pabs_accumulate!",
             "_ka_multiphysics_step!", "_ka_probe_dft_accumulate!", "_ka_probe_trace_plane!")


@pytest.mark.parametrize("name", HOOKS + LAUNCHERS + ("inject!", "_step_typed!", "run!"))
def test_extension_defines_method(name):
    txt = _ext_text()
    pat = r"^(?:function )?MagnetoPhotonic\." + re.escape(name) + r"\("
    assert re.search(pat, txt, re.M), name


def test_device_index_arrays_round_trip_one_based():
    """Index arrays go to the device 0-based (absent 0 -> -1) and come back 1-based."""
    txt = _ext_text()
    assert "_upload_index(x::AbstractArray{<:Integer}) = HIPIndexArray(_upload(Int32.(x) .- Int32(1)))" in txt
    assert "MagnetoPhotonic.to_host(x::HIPIndexArray) = Int.(Array(x.data)) .+ 1" in txt
