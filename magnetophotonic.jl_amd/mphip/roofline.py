"""Algorithmic (compulsory) HBM bytes of each kernel of the coupled step.

Each array a kernel touches is counted once per read and once per write at its storage
width; halo re-reads and cache effects are excluded.  These are the per-launch figures
behind the roofline fractions the benchmark reports.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from .cpml import PSI_AXIS

_SZ = {"f32": 4, "f64": 8, "bf16": 2, "f16": 2}


def _psi_cells(spec) -> int:
    """Cells of the H (or E) update range that carry one psi term, summed over the 12
    psi slabs: every slab cell is read and written once per step."""
    nx, ny, nz = spec.shape
    w = spec.cpml.widths
    tot = 0
    for s in range(12):
        a = PSI_AXIS[s]
        # each slab (lo or hi) holds w cells along its axis; the update range is n-1 wide
        # along the other two axes
        other = [(nx - 1), (ny - 1), (nz - 1)]
        other[a] = w[a]
        tot += other[0] * other[1] * other[2]
    return tot


def interior_cells(spec) -> int:
    """Cells of the interior box the split Yee launches give their interior kernels: whole
    float4 groups off every CPML slab and boundary cell, x edges on 128-byte lines."""
    nx, ny, nz = spec.shape
    mx, my, mz = (max(int(w), 1) for w in spec.cpml.widths)
    vx = 4
    gi0 = ((mx + vx - 1) // vx + 7) // 8 * 8
    ngi = max((nx - mx) // vx // 8 * 8 - gi0, 0)
    return ngi * vx * max(ny - 2 * my, 0) * max(nz - 2 * mz, 0)


def kernel_bytes(spec, eps_index: bool = False, fused_epilogue: bool = False, pingpong: bool = False) -> dict:
    """Algorithmic bytes per launch of each kernel of one step.  eps_index: the E update
    reads inv_eps as a 1-byte material index per cell instead of three volumes.
    fused_epilogue: the ADE, gyration and absorbed-power list work runs as one pass over
    the film cells (E read and written once, J summed in registers, no position maps)."""
    nx, ny, nz = spec.shape
    fs = spec.field_dtype.itemsize
    cs = spec.compute_dtype.itemsize
    ps = _SZ[spec.psi_dtype]
    n_upd = (nx - 1) * (ny - 1) * (nz - 1)
    out = {}
    # H: read E(3), H(3); write H(3).  E: read H(3), D(3), inv_eps(3); write D(3), E(3).
    e_cell = (12 * fs + 1) if eps_index else 15 * fs
    out["yee_H"] = 9 * fs * n_upd + 2 * ps * _psi_cells(spec)
    out["yee_E"] = e_cell * n_upd + 2 * ps * _psi_cells(spec)
    if pingpong:
        # the interior as one fused launch (read E, H, D and the material, write H, D, E);
        # the H and E launches take the CPML shell only; E on the three planes no update
        # writes is carried to the other buffer set each step
        n_int = interior_cells(spec)
        out["yee_HE"] = (6 * fs + e_cell) * n_int      # the E update's H read is the H just formed
        out["yee_H"] -= 9 * fs * n_int
        out["yee_E"] -= e_cell * n_int
        out["pingpong_sync"] = 3 * 2 * fs * (ny * nz + nx * nz + nx * ny)
    if spec.ade is not None:
        np_ = len(spec.diag_poles) // 3
        ss = _SZ[spec.ade_state_dtype]
        per = 4 + 2 * cs + 2 * cs + 2 * fs + 2 * 2 * np_ * ss      # idx, fill+inv, E_old r/w, E r/w, P,J r/w
        out["ade_patch"] = sum(per * len(a.idx) for a in spec.ade)
    if spec.mo is not None:
        np_ = len(spec.mo_poles) // 3
        per = 4 + 4 + cs + 2 * cs + 4 * fs + 4 * cs + 2 * fs + 4 * 2 * np_ * cs
        out["mo_gyration"] = per * len(spec.mo.idx)
    if spec.has_film:
        nf = spec.n_film
        np_ = (len(spec.diag_poles) // 3) if spec.ade is not None else 0
        ss = _SZ[spec.ade_state_dtype]
        work = (12 + 3 * np_ * ss) if spec.has_ade_work() else 0
        out["pabs"] = nf * (4 + 3 * fs + work + 6 * 8)
        out["multiphysics"] = nf * (20 * cs + 2 * 8)
        if fused_epilogue and spec.ade is not None:
            per = 4 + 6 * fs + 3 * (4 * cs + 4 * np_ * ss) + 6 * 8   # cell, E r/w, 3 ADE entries, U/window/peak
            if spec.mo is not None:
                npm = len(spec.mo_poles) // 3
                per += 3 * cs + 4 * cs + 8 * npm * cs + 2 * fs       # fill, m x2, E_old r/w x2, P/J, inv_eps x2
            for k in ("ade_patch", "mo_gyration", "pabs"):
                out.pop(k, None)
            out["film_epilogue"] = per * nf
    return out


def plane_bytes(spec, planes, nb: int) -> int:
    """DFT record: per plane cell 4 tangential fields read, 4 complex accumulators read
    and written per bin, the flux integral read and written."""
    nx, ny, nz = spec.shape
    fs = spec.field_dtype.itemsize
    dims = {0: ny * nz, 1: nz * nx, 2: nx * ny}
    return sum(dims[a] * (4 * fs + 4 * nb * 32 + 16) for a, _ in planes)


def plane_cells(spec, planes) -> int:
    nx, ny, nz = spec.shape
    dims = {0: ny * nz, 1: nz * nx, 2: nx * ny}
    return sum(dims[a] for a, _ in planes)


def step_bytes(spec, planes=(), nb: int = 0, eps_index: bool = False, dft_batch: int = 1,
               fused_epilogue: bool = False, pingpong: bool = False) -> dict:
    """Algorithmic bytes per step.  With dft_batch > 1 the DFT record is a per-step
    snapshot (4 fields read, 4 written) plus, once per batch, the stored steps read and
    the accumulators read and written once."""
    kb = kernel_bytes(spec, eps_index, fused_epilogue, pingpong)
    if planes and nb:
        if dft_batch > 1:
            fs = spec.field_dtype.itemsize
            cells = plane_cells(spec, planes)
            kb["dft_snapshot"] = cells * 8 * fs
            kb["dft_batch"] = cells * (4 * fs + 2 * (4 * nb * 16 + 8) // dft_batch)
        else:
            kb["dft_accumulate"] = plane_bytes(spec, planes, nb)
    kb["total"] = int(sum(v for v in kb.values()))
    return kb


# Counter measurements of the kernels, each valid only for the kernel sources it was taken
# on: `digest` is source_digest() of those files at the time, and a lookup whose current
# digest differs returns nothing, so a figure from older kernels is never reported as one
# of the running build.
YEE_SOURCES = ("yee.hip", "common.h")
FILM_SOURCES = ("thermo.hip", "film_physics.h", "common.h")

_CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc")


def compile_flags() -> bytes:
    """The compiler and flag definitions of csrc/Makefile (HIPCC, ARCH, HIPFLAGS with its
    continuation lines): what, besides the sources, decides the kernels' code."""
    out, cont = [], False
    with open(os.path.join(_CSRC, "Makefile"), "rb") as f:
        for line in f:
            s = line.strip()
            if cont or s.startswith((b"HIPCC", b"ARCH", b"HIPFLAGS")):
                out.append(s)
                cont = s.endswith(b"\\")
    return b"\n".join(out)


def source_digest(names) -> str:
    """First 16 hex digits of the sha256 of the named csrc files, concatenated in order, and
    the compile flags."""
    h = hashlib.sha256()
    for n in names:
        with open(os.path.join(_CSRC, n), "rb") as f:
            h.update(f.read())
    h.update(compile_flags())
    return h.hexdigest()[:16]


# HBM bytes per launch from rocprofv3 PMC runs (FETCH_SIZE x 2.0 + WRITE_SIZE, KiB per
# dispatch; the factors are what the counters read for a plain 1 GiB device copy on gfx950).
MEASURED_TRAFFIC = {
    ("coupled_pump_512x512x256", "plain", "index"): {
        "digest": "6465cab51f2a17f7", "profile": "profiles/coupled_pump_512x512x256_plain_pmc.md",
        "bytes": {"yee_E": 4133446194, "yee_H": 3247387641}},
    ("em_stencil_384^3", "plain", "index"): {
        "digest": "6465cab51f2a17f7", "profile": "profiles/em_stencil_384_pmc.md",
        "bytes": {"yee_E": 3635337602, "yee_H": 2865344936}},
}

# FP64 operations of one film cell-substep of the 4TM + LLB update from SQ_INSTS_VALU_*_F64
# counters (2 x FMA + ADD + MUL, x 64 lanes, per dispatch / (cells x substeps)).
FP64_OPS = {
    "coupled_pump_512x512x256": {"digest": "d7025c771f2d0f28", "ops": 1143.4,
                                 "profile": "profiles/coupled_pump_512x512x256_plain_pmc.md"},
}


def measured_traffic(workload: str, form: str, kernel: str, eps_index: bool = False):
    """(bytes per launch of `kernel`, profile file) measured on the current Yee sources, or
    (None, None) where none was."""
    m = MEASURED_TRAFFIC.get((workload, form, "index" if eps_index else "volumes"))
    if not m or kernel not in m["bytes"] or m["digest"] != source_digest(YEE_SOURCES):
        return None, None
    return m["bytes"][kernel], m["profile"]


def fp64_ops_per_cell_substep(workload: str):
    """(FP64 operations per film cell-substep, profile file) measured on the current film
    sources, or (None, None)."""
    m = FP64_OPS.get(workload)
    if not m or m["digest"] != source_digest(FILM_SOURCES):
        return None, None
    return m["ops"], m["profile"]
