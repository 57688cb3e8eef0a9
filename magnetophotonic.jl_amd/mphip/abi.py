"""ctypes mirror of include/mphip.h (struct layouts and the C-ABI function table).

The layouts are shared by anything that calls a library exporting this ABI; the loader
for libmphip itself lives in `_lib.py`.
"""
from __future__ import annotations

import ctypes as C

MP_F32, MP_F64, MP_BF16, MP_F16 = 0, 1, 2, 3
MP_YEE_AUTO, MP_YEE_WHOLE, MP_YEE_SCALAR = 0, 1, 2
MP_MAX_PLANES = 8

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double


class Yee(C.Structure):
    _fields_ = [("nx", i64), ("ny", i64), ("nz", i64), ("field_dtype", i32), ("compute_dtype", i32),
                ("ex", vp), ("ey", vp), ("ez", vp), ("hx", vp), ("hy", vp), ("hz", vp),
                ("dx", vp), ("dy", vp), ("dz", vp),
                ("inv_eps_x", vp), ("inv_eps_y", vp), ("inv_eps_z", vp),
                ("inv_d_cell_x", vp), ("inv_d_cell_y", vp), ("inv_d_cell_z", vp),
                ("inv_d_dual_x", vp), ("inv_d_dual_y", vp), ("inv_d_dual_z", vp),
                ("eps_index", vp), ("eps_table", vp), ("n_eps", i32), ("kernel_form", i32)]


class Cpml(C.Structure):
    _fields_ = [("npml_x", i32), ("npml_y", i32), ("npml_z", i32), ("psi_dtype", i32),
                ("inv_kappa", vp * 3), ("a", vp * 3), ("b", vp * 3),
                ("psi_h", vp * 12), ("psi_d", vp * 12), ("psi_scale", f64)]


class Ade(C.Structure):
    _fields_ = [("n", i64), ("npoles", i32), ("dtype", i32), ("state_dtype", i32), ("_pad", i32),
                ("state_scale", f64), ("P", vp), ("J", vp), ("E_old", vp), ("idx", vp),
                ("fill", vp), ("inv_eps", vp), ("poles", vp)]


class Mo(C.Structure):
    _fields_ = [("n", i64), ("npoles", i32), ("dtype", i32),
                ("Pa_from_b", vp), ("Ja_from_b", vp), ("Pb_from_a", vp), ("Jb_from_a", vp),
                ("E_old_a", vp), ("E_old_b", vp), ("idx", vp), ("fill", vp), ("mpos", vp),
                ("poles", vp)]


class Film(C.Structure):
    _fields_ = [("n", i64), ("dtype", i32), ("_pad", i32), ("cells", vp),
                ("pos_x", vp), ("pos_y", vp), ("pos_z", vp),
                ("Te", vp), ("Tl", vp), ("Ts_TM", vp), ("Ts_RE", vp),
                ("m_TM_x", vp), ("m_TM_y", vp), ("m_TM_z", vp),
                ("m_RE_x", vp), ("m_RE_y", vp), ("m_RE_z", vp),
                ("U_abs", vp), ("window_sum", vp), ("peak", vp)]


class Lut(C.Structure):
    _fields_ = [("tm", vp), ("re", vp), ("T_min", f64), ("inv_dT", f64), ("n", i32), ("dtype", i32)]


class Plane(C.Structure):
    _fields_ = [("axis", i32), ("_pad", i32), ("index", i64), ("dft_ea", vp), ("dft_eb", vp),
                ("dft_ha", vp), ("dft_hb", vp), ("energy", vp)]


class Pulse(C.Structure):
    _fields_ = [("kind", i32), ("_pad", i32), ("amplitude", f64), ("t0", f64), ("tau", f64),
                ("omega", f64), ("phase", f64)]


class Source(C.Structure):
    _fields_ = [("kind", i32), ("component", i32), ("axis", i32), ("_pad", i32),
                ("i", i64), ("j", i64), ("k", i64), ("i0", i64), ("i1", i64), ("w0", f64), ("w1", f64),
                ("profile", vp), ("hcorr", f64), ("pulse", Pulse)]


class Mlp(C.Structure):
    _fields_ = [("params", vp), ("packed", vp), ("norm", vp)]


MLP_NPARAM, MLP_NNORM, MLP_PACKED_BYTES = 6288, 80, 22528


class State(C.Structure):
    _fields_ = [("yee", C.POINTER(Yee)), ("cpml", C.POINTER(Cpml)), ("ade", C.POINTER(Ade) * 3),
                ("mo", C.POINTER(Mo)), ("mo_ea", vp), ("mo_eb", vp), ("mo_iea", vp), ("mo_ieb", vp),
                ("film", C.POINTER(Film)), ("params", vp), ("lut", C.POINTER(Lut)), ("source", Source),
                ("planes", C.POINTER(Plane)), ("nplanes", i32), ("nb", i32), ("omega_bins", vp),
                ("phase", vp), ("dt", f64), ("mu0", f64), ("pabs_factor", f64), ("Q_TM", f64),
                ("Q_RE", f64), ("t", f64), ("n", i64), ("last_mp_n", i64),
                ("multiphysics_every", i32), ("subcycles", i32), ("brillouin_iters", i32),
                ("multiphysics_enabled", i32), ("has_ade_work", i32), ("record_dft", i32),
                ("surrogate", C.POINTER(Mlp)), ("surrogate_tol", f64), ("surrogate_fallbacks", vp),
                ("film_lists_coincide", i32), ("film_serial", i32), ("yee_alt", vp * 6)]


MP_PULSE_GAUSSIAN, MP_PULSE_CARRIER_GAUSSIAN, MP_PULSE_CONTINUOUS = 1, 2, 3
MP_SRC_NONE, MP_SRC_SOFT, MP_SRC_PLANE, MP_SRC_MODE_X = 0, 1, 2, 3

P = C.POINTER
SIGNATURES = {
    "update_H_3d": (C.c_int, [P(Yee), P(Cpml), f64, f64]),
    "update_E_3d": (C.c_int, [P(Yee), P(Cpml), f64]),
    "inject_soft_3d": (C.c_int, [P(Yee), i32, i64, i64, i64, f64]),
    "inject_mode_x_3d": (C.c_int, [P(Yee), i32, i64, vp, f64, f64]),
    "inject_plane_3d": (C.c_int, [P(Yee), i32, i32, i64, f64, i64, f64, f64]),
    "patch_E_dispersive": (C.c_int, [vp, P(Ade), i32, f64]),
    "patch_E_mo_gyration": (C.c_int, [vp, vp, P(Mo), vp, vp, i32, vp, vp, f64, f64, f64]),
    "pabs_accumulate": (C.c_int, [P(Film), P(Yee), P(Ade), P(Ade), P(Ade), i32, f64, f64]),
    "film_epilogue": (C.c_int, [P(Yee), P(Ade), P(Ade), P(Ade), P(Mo), vp, vp, vp, vp, P(Film), i32, f64, f64,
                                f64, f64]),
    "multiphysics_step": (C.c_int, [P(Film), vp, P(Lut), f64, i32, i32, i32]),
    "probe_dft_accumulate": (C.c_int, [P(Yee), P(Plane), i32, vp, vp, i32, f64, f64]),
    "probe_trace_plane": (C.c_int, [P(Yee), i32, i64, i64, vp, vp, vp]),
    "flux_plane": (C.c_int, [P(Yee), i32, i64, vp, vp, i64]),
    "probe_shot_reduce": (C.c_int, [P(Plane), P(Plane), i64, i64, i32, i32, vp, vp, vp]),
    "plane_overlap": (C.c_int, [P(Yee), i32, i32, i64, vp, vp]),
    "film_reduce": (C.c_int, [P(Film), vp]),
    "film_average": (C.c_int, [P(Film), vp]),
    "switching_metrics": (C.c_int, [P(Film), f64, vp]),
    "last_error": (C.c_char_p, []),
}

# Device reductions and plane copies behind the backend hooks (checked against numpy).
MP_RED_SUM, MP_RED_COUNT_NEGATIVE, MP_RED_NORM_SUM = 0, 1, 2
HOOK_SIGNATURES = {
    "reduce": (C.c_int, [vp, vp, vp, i32, i64, i32, vp]),
    "sum_blocks": (C.c_int, [vp, i64, i64, vp]),
    "plane_gather": (C.c_int, [vp, i32, i64, i64, i64, i32, i64, vp]),
    "plane_abs": (C.c_int, [vp, i32, i64, i64, i64, i32, i64, vp]),
    "probe_trace_samples": (C.c_int, [P(Yee), i64, i64, vp, vp, vp]),
    "probe_frames": (C.c_int, [P(Yee), i64, i64, i64, vp, vp, vp, vp, vp, vp]),
}

# No-allocation variants of entry points (fixed library scratch); no oracle counterpart.
NOALLOC_SIGNATURES = {
    "probe_shot_reduce_noalloc": (C.c_int, [P(Plane), P(Plane), i64, i64, i32, i32, vp, vp, vp]),
}
MP_SHOT_NOALLOC_MAX_BINS = 63

DRIVER_SIGNATURES = {
    "run_steps": (C.c_int, [P(State), i64]),
    "pulse_value": (C.c_double, [P(Pulse), f64]),
}

SURROGATE_SIGNATURES = {
    "mp_mlp_pack": (C.c_int, [P(Mlp)]),
    "mp_mlp_forward": (C.c_int, [P(Mlp), vp, vp, i64]),
    "mp_mlp_backward": (C.c_int, [P(Mlp), vp, vp, i64, vp, vp]),
    "mp_mlp_adam": (C.c_int, [vp, vp, vp, vp, i64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "mp_surrogate_dataset": (C.c_int, [P(Film), vp, P(Lut), f64, i32, i32, i32, i64, vp, vp, i64]),
    "mp_surrogate_step": (C.c_int, [P(Film), vp, P(Lut), P(Mlp), f64, i32, i32, i32, f64, vp]),
}

RUNTIME_SIGNATURES = {
    "mp_init": (C.c_int, [C.c_int]),
    "mp_malloc": (C.c_int, [P(vp), C.c_size_t]),
    "mp_free": (C.c_int, [vp]),
    "mp_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
    "mp_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
    "mp_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
    "mp_memset": (C.c_int, [vp, C.c_int, C.c_size_t]),
    "mp_fill": (C.c_int, [vp, i32, i64, f64]),
    "mp_synchronize": (C.c_int, []),
    "mp_mem_info": (C.c_int, [P(C.c_size_t), P(C.c_size_t)]),
    "mp_set_stream": (C.c_int, [vp]),
    "mp_get_stream": (vp, []),
    "mp_event_create": (C.c_int, [P(vp)]),
    "mp_event_destroy": (C.c_int, [vp]),
    "mp_event_record": (C.c_int, [vp]),
    "mp_event_elapsed_ms": (C.c_int, [vp, vp, P(C.c_float)]),
    "mp_profile_enable": (C.c_int, [C.c_int]),
    "mp_kernel_stats": (C.c_int, [C.c_char_p, P(i64), P(f64)]),
    "mp_profile_select": (C.c_int, [C.c_char_p]),
    "mp_profile_sample": (C.c_int, [C.c_int32]),
    "mp_any_nonfinite": (C.c_int, [vp, i32, i64, vp]),
}


def bind(lib, prefix: str, table: dict) -> None:
    """Attach restype/argtypes for every `prefix + name` entry of `table`."""
    for name, (res, args) in table.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
