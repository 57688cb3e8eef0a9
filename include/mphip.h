/*
 * mphip.h -- C-ABI of libmphip, the MI355X (gfx950) backend for MagnetoPhotonic.jl's
 * coupled 3-D time step.
 *
 * Every entry point replaces one launcher of the reference's kernel layer
 * (src/fdtd/Kernels.jl `_ka_*` functions, the per-backend override points) or one
 * backend hook of src/core/Backend.jl.  The reference file:line each one replaces is
 * given beside it.  Conventions:
 *
 *  - Plain pointers and sizes only.  Device pointers are `void*` obtained from
 *    mp_malloc (or any HIP allocation on the current device).
 *  - Volumes are column-major with x fastest, as Julia arrays are: cell (i,j,k),
 *    0-based, lives at i + nx*(j + ny*k).  The reference is 1-based; every index
 *    crossing this ABI is 0-based.
 *  - Cell lists use int32 linear indices (grids up to 2^31 cells); the reference
 *    keeps Int64 on the host.
 *  - Composite Julia isbits types map to C structs with identical field order:
 *    DLPole{T} -> {T C1, C2, C3}; FerrimagnetParameters{Float64} -> mp_ferri_params.
 *  - Every entry returns 0 on success and a negative code on failure; the message is
 *    available from mp_last_error() (thread-local).  No C++ exception crosses the ABI.
 *  - All launches are asynchronous on the library stream (mp_set_stream); nothing
 *    synchronises unless the call says so.  The library never keeps caller pointers
 *    past a call.  One host thread drives a stream; calls are not re-entrant per stream.
 */
#ifndef MPHIP_H
#define MPHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- dtypes ---------------------------------------------------------------------- */
enum mp_dtype {
    MP_F32 = 0,   /* Float32 */
    MP_F64 = 1,   /* Float64 */
    MP_BF16 = 2,  /* bfloat16 storage (FP32 arithmetic) */
    MP_F16 = 3    /* IEEE half storage with a power-of-two scale (FP32 arithmetic) */
};

/* Field component codes: src/fdtd/Kernels.jl:1-7 (_component_code). */
enum mp_component { MP_EX = 1, MP_EY = 2, MP_EZ = 3, MP_HX = 4, MP_HY = 5, MP_HZ = 6 };

enum mp_status {
    MP_OK = 0,
    MP_ERR_ARG = -1,      /* bad argument (ArgumentError / DimensionMismatch in the reference) */
    MP_ERR_HIP = -2,      /* HIP runtime error */
    MP_ERR_UNSUPPORTED = -3
};

/* ---- data model -------------------------------------------------------------------- */

/* Yee volumes and 1-D metric arrays of an FDTDState (src/fdtd/Solver.jl:9-53,78-94;
 * src/fdtd/Fields.jl:1-11).  inv_eps_* are ABSOLUTE 1/(eps0*eps_r) (Solver.jl:87-94),
 * stored in field_dtype; inv_d_* are the CT copies of Axis1D.inv_d_cell/inv_d_dual
 * (Solver.jl:81-86, src/grid/Grid.jl:24-38). */
typedef struct mp_yee {
    int64_t nx, ny, nz;
    int32_t field_dtype;    /* storage of E, H, D, inv_eps: MP_F32 or MP_F64 */
    int32_t compute_dtype;  /* arithmetic type CT: MP_F32 or MP_F64 */
    void *ex, *ey, *ez, *hx, *hy, *hz, *dx, *dy, *dz;
    const void *inv_eps_x, *inv_eps_y, *inv_eps_z;
    const void *inv_d_cell_x, *inv_d_cell_y, *inv_d_cell_z;
    const void *inv_d_dual_x, *inv_d_dual_y, *inv_d_dual_z;
    /* Optional compact copy of inv_eps for the FP32 E update: eps_index holds one byte
     * per cell (volume layout) naming an entry of eps_table, n_eps <= 256 entries of 4
     * FP32 (inv_eps_x, inv_eps_y, inv_eps_z, unused).  The entries are the volumes' own
     * values, so the update is the same; it reads 1 byte per cell instead of 12.  NULL /
     * 0 reads the volumes.  The volumes stay required (sources, gyration, other paths). */
    const uint8_t *eps_index;
    const float *eps_table;
    int32_t n_eps;
    /* Kernel family of the FP32 updates (enum mp_yee_form): MP_YEE_AUTO picks the
     * interior/shell split where it applies; the other values force one form so that each
     * can be checked against the others.  Every form computes the same values. */
    int32_t kernel_form;
} mp_yee;

enum mp_yee_form {
    MP_YEE_AUTO = 0,    /* interior + CPML-shell launches where the box is non-empty */
    MP_YEE_WHOLE = 1,   /* one vector launch over the whole volume */
    MP_YEE_SCALAR = 2   /* the scalar (one cell per lane) kernels */
};

/* CFS-CPML state (src/fdtd/CPML.jl:79-189).  Per-axis coefficient vectors inv_kappa,
 * a, b (length n of that axis, CT) and the 24 convolution-memory slabs: an x-slab is
 * (w, ny, nz), a y-slab (nx, w, nz), a z-slab (nx, ny, w) (CPML.jl:183-188).
 * psi_h / psi_d order: xy_lo, xy_hi, xz_lo, xz_hi, yz_lo, yz_hi, yx_lo, yx_hi,
 * zx_lo, zx_hi, zy_lo, zy_hi (psi_<component><derivative axis>_<side>, CPML.jl:133-146).
 * npml_* = 0 disables that axis (CPML.jl:151 _cpml_width).  For MP_BF16 / MP_F16
 * storage the slab holds psi * psi_scale. */
typedef struct mp_cpml {
    int32_t npml_x, npml_y, npml_z;
    int32_t psi_dtype;
    const void *inv_kappa[3], *a[3], *b[3];
    void *psi_h[12];
    void *psi_d[12];
    double psi_scale;
} mp_cpml;

/* One component's Drude-Lorentz ADE list (src/fdtd/Dispersion.jl:93-113;
 * Solver.jl:125-169).  P, J are (n, npoles) column-major: P[m + p*n]. */
typedef struct mp_ade {
    int64_t n;
    int32_t npoles;
    int32_t dtype;        /* CT: type of E_old, fill, inv_eps, poles and the arithmetic */
    int32_t state_dtype;  /* storage of P and J: CT, or MP_BF16 / MP_F16 (mixed form) */
    int32_t _pad;
    double state_scale;   /* stored = value * state_scale for BF16/F16 storage */
    void *P, *J, *E_old;
    const int32_t *idx;   /* 0-based linear cell index */
    const void *fill, *inv_eps;   /* per active cell, CT; inv_eps = 1/(eps0*eps) */
    const void *poles;            /* npoles x {C1, C2, C3}, CT (DLPole{CT}) */
} mp_ade;

/* Magneto-optic gyration ADE state (src/fdtd/MagnetoOptic.jl:1-39).  The reference
 * hard-wires the pair (a, b) = (Ey, Ez) driven by m_x (Kernels.jl:233-305); here the
 * pair is whatever two E volumes the caller passes.  mpos maps list position -> film
 * cell (-1 = none). */
typedef struct mp_mo {
    int64_t n;
    int32_t npoles;
    int32_t dtype;  /* CT */
    void *Pa_from_b, *Ja_from_b, *Pb_from_a, *Jb_from_a;  /* (n, npoles) */
    void *E_old_a, *E_old_b;
    const int32_t *idx;
    const void *fill;
    const int32_t *mpos;
    const void *poles;
} mp_mo;

/* Film (material-cell) state: ThermalState, MagnetizationState, AbsorptionState and
 * the region maps (src/physics/Thermal.jl:5-61, src/physics/Magnetization.jl:265-284,
 * Solver.jl:114-159).  pos_* map a film cell to its position in each ADE list (-1 =
 * inactive there), the reference's active_axis_position_map (ReferenceRaster.jl:218). */
typedef struct mp_film {
    int64_t n;
    int32_t dtype;  /* CT: storage of Te, Tl, Ts_*, m_* */
    int32_t _pad;
    const int32_t *cells;
    const int32_t *pos_x, *pos_y, *pos_z;
    void *Te, *Tl, *Ts_TM, *Ts_RE;
    void *m_TM_x, *m_TM_y, *m_TM_z, *m_RE_x, *m_RE_y, *m_RE_z;
    double *U_abs, *window_sum, *peak;   /* always FP64 */
} mp_film;

/* FerrimagnetParameters{T} field order (src/physics/Models.jl:5-103), as FP64. */
typedef struct mp_ferri_params {
    double gamma_e, Cl, Cs_TM, Cs_RE, Gel, Ges_TM, Ges_RE, Gamma_E;
    double Hsw0, Tsw, dTsw, gamma_gyro, Ms_TM, Ms_RE, mu_TM, mu_RE;
    double alpha0_TM, alpha0_RE, J_TMRE, J_TMTM, J_RERE, K_TM, K_RE;
    double T_Curie, T_comp, T0, TcT, TcR, inv_TcT, inv_TcR;
    double beta_T, beta_R, floor_T, floor_R, val300_T, val300_R, inv_val300_T, inv_val300_R;
    double GammaT0, GammaT_hot, GammaR0, GammaR_hot, Tsw_long, dTsw_long, inv_dTsw_long;
    double ellT, ellR, inv_ellT, inv_ellR, GammaEx, eta_quench, inv_eta_quench;
    double GammaT_tail, eta_tail, inv_eta_tail, mTM_tail_target, field_to_K;
    double GTlT, GTlR, GsTR, S_TM, S_RE, Q_voigt_TM, Q_voigt_RE;
    double omega_pump, eps_imag_pump, eps_real_pump, omega_probe, eps_imag_probe, eps_real_probe;
    double kb, moment_ratio_TM_RE, inv_Ms_sum;
    double inv_T_Curie, inv_Cl, inv_Cs_TM, inv_Cs_RE;
    double ferri_RE_scale, re_exchange_lag, inv_mu_TM, inv_mu_RE;
    double inv_Ms_TM, inv_Ms_RE, twelve_J_TMTM, twelve_J_RERE;
    double two_K_over_Ms_TM, two_K_over_Ms_RE;
    double inv_tau_TM_demag, inv_tau_RE_demag, moment_ratio_RE_TM;
    double HK_T_field, HK_R_field, HAF_T_field, HAF_R_field, alphaT_hot, alphaR_hot, seed_tilt;
} mp_ferri_params;

/* Brillouin m_eq lookup table (src/physics/Magnetization.jl:50-86), stored in CT. */
typedef struct mp_lut {
    const void *tm, *re;
    double T_min, inv_dT;
    int32_t n;
    int32_t dtype;
} mp_lut;

/* One DFT / flux monitor plane (src/sim/ProbeReadout.jl:11-170, generalised from the
 * reference's two x-normal planes to any axis).  The tangential pair (a, b) is the
 * cyclic pair after the normal axis: x -> (y, z), y -> (z, x), z -> (x, y), so the
 * signed forward flux is Ea*Hb - Eb*Ha (x: Ey*Hz - Ez*Hy, Kernels.jl:447).
 * dft holds four ComplexF64 arrays (Ea, Eb, Ha, Hb), each (n1, n2, nb) column-major
 * with interleaved re/im; energy is (n1, n2); n1, n2 are the plane's two in-plane
 * sizes in axis order (x-plane: ny, nz). */
typedef struct mp_plane {
    int32_t axis;    /* 0 = x, 1 = y, 2 = z (plane normal) */
    int32_t _pad;
    int64_t index;   /* 0-based plane index along the normal axis */
    double *dft_ea, *dft_eb, *dft_ha, *dft_hb;
    double *energy;
} mp_plane;

#define MP_MAX_PLANES 8

/* Source waveform evaluated on the host in FP64 once per step (src/fdtd/Source.jl:1-33). */
enum mp_pulse_kind {
    MP_PULSE_GAUSSIAN = 1,          /* A exp(-((t-t0)/tau)^2) sin(w (t-t0) + phase)  */
    MP_PULSE_CARRIER_GAUSSIAN = 2,  /* A sin(w t) exp(-1/2 ((t-t0)/tau)^2)           */
    MP_PULSE_CONTINUOUS = 3         /* A sin(w t + phase)                             */
};
typedef struct mp_pulse {
    int32_t kind;
    int32_t _pad;
    double amplitude, t0, tau, omega, phase;
} mp_pulse;

/* Source of a state (Solver.jl:296-311): soft point, plane (any axis, two weights) or
 * x-normal mode plane. */
enum mp_source_kind { MP_SRC_NONE = 0, MP_SRC_SOFT = 1, MP_SRC_PLANE = 2, MP_SRC_MODE_X = 3 };
typedef struct mp_source {
    int32_t kind, component, axis, _pad;
    int64_t i, j, k;        /* soft: the cell; mode: plane index in i */
    int64_t i0, i1;         /* plane: cells along `axis` (i1 = -1 for one) */
    double w0, w1;
    const void *profile;    /* mode: (ny, nz) CT */
    double hcorr;
    mp_pulse pulse;
} mp_source;

/* Learned stand-in for the sub-stepped 4TM + LLB film integrator: an MLP 11 -> 64 -> 64
 * -> 6 with SiLU, BF16 operands on MFMA with FP32 accumulation.  The reference has no
 * such model; it replaces the substep loop of multiphysics_kernel! (Kernels.jl:362-412).
 *
 * Features (raw, FP32), from the film state before the update:
 *   Te, Tl, Ts_TM, Ts_RE, m_TM_x, m_RE_x, asinh(pabs / 1e18 W m^-3), log(dt / 1e-18 s),
 *   gamma_e, log(Gel), T_Curie
 * Targets (raw rates over dt, FP32):
 *   (du_e/dt - pabs) with u_e = gamma_e Te^2 / 2, dTl/dt, dTs_TM/dt, dTs_RE/dt,
 *   dm_TM_x/dt, dm_RE_x/dt
 * The network sees x = (f - x_mean) * x_inv_std and predicts y = (t - y_mean) * y_inv_std.
 *
 * FP32 parameter layout (row-major, out x in; padding rows/columns stay zero):
 *   W1 [64][16] @0, b1 [64] @1024, W2 [64][64] @1088, b2 [64] @5184,
 *   W3 [16][64] @5248, b3 [16] @6272.
 * norm (FP32): x_mean[16], x_inv_std[16], y_mean[16], y_std[16], y_inv_std[16]. */
#define MP_MLP_IN 11
#define MP_MLP_HIDDEN 64
#define MP_MLP_OUT 6
#define MP_MLP_NPARAM 6288
#define MP_MLP_NNORM 80
#define MP_MLP_PACKED_BYTES 22528   /* 44 BF16 MFMA fragments of 64 lanes x 4 values */
typedef struct mp_mlp {
    const float *params;   /* MP_MLP_NPARAM FP32 master parameters (device) */
    void *packed;          /* MP_MLP_PACKED_BYTES, written by mp_mlp_pack (device) */
    const float *norm;     /* MP_MLP_NNORM FP32 normalisation constants (device) */
} mp_mlp;

/* Everything one step of an FDTDState touches (Solver.jl:9-53).  Pointers to the
 * component structs above; NULL disables that part (no dispersion, no MO, no film, no
 * monitor).  t, n and last_mp_n are advanced by mp_run_steps. */
typedef struct mp_state {
    mp_yee *yee;
    mp_cpml *cpml;
    mp_ade *ade[3];
    mp_mo *mo;
    void *mo_ea, *mo_eb;
    const void *mo_iea, *mo_ieb;
    mp_film *film;
    const mp_ferri_params *params;
    const mp_lut *lut;
    mp_source source;
    const mp_plane *planes;
    int32_t nplanes, nb;
    const double *omega_bins;
    double *phase;
    double dt, mu0, pabs_factor, Q_TM, Q_RE, t;
    int64_t n, last_mp_n;
    int32_t multiphysics_every, subcycles, brillouin_iters, multiphysics_enabled;
    int32_t has_ade_work, record_dft;
    /* Film update by the surrogate (mp_surrogate_step) instead of the plain integrator
     * when non-NULL; surrogate_fallbacks (device int64, may be NULL) counts the cells
     * that fell back to the plain integrator. */
    const mp_mlp *surrogate;
    double surrogate_tol;
    int64_t *surrogate_fallbacks;
    /* Non-zero when every ADE list and the MO list are the film list (entry n is film
     * cell n) and the three ADE lists share one storage form: the step then runs the list
     * work after the E update as one mp_film_epilogue pass. */
    int32_t film_lists_coincide;
    /* Non-zero runs the film update on the library stream in step order; zero (default)
     * runs it on a side stream beside the next step's Yee updates, joined before the
     * next gyration / absorbed-power kernels, which gives the same values. */
    int32_t film_serial;
    /* Optional second buffer set of the six E and H volumes (ex, ey, ez, hx, hy, hz; same
     * shape and type as the yee's).  When all six are set, the field type and arithmetic
     * are FP32 on the vector kernels with an interior/CPML-shell split and the source is a
     * soft or plane source on an E component (or none), mp_run_steps advances the steps in
     * pairs of ping-pong steps: each reads E and H from one set and writes them to the
     * other, so the interior H and E updates run as one fused launch (73 instead of 85
     * bytes per interior cell).  Each pair ends with the fields back in the yee's volumes;
     * an odd last step runs in place.  The results are those of the in-place sequence bit
     * for bit.  The library keeps no pointer after the call.  NULL keeps the in-place steps.
     * Should a launch fail inside a pair, the call returns its error with the fields of the
     * step in progress possibly in this second set. */
    void *yee_alt[6];
} mp_state;

/* ---- runtime / backend hooks (src/core/Backend.jl:19-126) ---------------------------- */

/* Select the device and create the library state.  Backend.jl:19-33 (__init__ /
 * _GPU_FUNCTIONAL) and ext/MagnetoPhotonicCUDAExt.jl:328-331. */
int mp_init(int device);
/* Thread-local message of the last failing call ("" if none). */
const char *mp_last_error(void);
/* zeros_backend / free_device! (Backend.jl:31-33,126). */
int mp_malloc(void **ptr, size_t bytes);
int mp_free(void *ptr);
/* adapt_backend / to_host (Backend.jl:41-47); synchronous copies. */
int mp_memcpy_h2d(void *dst, const void *src, size_t bytes);
int mp_memcpy_d2h(void *dst, const void *src, size_t bytes);
int mp_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Byte fill on the library stream (fill! of zeros). */
int mp_memset(void *ptr, int value, size_t bytes);
/* Typed fill on the library stream: fill_backend (Backend.jl:37-48). */
int mp_fill(void *ptr, int32_t dtype, int64_t n, double value);
/* synchronize(backend) (Backend.jl:85-95). */
int mp_synchronize(void);
/* device_memory_info (Backend.jl:106-119). */
int mp_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Launch stream for every later call (a hipStream_t, or NULL for the null stream). */
int mp_set_stream(void *stream);
void *mp_get_stream(void);

/* Event timing on the library stream (replaces step_profiling!, Solver.jl:195-252). */
int mp_event_create(void **event);
int mp_event_destroy(void *event);
int mp_event_record(void *event);
int mp_event_elapsed_ms(void *start, void *stop, float *ms);  /* synchronises on stop */

/* Per-kernel timing: when on, each launch is bracketed by events; totals are read with
 * mp_kernel_stats (names are the kernel names below).  Off by default. */
int mp_profile_enable(int on);
int mp_kernel_stats(const char *name, int64_t *launches, double *total_ms);
/* Restrict the per-kernel timing to a comma-separated list of kernel names (NULL or ""
 * times every kernel), so that a timed run brackets only the launches it reports.  No
 * counterpart in the reference (instrumentation of this library). */
int mp_profile_select(const char *names);
/* Time only every `every`-th launch of each timed kernel name, counted from this call
 * (1, the default, times every launch).  Each bracket costs the launch stream two event
 * records, so a timed run can sample its launches instead of delaying each step.  No
 * counterpart in the reference (instrumentation of this library). */
int mp_profile_sample(int32_t every);

/* ---- hot-path kernels ---------------------------------------------------------------- */

/* Yee H update with CFS-CPML psi_H: H -= dt/mu0 * curl E over i<nx-1, j<ny-1, k<nz-1.
 * Replaces _ka_update_H_3d! (src/fdtd/Kernels.jl:522-541; kernel :60-101).
 * dt_mu0 is formed as CT(dt)/CT(mu0) as the launcher does (Kernels.jl:539). */
int mp_update_H_3d(const mp_yee *g, mp_cpml *c, double dt, double mu0);

/* Yee E update with CFS-CPML psi_D: D += dt*curl H, E = D*inv_eps over i,j,k >= 1.
 * Replaces _ka_update_E_3d! (Kernels.jl:543-565; kernel :103-152). */
int mp_update_E_3d(const mp_yee *g, mp_cpml *c, double dt);

/* Soft point source at one cell: E += v, D += v/inv_eps (or H += v), in CT.
 * Replaces _ka_inject_soft_3d! (Kernels.jl:567-576; kernel :154-174). */
int mp_inject_soft_3d(const mp_yee *g, int32_t component, int64_t i, int64_t j, int64_t k,
                      double value);

/* x-normal mode-source plane: E_c += v*profile, D_c += .../inv_eps, H -/+= hcorr*amp.
 * profile is (ny, nz) in CT.  Replaces _ka_inject_mode_x_3d! (Kernels.jl:578-593;
 * kernel :176-192). */
int mp_inject_mode_x_3d(const mp_yee *g, int32_t component, int64_t i, const void *profile,
                        double value, double hcorr);

/* Plane source on any axis with the two-cell linear interpolation weights of
 * _axis_weights (src/fdtd/Source.jl:74-86).  Arithmetic follows the host loop
 * inject!(::FieldState, ::Grid3D, ::PlaneSource) (Source.jl:188-206, 216-236): the
 * increment w*value and D += w*value/inv_eps are formed in FP64 and rounded once into
 * the stored field.  The reference has no device kernel for this (it is a host loop);
 * pass i1 = -1 for a single-cell plane. */
int mp_inject_plane_3d(const mp_yee *g, int32_t component, int32_t axis, int64_t i0, double w0,
                       int64_t i1, double w1, double value);

/* Semi-implicit Drude-Lorentz ADE patch on one component's active list.
 * Replaces _ka_patch_E_dispersive! (Kernels.jl:595-604; kernel :194-231).
 * E is that component's field volume; inv_alpha is formed as CT(2)/CT(dt). */
int mp_patch_E_dispersive(void *E, mp_ade *s, int32_t field_dtype, double dt);

/* Crank-Nicolson gyrotropic exchange between Ea and Eb on the MO list.
 * Replaces _ka_patch_E_mo_gyration! (Kernels.jl:606-619; kernel :233-305).
 * m_TM / m_RE are the film magnetisation components that drive the gyration (m_x in the
 * reference).  inv_eps_a / inv_eps_b are the full volumes (field dtype). */
int mp_patch_E_mo_gyration(void *Ea, void *Eb, mp_mo *s, const void *inv_eps_a,
                           const void *inv_eps_b, int32_t field_dtype, const void *m_TM,
                           const void *m_RE, double Q_voigt_TM, double Q_voigt_RE, double dt);

/* Per-EM-step absorbed power: pabs = max(E.sum(J),0) (ade_work) or pabs_factor*|E|^2;
 * U_abs, window_sum += pabs*dt; peak = max.  FP64.  ax/ay/az may be NULL when
 * has_ade_work == 0.  Replaces _ka_pabs_accumulate! (Kernels.jl:623-646; :313-355). */
int mp_pabs_accumulate(mp_film *f, const mp_yee *g, const mp_ade *ax, const mp_ade *ay,
                       const mp_ade *az, int32_t has_ade_work, double pabs_factor, double dt);

/* The step's film list work after the E update in one pass over the film cells: the ADE
 * patch of Ex, Ey and Ez, the gyrotropic exchange of the pair (mo_ea, mo_eb) and the
 * absorbed-power accumulation, each cell in the order of _step_typed!
 * (src/fdtd/Solver.jl:310-345).  The result is that of mp_patch_E_dispersive on x, y, z,
 * then mp_patch_E_mo_gyration, then mp_pabs_accumulate, bit for bit.  It requires every
 * list to be the film list: entry n of each ADE list and of the MO list is film cell n,
 * as FDTDState builds them from the material cells (Solver.jl:114-159); the caller
 * vouches for that (the lengths are checked).  The three ADE lists share their storage:
 * the compute type, or BF16 / F16 with scale (FP32 compute), decoded and encoded as
 * mp_patch_E_dispersive does; mo may be NULL (no gyration).  mo_ea / mo_eb are two of g's E volumes and
 * mo_iea / mo_ieb their inv_eps volumes.  Replaces the launch sequence of
 * _ka_patch_E_dispersive! x3, _ka_patch_E_mo_gyration! and _ka_pabs_accumulate!
 * (Kernels.jl:595-646) when the lists coincide. */
int mp_film_epilogue(const mp_yee *g, mp_ade *ax, mp_ade *ay, mp_ade *az, mp_mo *mo, void *mo_ea,
                     void *mo_eb, const void *mo_iea, const void *mo_ieb, mp_film *f,
                     int32_t has_ade_work, double pabs_factor, double Q_voigt_TM, double Q_voigt_RE,
                     double dt);

/* 4TM + two-sublattice LLB, `subcycles` substeps of dt/subcycles, FP64 arithmetic.
 * has_pabs = 1 consumes window_sum/dt and resets it; 0 is the cool-down form
 * (relax_step!).  dt, T_min, inv_dT are rounded to the film CT first, as the launcher
 * does.  p holds FerrimagnetParameters{CT} widened to FP64.
 * Replaces _ka_multiphysics_step! (Kernels.jl:648-666; kernel :362-412) with
 * update_4tm (src/physics/Thermal.jl:24-40) and llb_step (Magnetization.jl:105-263). */
int mp_multiphysics_step(mp_film *f, const mp_ferri_params *p, const mp_lut *lut, double dt,
                         int32_t subcycles, int32_t brillouin_iters, int32_t has_pabs);

/* Per-step DFT of the tangential fields on up to MP_MAX_PLANES planes: the nb phase
 * factors (cos wt*dt, -sin wt*dt) are formed once into phase (device, 2*nb doubles),
 * then every plane cell accumulates 4 complex values per bin plus the signed flux
 * integral.  omega_bins is a device array of nb doubles.  Replaces
 * _ka_probe_dft_accumulate! (Kernels.jl:671-689; kernels :420-461). */
int mp_probe_dft_accumulate(const mp_yee *g, const mp_plane *planes, int32_t nplanes,
                            const double *omega_bins, double *phase, int32_t nb, double t,
                            double dt);

/* Mode-weighted plane sums of Ea and Eb on two planes of one axis, reduced on device:
 * out[0..3] = sum(Ea_pre*w), sum(Eb_pre*w), sum(Ea_post*w), sum(Eb_post*w) with
 * w = mode_w*area (n1*n2 doubles each, device).  out is a device array of 4 doubles.
 * Replaces _ka_probe_trace_plane! + the host sum (Kernels.jl:691-699, :467-478;
 * ProbeReadout.jl:286-298). */
int mp_probe_trace_plane(const mp_yee *g, int32_t axis, int64_t i_pre, int64_t i_post,
                         const double *mode_w, const double *area, double *out);

/* Plane-integrated Poynting flux sum((Ea*Hb - Eb*Ha)*area) on one plane, written to
 * out[slot] (device).  Replaces the host loop of record!(::FluxMonitor)
 * (src/sim/Monitors.jl:263-341) with an on-device reduction. */
int mp_flux_plane(const mp_yee *g, int32_t axis, int64_t index, const double *area,
                  double *out, int64_t slot);

/* The device reductions of one probe shot over the pre and post DFT planes (buffers as
 * mp_probe_dft_accumulate fills them for planes of n1 x n2 cells and nb bins; only the
 * buffer pointers of pre and post are read), written to out (device, 4*nb + 10 doubles):
 *   out[2q], out[2q+1]              Sx_pre[q]  = sum((Ea*conj(Hb) - Eb*conj(Ha)) * dA)
 *   out[2nb+2q], out[2nb+2q+1]      Sx_post[q]
 *   out[4nb], out[4nb+1]            net_pre = sum(energy_pre * dA), net_post
 *   out[4nb+2 .. 4nb+9]             sum(dft[., center_bin] * mode_w * dA) for post Ea,
 *                                   post Eb, pre Ea, pre Eb (re, im each; the caller
 *                                   divides by the mode norm)
 * with dA = area and mode_w device arrays of n1*n2 doubles.  Fixed partition and
 * reduction order: the result is the same on every call.  Replaces _plane_poynting,
 * _plane_net_flux and _jones_overlap on host downloads in probe_shot
 * (src/sim/ProbeReadout.jl:322-351, :362-420). */
int mp_probe_shot_reduce(const mp_plane *pre, const mp_plane *post, int64_t n1, int64_t n2,
                         int32_t nb, int32_t center_bin, const double *mode_w, const double *area,
                         double *out);

/* mp_probe_shot_reduce without any allocation: the same kernels, partition and output,
 * with the partial sums in a fixed device array that the library holds (one per GPU,
 * 320 KiB, kept until the process ends) instead of a buffer it allocates and regrows.
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  nb must not exceed
 * MP_SHOT_NOALLOC_MAX_BINS; larger nb is rejected with MP_ERR_ARG.  Calls on one GPU
 * share the array: calls that overlap on one GPU (several streams or host threads) race. */
#define MP_SHOT_NOALLOC_MAX_BINS 63
int mp_probe_shot_reduce_noalloc(const mp_plane *pre, const mp_plane *post, int64_t n1, int64_t n2,
                                 int32_t nb, int32_t center_bin, const double *mode_w,
                                 const double *area, double *out);

/* Weighted plane sum of one field component (0..5 = Ex, Ey, Ez, Hx, Hy, Hz) on the plane
 * `index` normal to `axis`: out[0] = sum(F * weight) over the plane cells in (n1, n2)
 * order, weight = 1 when null (device, n1*n2 doubles).  Replaces the host loops of
 * _mode_overlap (src/sim/Monitors.jl:350-368) and _mean_plane_components
 * (Monitors.jl:461-473) behind Polarimetry and mode-resolved Transmission/Reflection. */
int mp_plane_overlap(const mp_yee *g, int32_t component, int32_t axis, int64_t index,
                     const double *weight, double *out);

/* nsteps of the fixed step sequence, all launches on the library stream, no host sync:
 * H -> source(t) -> E -> ADE x,y,z -> MO -> n,t += 1,dt -> pabs -> 4TM+LLB every
 * multiphysics_every steps over the window since the last consume -> DFT record.
 * Replaces step!/_step_typed! + run! (src/fdtd/Solver.jl:277-383, 557-563) and the
 * per-step record! of ProbeReadout (src/sim/ProbeReadout.jl:303-307). */
int mp_run_steps(mp_state *s, int64_t nsteps);

/* Value of a pulse at time t (FP64), as the step driver evaluates it. */
double mp_pulse_value(const mp_pulse *p, double t);

/* ---- surrogate of the film integrator (no counterpart in the reference) ----------- */

/* Re-pack the FP32 parameters into the BF16 MFMA fragments the kernels read
 * (forward and transposed weights).  Call after every parameter change. */
int mp_mlp_pack(mp_mlp *m);

/* Inference on a batch: X raw features (B rows of MP_MLP_IN FP32), Y raw predicted
 * rates (B rows of MP_MLP_OUT FP32). */
int mp_mlp_forward(const mp_mlp *m, const float *X, float *Y, int64_t B);

/* Forward + backward of the mean-squared error over the normalised targets:
 * loss = mean_{b,o} (y_b,o - (T_b,o - y_mean_o) * y_inv_std_o)^2.  grad (device,
 * MP_MLP_NPARAM FP32) receives dloss/dparams; *loss (device double, may be NULL) the
 * loss.  Deterministic (fixed partition, fixed reduction order). */
int mp_mlp_backward(const mp_mlp *m, const float *X, const float *T, int64_t B, float *grad,
                    double *loss);

/* One Adam step (decoupled weight decay wd) on the MP_MLP_NPARAM parameters; step >= 1
 * is the bias-correction count.  Re-pack afterwards. */
int mp_mlp_adam(float *params, const float *grad, float *m1, float *m2, int64_t step, float lr,
                float beta1, float beta2, float eps, float wd);

/* Training rows from a film state, which is not modified: for every stride-th cell
 * (at most max_rows), the features above and the rates of the plain integrator over dt
 * (subcycles substeps, FP64, as mp_multiphysics_step would compute them; pabs =
 * window_sum/dt when has_pabs).  X: rows x MP_MLP_IN, T: rows x MP_MLP_OUT (device). */
int mp_surrogate_dataset(const mp_film *f, const mp_ferri_params *p, const mp_lut *lut,
                         double dt, int32_t subcycles, int32_t brillouin_iters, int32_t has_pabs,
                         int64_t stride, float *X, float *T, int64_t max_rows);

/* Film update by the surrogate, in place of mp_multiphysics_step (same arguments).  Per
 * cell: the network's rates give the end state over dt.  The step is checked against one
 * plain substep of h = dt/subcycles taken at its midpoint (temperatures and m_x halfway
 * along the predicted step, m_y and m_z at their start values): with r_mid that
 * substep's rate, the change must equal dt r_mid to within tol (absolute in m, relative
 * in T), the midpoint rule; m_y and m_z advance by dt r_mid.  A cell that fails falls back
 * to the plain integrator from its start state, so its result is the plain one.  Cost:
 * the network and 1 substep per accepted cell, subcycles more per fallback.  tol < 0
 * forces the plain integrator everywhere.  *fallbacks (device int64, may be NULL) is
 * incremented by the number of fallbacks. */
int mp_surrogate_step(mp_film *f, const mp_ferri_params *p, const mp_lut *lut,
                      const mp_mlp *m, double dt, int32_t subcycles, int32_t brillouin_iters,
                      int32_t has_pabs, double tol, int64_t *fallbacks);

/* Film reductions on device (src/sim/Monitors.jl:487-578 FilmAverage/SwitchedFraction,
 * src/physics/Magnetization.jl:280-283 norms): out[0..9] = means of m_TM_xyz, m_RE_xyz,
 * Te, Tl, then the fraction of cells with m_TM_x < 0, then max Te.  out: 10 device
 * doubles. */
int mp_film_reduce(const mp_film *f, double *out);

/* The sums behind one FilmAverage / SwitchedFraction record (src/sim/Monitors.jl:
 * 518-570, 487-497; reduce_mean / reduce_norm_mean / reduce_count_negative,
 * src/core/Backend.jl:49-57), over the f->n film cells in FP64: out[0..5] = sum of
 * m_TM_x, m_TM_y, m_TM_z, m_RE_x, m_RE_y, m_RE_z; out[6], out[7] = sum of |m_TM|, |m_RE|;
 * out[8..11] = sum of Te, Tl, Ts_TM, Ts_RE; out[12] = count of m_TM_x < 0.  The caller
 * divides by n.  out: 13 device doubles (zeros when n == 0).  Fixed reduction order. */
int mp_film_average(const mp_film *f, double *out);

/* Final-state switching diagnostics (switching_metrics, src/sim/Monitors.jl:682-727):
 * a cell is switched when m_TM_x < 0 and m_RE_x > 0; the core is U_abs >=
 * core_fraction * max(U_abs).  out (13 device doubles): [0] switched count, [1] count
 * m_TM_x < 0, [2] count m_RE_x > 0, [3], [4] sums of m_TM_x, m_RE_x over switched cells,
 * [5], [6] the same over unswitched cells, [7] max U_abs, [8] its first 0-based index,
 * [9] core count, [10] switched core count, [11], [12] sums of m_TM_x, m_RE_x over the
 * core.  The caller forms the fractions and means.  n must be positive. */
int mp_switching_metrics(const mp_film *f, double core_fraction, double *out);

/* Non-finite scan of a device array (NaNGuard, Monitors.jl:580-606): *flag (device
 * int32) is set to 1 if any element is NaN or Inf; the caller clears it.  Replaces
 * any_nonfinite (src/core/Backend.jl:59). */
int mp_any_nonfinite(const void *x, int32_t dtype, int64_t n, int32_t *flag);

/* One FP64 reduction of a device array of n elements (dtype MP_F32 or MP_F64) into
 * out[0] (device): MP_RED_SUM sum(x); MP_RED_COUNT_NEGATIVE the count of x < 0;
 * MP_RED_NORM_SUM sum(sqrt(x^2 + y^2 + z^2)) over three arrays.  0 when n == 0.  Fixed
 * partition and reduction order.  Replaces reduce_sum, reduce_mean (the caller divides
 * by n), reduce_count_negative and reduce_norm_mean (src/core/Backend.jl:49-57), which
 * the reference evaluates with GPUArrays mapreduce on device arrays. */
enum mp_reduce_op { MP_RED_SUM = 0, MP_RED_COUNT_NEGATIVE = 1, MP_RED_NORM_SUM = 2 };
int mp_reduce(const void *x, const void *y, const void *z, int32_t dtype, int64_t n, int32_t op,
              double *out);

/* out[b] = sum of x[b*block .. (b+1)*block - 1] for b < nblocks (FP64, device arrays):
 * sum(work_trace; dims=(1, 2)) of the probe trace buffer (src/sim/ProbeReadout.jl:291). */
int mp_sum_blocks(const double *x, int64_t block, int64_t nblocks, double *out);

/* plane_to_host's device side (src/core/Backend.jl:66-83): the plane `index` (0-based)
 * normal to `axis` of a (nx, ny, nz) field copied to out (device) in the order of the
 * reference's views: x -> field[i,:,:] as (ny, nz), y -> field[:,i,:] as (nx, nz),
 * z -> field[:,:,i] as (nx, ny), first index fastest. */
int mp_plane_gather(const void *field, int32_t dtype, int64_t nx, int64_t ny, int64_t nz,
                    int32_t axis, int64_t index, void *out);
/* |field| of one plane as float (device), in the plane order of mp_plane_gather: the movie
 * frame staging of the HDF5 frame writers, `dev .= abs.(view(field, ...))` followed by the
 * Float32 conversion (_stage_abs_frame!, ext/MagnetoPhotonicHDF5Ext.jl:619-624, in
 * record_production_frame! :626-636 and record_probe_frame! :685-697). */
int mp_plane_abs(const void *field, int32_t dtype, int64_t nx, int64_t ny, int64_t nz,
                 int32_t axis, int64_t index, float *out);

/* probe_trace_plane_kernel! (src/fdtd/Kernels.jl:467-478): work (ny, nz, 4), device FP64,
 * = Ey[i_pre]*w, Ez[i_pre]*w, Ey[i_post]*w, Ez[i_post]*w with w = mode_w*area (ny*nz
 * doubles each) on the two x-normal planes.  Replaces _ka_probe_trace_plane!
 * (Kernels.jl:691-699); mp_probe_trace_plane is the same with the plane sums fused. */
int mp_probe_trace_samples(const mp_yee *g, int64_t i_pre, int64_t i_post, const double *mode_w,
                           const double *area, double *work);

/* One probe movie frame, staged on the device (_record_probe_frame!,
 * src/sim/ProbeReadout.jl:193-215, which downloads the six planes with plane_to_host and
 * takes Float32.(abs.(...)) on the host): |Ey| and |Ez| as float on the z-plane kc (0-based)
 * into xy_* (nx*ny, x fastest, the reference's (Nx, Ny) slice) and on the x-planes i_pre,
 * i_post into yz_* (ny*nz, y fastest, (Ny, Nz)).  The outputs are device arrays, typically
 * frame f of a frame-last (d1, d2, frames) cube, the layout _stack_frames builds
 * (ProbeReadout.jl:172-181). */
int mp_probe_frames(const mp_yee *g, int64_t kc, int64_t i_pre, int64_t i_post, float *xy_ey, float *xy_ez,
                    float *yz_ey_pre, float *yz_ez_pre, float *yz_ey_post, float *yz_ez_post);

#ifdef __cplusplus
}
#endif
#endif /* MPHIP_H */
