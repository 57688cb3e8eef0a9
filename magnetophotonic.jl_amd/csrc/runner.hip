// Host-side step driver: the fixed per-step launch sequence of an FDTDState, issued
// from C so a time loop costs no interpreter round trips.
//
// Reference: step!/_step_typed! (src/fdtd/Solver.jl:277-369), the multiphysics window
// _consume_multiphysics_window! (:373-383), run! (:557-563), the pulse values of
// src/fdtd/Source.jl:1-33 and the per-step device record! of ProbeReadout
// (src/sim/ProbeReadout.jl:274-307).
#include <cmath>
#include <cstdlib>

#include "common.h"
#include "monitors.h"
#include "sources.h"
#include "yee.h"

using namespace mp;

MP_API double mp_pulse_value(const mp_pulse *p, double t) {
    switch (p->kind) {
    case MP_PULSE_GAUSSIAN: {
        const double u = (t - p->t0) / p->tau;
        return p->amplitude * std::exp(-(u * u)) * std::sin(p->omega * (t - p->t0) + p->phase);
    }
    case MP_PULSE_CARRIER_GAUSSIAN: {
        const double u = (t - p->t0) / p->tau;
        return p->amplitude * std::sin(p->omega * t) * std::exp(-0.5 * (u * u));
    }
    case MP_PULSE_CONTINUOUS:
        return p->amplitude * std::sin(p->omega * t + p->phase);
    default:
        return 0.0;
    }
}

// The source of the step, reading the field from `in` and writing it to `out` (the same
// volumes except in a ping-pong step).
static int inject(mp_state *s, const mp_yee *in, const mp_yee *out) {
    const mp_source &src = s->source;
    if (src.kind == MP_SRC_NONE) return MP_OK;
    const double v = mp_pulse_value(&src.pulse, s->t);
    switch (src.kind) {
    case MP_SRC_SOFT: return inject_soft_pp(in, out, src.component, src.i, src.j, src.k, v);
    case MP_SRC_PLANE: return inject_plane_pp(in, out, src.component, src.axis, src.i0, src.w0, src.i1, src.w1, v);
    case MP_SRC_MODE_X: return mp_inject_mode_x_3d(in, src.component, src.i, src.profile, v, src.hcorr);
    default: return fail_arg("mp_run_steps: unknown source kind %d", src.kind);
    }
}

// Whether the steps may run as ping-pong pairs: a second buffer set, the fused Yee form for
// this grid, and a source that writes only E and D (an H source, or the mode source's H
// correction, lands between the H and E updates the fused launch joins).
static bool pingpong_ok(const mp_state *s) {
    for (int c = 0; c < 6; ++c)
        if (!s->yee_alt[c]) return false;
    const mp_source &src = s->source;
    if (src.kind == MP_SRC_MODE_X) return false;
    if ((src.kind == MP_SRC_SOFT || src.kind == MP_SRC_PLANE) && (src.component < MP_EX || src.component > MP_EZ))
        return false;
    return yee_fused_ok(s->yee, s->cpml);
}

// E volume `p` of the caller's yee as the same component of `y`.
static void *same_component(const mp_yee *caller, const mp_yee *y, const void *p) {
    if (p == caller->ex) return y->ex;
    if (p == caller->ey) return y->ey;
    if (p == caller->ez) return y->ez;
    return const_cast<void *>(p);
}

// Side stream for the film integrator.  The 4TM + LLB update of step n is FP64-compute
// bound while the next step's Yee updates are HBM bound; its only consumers are that
// step's MO gyration (reads m) and absorbed-power accumulation (writes the window it
// reads), so it runs on a second stream and the main stream waits on its event just
// before those two kernels.  Both streams keep the default priority.
struct Overlap {
    hipStream_t side = nullptr;
    hipEvent_t main_done = nullptr, film_done = nullptr;
    bool pending = false;     // film_done recorded and not yet waited on
};
static Overlap g_ov;

static int overlap_init() {
    if (g_ov.side) return MP_OK;
    MP_HIP(hipStreamCreateWithFlags(&g_ov.side, hipStreamNonBlocking));
    MP_HIP(hipEventCreateWithFlags(&g_ov.main_done, hipEventDisableTiming));
    MP_HIP(hipEventCreateWithFlags(&g_ov.film_done, hipEventDisableTiming));
    return MP_OK;
}

static int overlap_join() {
    if (!g_ov.pending) return MP_OK;
    MP_HIP(hipStreamWaitEvent(stream(), g_ov.film_done, 0));
    g_ov.pending = false;
    return MP_OK;
}

static int film_update(mp_state *s, double dt_mp) {
    if (s->surrogate)
        return mp_surrogate_step(s->film, s->params, s->lut, s->surrogate, dt_mp, s->subcycles, s->brillouin_iters,
                                 1, s->surrogate_tol, s->surrogate_fallbacks);
    return mp_multiphysics_step(s->film, s->params, s->lut, dt_mp, s->subcycles, s->brillouin_iters, 1);
}

// About 10 us on one wave of the main stream (3 x s_sleep 127, 64 clocks each at ~2.4 GHz).
// Issued right after the film update: the film update's workgroups then reach the CUs
// before those of the next launches on the main stream (the DFT snapshot, the H update).
// Without a DFT the 512 x 512 x 256 step measured 1.74 ms with both dispatched together,
// 1.51 ms with this head start.  With the DFT the snapshot alone gave the head start in
// some builds (1.487 ms) and not in others that differed only in the monitor code object's
// layout (1.52-1.56 ms); with this launch every such build measured 1.489 ms.  Shorter
// sleeps (one s_sleep, or an empty launch) measured 1.60 ms, longer ones (4-8) 1-15 us more
// per step than two or three.
__global__ void head_start_kernel() {
    for (int r = 0; r < 3; ++r) __builtin_amdgcn_s_sleep(127);
}

// The film update on the side stream, after everything issued so far on the main stream.
static int launch_film_side(mp_state *s, double dt_mp) {
    if (int rc = overlap_init()) return rc;
    MP_HIP(hipEventRecord(g_ov.main_done, stream()));
    MP_HIP(hipStreamWaitEvent(g_ov.side, g_ov.main_done, 0));
    hipStream_t prev = swap_stream(g_ov.side);
    const int rc = film_update(s, dt_mp);
    swap_stream(prev);
    if (rc) return rc;
    MP_HIP(hipEventRecord(g_ov.film_done, g_ov.side));
    g_ov.pending = true;
    head_start_kernel<<<1, 64, 0, stream()>>>();
    MP_LAUNCH_CHECK("head_start_kernel");
    return MP_OK;
}

static int film_list_work(mp_state *s, const mp_yee *y);
static int film_update_due(mp_state *s);

// One step from the fields in `in` to the fields in `out`: in place (in == out, the
// mp_update_H_3d / inject / mp_update_E_3d sequence) or ping-pong (the H shell, the
// source, the fused interior H + E, the E shell).  Everything after the Yee updates reads
// and writes `out`.
static int one_step(mp_state *s, const mp_yee *in, const mp_yee *out) {
    int rc;
    if (in == out) {
        yee_shell_split_x(g_ov.pending);   // the film update of the previous step runs beside
        rc = mp_update_H_3d(in, s->cpml, s->dt, s->mu0);
        yee_shell_split_x(false);
        if (rc) return rc;
        if ((rc = inject(s, in, in))) return rc;
        if ((rc = mp_update_E_3d(in, s->cpml, s->dt))) return rc;
    } else {
        // E on the planes no update writes carries over first: the source and the film list
        // work may have changed it there in the set this step reads
        if ((rc = yee_sync_boundary(in, out, true))) return rc;
        if ((rc = yee_fused_H_shell(in, out, s->cpml, s->dt, s->mu0))) return rc;
        if ((rc = inject(s, in, out))) return rc;
        if ((rc = yee_fused_interior(in, out, s->cpml, s->dt, s->mu0))) return rc;
        if ((rc = yee_fused_E_shell(out, s->cpml, s->dt))) return rc;
    }
    const mp_yee *y = out;
    // a full DFT batch (its last step recorded at the end of the previous step) is added to
    // the accumulators here, after the previous film update has finished: alone it takes
    // about 220 us, beside that update more than twice as long
    if (s->record_dft && s->nplanes > 0 && dft_batch_full()) {
        if ((rc = overlap_join())) return rc;
        if ((rc = dft_flush(y, s->planes, s->nplanes, s->nb, s->dt))) return rc;
    }
    if (s->film_lists_coincide && s->ade[0] && s->film) {
        // the film list work in one pass; it contains the gyration, which reads the film
        // update of the previous step
        if ((rc = overlap_join())) return rc;
        if ((rc = mp_film_epilogue(y, s->ade[0], s->ade[1], s->ade[2], s->mo, same_component(s->yee, y, s->mo_ea),
                                   same_component(s->yee, y, s->mo_eb), s->mo_iea, s->mo_ieb, s->film,
                                   s->has_ade_work, s->pabs_factor, s->Q_TM, s->Q_RE, s->dt)))
            return rc;
        s->n += 1;
        s->t += s->dt;
        if ((rc = film_update_due(s))) return rc;
    } else {
        if ((rc = film_list_work(s, y))) return rc;
    }
    if (s->record_dft && s->nplanes > 0) {
        if ((rc = dft_record_batched(y, s->planes, s->nplanes, s->omega_bins, s->phase, s->nb, s->t, s->dt)))
            return rc;
    }
    return MP_OK;
}

// The list kernels after the E update, one launch each (the lists need not coincide).
static int film_list_work(mp_state *s, const mp_yee *y) {
    int rc;
    if (s->ade[0]) {
        void *E[3] = {y->ex, y->ey, y->ez};
        for (int c = 0; c < 3; ++c)
            if ((rc = mp_patch_E_dispersive(E[c], s->ade[c], y->field_dtype, s->dt))) return rc;
    }
    if ((rc = overlap_join())) return rc;   // the previous step's film update is complete
    if (s->mo && s->film) {
        if ((rc = mp_patch_E_mo_gyration(same_component(s->yee, y, s->mo_ea), same_component(s->yee, y, s->mo_eb),
                                         s->mo, s->mo_iea, s->mo_ieb, y->field_dtype, s->film->m_TM_x,
                                         s->film->m_RE_x, s->Q_TM, s->Q_RE, s->dt)))
            return rc;
    }
    s->n += 1;
    s->t += s->dt;
    if (s->film) {
        if ((rc = mp_pabs_accumulate(s->film, y, s->ade[0], s->ade[1], s->ade[2], s->has_ade_work,
                                     s->pabs_factor, s->dt)))
            return rc;
        if ((rc = film_update_due(s))) return rc;
    }
    return MP_OK;
}

// The 4TM + LLB update when the multiphysics window closes at step n (Solver.jl:356-367).
// Launched at once: on the side stream it runs beside step n+1's H update (launching it
// after that H update instead, beside the E update, measured 2.28 against 2.19 ms per
// 512x512x256 step: the E update suffers more from the company).
static int film_update_due(mp_state *s) {
    int rc;
    if (s->film) {
        if (s->multiphysics_enabled && s->multiphysics_every > 0 && s->n % s->multiphysics_every == 0) {
            const int64_t window = s->n - s->last_mp_n;
            if (window > 0) {
                const double dt_mp = s->dt * (double)window;
                if (s->film_serial) {
                    if ((rc = film_update(s, dt_mp))) return rc;
                } else if ((rc = launch_film_side(s, dt_mp))) {
                    return rc;
                }
                s->last_mp_n = s->n;
            }
        }
    }
    return MP_OK;
}

MP_API int mp_run_steps(mp_state *s, int64_t nsteps) {
    if (!s || !s->yee || !s->cpml) return fail_arg("mp_run_steps: null state");
    if (s->film && (!s->params || !s->lut)) return fail_arg("mp_run_steps: film state needs params and LUT");
    if (s->ade[0] && (!s->ade[1] || !s->ade[2])) return fail_arg("mp_run_steps: ADE needs all three lists");
    // pairs of ping-pong steps (the yee's volumes -> the second set -> the yee's volumes),
    // then any last step in place; the second set first takes the cells no update writes
    const bool pp = nsteps >= 2 && pingpong_ok(s);
    mp_yee alt = *s->yee;
    int rc = MP_OK;
    int64_t k = 0;
    if (pp) {
        alt.ex = s->yee_alt[0]; alt.ey = s->yee_alt[1]; alt.ez = s->yee_alt[2];
        alt.hx = s->yee_alt[3]; alt.hy = s->yee_alt[4]; alt.hz = s->yee_alt[5];
        rc = yee_sync_boundary(s->yee, &alt, false);
        for (; !rc && k + 2 <= nsteps; k += 2)
            if (!(rc = one_step(s, s->yee, &alt))) rc = one_step(s, &alt, s->yee);
    }
    for (; !rc && k < nsteps; ++k) rc = one_step(s, s->yee, s->yee);
    if (rc) {
        if (s->record_dft && s->nplanes > 0) dft_flush(s->yee, s->planes, s->nplanes, s->nb, s->dt);
        overlap_join();
        return rc;
    }
    // DFT steps still held in the batch reach the accumulators before the call returns
    if (s->record_dft && s->nplanes > 0) {
        if ((rc = dft_flush(s->yee, s->planes, s->nplanes, s->nb, s->dt))) {
            overlap_join();
            return rc;
        }
    }
    // everything issued here is ordered before later work on the caller's stream
    return overlap_join();
}
