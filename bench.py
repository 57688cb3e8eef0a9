"""Benchmark of the coupled 3-D time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload coupled_pump|em_stencil]
                    [--form plain|mixed|surrogate|surrogate+mixed] [--storage bf16|f16]
                    [--no-accuracy] [--no-cpu]

A step is one pass of the full coupled sequence (Yee H, pump source, Yee E, ADE x/y/z,
gyrotropic MO, absorbed power, 4TM + LLB with 4 substeps, DFT on 6 planes x 4 bins)
over the 512 x 512 x 256 synthetic pump workload, inputs resident in HBM.  Prints one
JSON line: Mcell-updates/s = Nx*Ny*Nz*K / wall, the roofline of the dominant kernel
(algorithmic bytes per launch / mean launch time from HIP events recorded around its
launches in every 4th step of the timed region; the per-kernel breakdown comes from further steps with
every launch bracketed, then from as many steps with the film update serial, each kernel
alone), the FP64 roofline of the film update, and the CPU oracle timed on the host cores
(FP32 on every core, FP32 on one thread, FP64 on every core).  Counter figures (measured
HBM traffic, FP64 operations) are printed only while the kernel sources have the digest
they were measured on, else null.

--form mixed stores CPML psi and the ADE state in BF16 (--storage f16: binary16 with
scales).  --form surrogate replaces the 4TM + LLB substeps with the learned film update
(checked against plain substeps, plain fallback); its network is first trained on the
GPU, outside the timed region, on rows from a 96 x 96 x 256 slice of the same workload.
--form surrogate+mixed is both.

The default plain run also prints `form_comparison`, after the timed region: the
surrogate + BF16 and surrogate + binary16 forms timed on the same step, and the final
film-mean |dm| of each form on the 128 x 128 x 64 accuracy case against the same path in
FP64 on the GPU (--no-accuracy skips it).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "magnetophotonic.jl_amd"))

METRIC = "coupled Mcell-updates/s + % HBM roofline; surrogate vs plain at |Δm|≤1e-3"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64
KERNELS = ("yee_H", "yee_HE", "yee_E", "pingpong_sync", "inject_plane", "ade_patch", "mo_gyration", "pabs",
           "film_epilogue", "multiphysics", "surrogate", "surrogate_fallback", "dft_phase", "dft_accumulate",
           "dft_snapshot", "dft_batch")
DFT_BATCH = 32   # steps per batched DFT pass in the library
PROFILE_EVERY = 4   # timed steps per bracketed launch of the dominant kernel


FORMS = ("plain", "mixed", "surrogate", "surrogate+mixed")


def build_workload(name: str, form: str, dtype: str = "f32", storage: str = "bf16"):
    from mphip import configs
    if name == "coupled_pump":
        kw = {}
        if form in ("mixed", "surrogate+mixed"):
            kw = dict(psi_precision=storage, ade_state_precision=storage)
        spec, meta = configs.coupled_pump(nx=512, ny=512, nz=256, steps=20000, dtype=dtype, **kw)
    elif name == "em_stencil":
        spec, meta = configs.em_stencil(n=384, npml=16, steps=5000, dtype=dtype)
        meta["planes"], meta["omega_bins"] = [], []
    else:
        raise SystemExit(f"unknown workload {name}")
    return spec, meta


def train_pump_surrogate(steps: int, seed: int = 1):
    """Network for the pump-phase film update, trained on the GPU on rows from a
    96 x 96 x 256 slice of the 512 x 512 x 256 pump (same fluence, source and film)."""
    from mphip import _lib, configs
    from mphip import surrogate as S
    from mphip.state import FDTDState
    t0 = time.perf_counter()
    spec, _ = configs.coupled_pump(nx=96, ny=96, nz=256, steps=20000)
    st = FDTDState(spec)
    X, T = S.collect_trajectory(st, steps, 10)
    st.free()
    X, T = S.shuffled(X, T, seed)
    sur = S.Surrogate(norm=S.make_norm(X[:200000], T[:200000]), seed=seed)
    data = S.Dataset(X, T)
    sur.train(data, 10000, batch=min(65536, X.shape[0]), lr=3e-3, lr_end=1e-5)
    loss = sur.evaluate(data, rows=262144)
    data.free()
    _lib.call("synchronize")
    return sur, dict(train_rows=int(X.shape[0]), adam_steps=10000, train_loss=round(loss, 6),
                     train_s=round(time.perf_counter() - t0, 1))


def relax_accuracy():
    """Surrogate against plain on the relax phase of the 128 x 128 x 64 accuracy case
    (held-out 60 mJ/cm^2; network trained on the GPU on rows from 35, 50, 70, 85 mJ/cm^2):
    film-mean |dm| against the same path run in FP64 on the GPU, relax wall time and
    fallback fraction.  Not part of the timed region."""
    from mphip import surrogate as S
    t0 = time.perf_counter()
    X, T = S.relax_rows((35.0, 50.0, 70.0, 85.0))
    rsur = S.train_relax_surrogate(X, T)
    ref = S.pump_relax_film("f64")
    res = S.relax_comparison(rsur, ref, tols=(1e-3, 1e-4), plain_subcycles=(10, 2, 3))
    rsur.free()
    res["rows"] = int(X.shape[0])
    res["reference"] = "same path in FP64 on the GPU"
    res["study_s"] = round(time.perf_counter() - t0, 1)
    return res


def form_comparison(plain_ms: float, steps: int, warmup: int, tol: float, kinds=("bf16", "f16")):
    """The surrogate + reduced-storage forms against plain FP32: ms per step of the
    512 x 512 x 256 pump step in each form (same DFT planes; the pump network trained on the
    GPU first, outside the timed steps), and the final film-mean |dm| of the 128 x 128 x 64
    accuracy case against the same path in FP64 on the GPU for each form."""
    from mphip import _lib
    from mphip import surrogate as S
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    t0 = time.perf_counter()
    res = {"plain": {"ms_per_step": round(plain_ms, 4)}}
    sur, info = train_pump_surrogate(warmup + steps)
    for kind in kinds:
        spec, meta = build_workload("coupled_pump", "surrogate+mixed", storage=kind)
        st = FDTDState(spec)
        st.attach_dft(DFTPlanes(st, meta["planes"], meta["omega_bins"]))
        st.attach_surrogate(sur, tol)
        st.run(warmup)
        st.synchronize()
        sur.reset_fallbacks()
        t1 = time.perf_counter()
        st.run(steps)
        st.synchronize()
        ms = (time.perf_counter() - t1) / steps * 1e3
        res[f"surrogate+{kind}"] = {"ms_per_step": round(ms, 4), "speedup_vs_plain": round(plain_ms / ms, 3),
                                    "fallback_frac": round(sur.fallback_count() / (spec.n_film * steps), 5),
                                    "film_list_work": "one fused pass" if st.fused_epilogue else "five launches"}
        st.free()
        _lib.call("synchronize")
    sur.free()
    res["pump_network"] = info
    res["steps"], res["warmup"], res["tol"] = steps, warmup, tol
    res["accuracy_128x128x64"] = S.accuracy_forms(kinds, tol)
    res["target"] = "surrogate+mixed >= 2x plain at film-mean |dm| <= 1e-3 against FP64"
    res["comparison_s"] = round(time.perf_counter() - t0, 1)
    return res


def _oracle_rate(driver, spec, max_steps: int, budget_s: float, warmup: int = 1):
    """Steps of the C oracle on `spec` after `warmup` steps: up to max_steps, stopping once
    budget_s has passed (at least one).  Returns (steps, seconds)."""
    st = driver.OracleState(spec)
    for _ in range(warmup):
        st.step()
    k, t0 = 0, time.perf_counter()
    while k < max_steps and (k == 0 or time.perf_counter() - t0 < budget_s):
        st.step()
        k += 1
    return k, time.perf_counter() - t0


def cpu_baseline(spec, meta, workload: str, steps: int, budget_s: float = 12.0):
    """The C oracle (a restatement of the reference CPU path) on the host cores: FP32 on
    every core (the headline figure), FP32 on one thread, FP64 on every core.  Each sample
    is bounded (steps or seconds, whichever ends first) after one warm-up step."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import driver
    from mphip import abi
    L = driver.lib(abi)
    L.orc_num_threads.restype = C.c_int
    L.orc_set_num_threads.restype = None
    L.orc_set_num_threads.argtypes = [C.c_int]
    nx, ny, nz = spec.shape
    threads = int(L.orc_num_threads())

    def entry(k, el, cores, what):
        return dict(value=round(nx * ny * nz * k / el / 1e6, 2), unit="Mcell-updates/s", cores=cores,
                    kind="port", sample=f"{k} steps of {meta['name']} after 1 warm-up step, {what}, "
                                        f"C oracle, {el:.1f} s")

    k, el = _oracle_rate(driver, spec, steps, 1e9)
    out = entry(k, el, threads, f"FP32, OpenMP on {threads} threads")
    L.orc_set_num_threads(1)
    try:
        k, el = _oracle_rate(driver, spec, steps, budget_s)
    finally:
        L.orc_set_num_threads(threads)
    out["one_thread"] = entry(k, el, 1, "FP32, one thread")
    spec64, _ = build_workload(workload, "plain", dtype="f64")
    k, el = _oracle_rate(driver, spec64, steps, budget_s)
    out["fp64"] = entry(k, el, threads, f"FP64, OpenMP on {threads} threads")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="coupled_pump")
    ap.add_argument("--form", default="plain", choices=FORMS)
    ap.add_argument("--storage", default="bf16", choices=("bf16", "f16"),
                    help="psi / ADE storage of the mixed forms")
    ap.add_argument("--tol", type=float, default=1e-4, help="surrogate check tolerance")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the form comparison (plain) / the relax accuracy study (surrogate)")
    ap.add_argument("--compare-steps", type=int, default=50, help="timed steps per form in the form comparison")
    ap.add_argument("--pingpong", action="store_true",
                    help="ping-pong steps: a second E/H buffer set and the fused interior H + E launch")
    ap.add_argument("--film-serial", action="store_true",
                    help="the film update on the launch stream after each step's other launches "
                         "(one stream; for counter runs, whose profiler serialises every launch)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB, the known-size "
                         "kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the coupled step would shard on z with per-step halo planes, not built)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    from mphip import _lib, roofline
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState

    if args.copy_probe_mib > 0:
        from mphip.state import DeviceBuffer
        nb = args.copy_probe_mib << 20
        a, b = DeviceBuffer.zeros(nb), DeviceBuffer.zeros(nb)
        _lib.call("memcpy_d2d", b.vp, a.vp, C.c_size_t(nb))
        _lib.call("synchronize")
        a.free()
        b.free()
    spec, meta = build_workload(args.workload, args.form, storage=args.storage)
    planes, bins = meta.get("planes", []), meta.get("omega_bins", [])
    st = FDTDState(spec, pingpong=args.pingpong)
    if planes:
        st.attach_dft(DFTPlanes(st, planes, bins))
    L = _lib.lib()
    sur_info = None
    surrogate_form = args.form.startswith("surrogate")
    if surrogate_form:
        sur, sur_info = train_pump_surrogate(args.warmup + args.steps)
        st.attach_surrogate(sur, args.tol)
    if args.film_serial:
        st.film_serial = True
    st.run(args.warmup)
    st.synchronize()
    if surrogate_form:
        sur.reset_fallbacks()

    nx, ny, nz = spec.shape
    ncell = nx * ny * nz
    eps_idx = st.yee.n_eps > 0
    kb = roofline.step_bytes(spec, planes, len(bins), eps_idx, DFT_BATCH, st.fused_epilogue, args.pingpong)
    if "multiphysics" in kb:
        kb["surrogate"] = kb["multiphysics"]
    # the dominant kernel is the largest HBM consumer of the step (algorithmic bytes);
    # its launch times do not depend on what the side stream runs beside it
    dom = max((k for k in kb if k in KERNELS), key=lambda k: kb[k])

    # timed region: K steps; HIP events on the launch stream bracket the dominant kernel's
    # launches only, and only in every PROFILE_EVERY-th step: each bracket is two event
    # records, which hold the launch stream ~12 us (the next launch waits for each record)
    L.mp_profile_select(dom.encode())
    every = PROFILE_EVERY if args.steps >= 2 * PROFILE_EVERY else 1
    L.mp_profile_sample(every)
    L.mp_profile_enable(1)
    st.synchronize()
    t0 = time.perf_counter()
    st.run(args.steps)
    st.synchronize()
    wall = time.perf_counter() - t0
    n, ms = C.c_int64(), C.c_double()
    L.mp_kernel_stats(dom.encode(), C.byref(n), C.byref(ms))
    L.mp_profile_enable(0)
    if not n.value:
        raise RuntimeError(f"bench.py: the dominant kernel {dom} did not run in the timed region")
    dom_launches, dom_avg_us = n.value, ms.value / n.value * 1e3
    fallbacks = sur.fallback_count() if sur_info is not None else 0

    # per-kernel breakdown: further steps after the timed region, every launch bracketed
    brk_steps = min(args.steps, 50)
    L.mp_profile_sample(1)
    L.mp_profile_select(None)
    L.mp_profile_enable(1)
    st.run(brk_steps)
    st.synchronize()
    stats = {}
    for k in KERNELS:
        n, ms = C.c_int64(), C.c_double()
        L.mp_kernel_stats(k.encode(), C.byref(n), C.byref(ms))
        if n.value:
            stats[k] = dict(launches=n.value, avg_us=round(ms.value / n.value * 1e3, 2))
    L.mp_profile_enable(0)
    # the same steps with the film update on the main stream after the step's launches:
    # every kernel's time alone, without the side stream's company
    serial = {}
    if spec.has_film and not args.film_serial:
        st.film_serial = True
        L.mp_profile_select(None)
        L.mp_profile_enable(1)
        st.run(brk_steps)
        st.synchronize()
        for k in KERNELS:
            n, ms = C.c_int64(), C.c_double()
            L.mp_kernel_stats(k.encode(), C.byref(n), C.byref(ms))
            if n.value:
                serial[k] = dict(launches=n.value, avg_us=round(ms.value / n.value * 1e3, 2))
        L.mp_profile_enable(0)
        st.film_serial = False
    mcell = ncell * args.steps / wall / 1e6
    # calls of the dominant kernel per step (one for a Yee update), from the sampled count
    calls_per_step = max(1, round(dom_launches * every / args.steps))
    per_launch = kb[dom] / calls_per_step
    achieved = per_launch / (dom_avg_us * 1e-6) / 1e9
    for k, v in stats.items():
        if k in kb:
            launches_per_step = v["launches"] / brk_steps
            v["alg_GBs"] = round(kb[k] / launches_per_step / (v["avg_us"] * 1e-6) / 1e9, 1)
    for k, v in serial.items():
        if k in kb:
            v["alg_GBs"] = round(kb[k] / (v["launches"] / brk_steps) / (v["avg_us"] * 1e-6) / 1e9, 1)
    step_gbs = kb["total"] / (wall / args.steps) / 1e9
    traffic, traffic_src = roofline.measured_traffic(meta["name"], args.form, dom, eps_idx)
    film_rf = None
    if "multiphysics" in stats and not surrogate_form:
        ops, ops_src = roofline.fp64_ops_per_cell_substep(meta["name"])
        flops = ops * spec.n_film * spec.subcycles if ops else None
        launch_us = {"beside the Yee updates": stats["multiphysics"]["avg_us"]}
        if "multiphysics" in serial:
            launch_us["alone"] = serial["multiphysics"]["avg_us"]
        film_rf = {"bound": "fp64", "kernel": "multiphysics", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "ops_per_cell_substep": ops, "flops_per_launch": flops, "ops_from": ops_src,
                   "avg_us": launch_us,
                   "achieved": {k: (round(flops / (u * 1e-6) / 1e12, 2) if flops else None)
                                for k, u in launch_us.items()},
                   "frac": {k: (round(flops / (u * 1e-6) / 1e12 / FP64_PEAK_TFLOPS, 4) if flops else None)
                            for k, u in launch_us.items()}}
    out = {
        "metric": METRIC,
        "value": round(mcell, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "f32", "mixed": f"f32 ({args.storage} psi/ADE storage)",
                  "surrogate": "f32 (film update: bf16 MFMA network, f64 check)",
                  "surrogate+mixed": f"f32 ({args.storage} psi/ADE storage; film update: bf16 MFMA network, "
                                     f"f64 check)"}[args.form],
        "data": "synthetic",
        "config": {"workload": meta["name"], "grid": list(spec.shape), "film_cells": spec.n_film,
                   "form": args.form, "dft_planes": len(planes), "dft_bins": len(bins),
                   "llb_substeps": spec.subcycles, "cpml": list(spec.cpml.widths),
                   "inv_eps": f"uint8 index, {st.yee.n_eps} triples" if eps_idx else "FP32 volumes",
                   "film_list_work": "one fused pass" if st.fused_epilogue else "five launches",
                   "yee_steps": "ping-pong pairs, fused interior H + E" if args.pingpong else "in place",
                   "film_update": "serial, on the launch stream" if args.film_serial else "side stream, beside the next step's Yee updates",
                   "dft_batch": DFT_BATCH},
        "roofline": {"bound": "hbm", "kernel": dom, "launches": dom_launches, "timed_every": every,
                     "avg_us": round(dom_avg_us, 2),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_from": traffic_src,
                     "alg_bytes_per_launch": int(per_launch)},
        "step_roofline": {"alg_bytes_per_step": kb["total"],
                          "alg_bytes_per_cell": round(kb["total"] / ncell, 2),
                          "achieved_GBs": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernels": stats,
        "kernels_from": f"{brk_steps} steps after the timed region, every launch bracketed by HIP events",
    }
    if serial:
        out["kernels_film_serial"] = serial
        out["kernels_film_serial_from"] = (f"{brk_steps} further steps with the film update on the main stream "
                                           f"after the step's launches (each kernel alone)")
    if film_rf is not None:
        out["film_roofline"] = film_rf
    if sur_info is not None:
        sur_info["tol"] = args.tol
        sur_info["fallback_frac"] = round(fallbacks / (spec.n_film * args.steps), 5)
        if not args.no_accuracy:
            sur_info["relax_accuracy"] = relax_accuracy()
        out["config"]["surrogate"] = sur_info
    if args.form == "plain" and args.workload == "coupled_pump" and not args.no_accuracy:
        st.free()
        out["form_comparison"] = form_comparison(wall / args.steps * 1e3, min(args.compare_steps, args.steps),
                                                 min(args.warmup, 10), args.tol)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, meta, args.workload, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
