"""Phase drivers around the step: pump with monitors, relax (cool-down) and the
equilibration dry-run.

Mirrors run_phase!(::Pump) and run_phase!(::Relax) (src/sim/Experiment.jl:294-345,
394-417), _phase_steps (:84-90) and _record_phase_monitors! (:242-252), and the
equilibration dry-run of examples/replicate_production.jl:428-440.  The steps run through
the C step driver (mp_run_steps) between the steps where a monitor is due; monitors keep
their records on the device.
"""
from __future__ import annotations

import math
from typing import Iterable, Optional

EQUIL_STEPS = 10000      # dry-run length of the production driver
EQUIL_DT = 0.5e-15       # and its relax step


def phase_steps(steps: Optional[int], until: Optional[float], dt: float, default_steps: int = 0) -> int:
    """_phase_steps (Experiment.jl:84-90): explicit steps, else ceil(until / dt)."""
    if steps is not None:
        return max(0, int(steps))
    if until is not None:
        return max(0, math.ceil(float(until) / float(dt)))
    return max(0, int(default_steps))


def record_monitors(monitors: Iterable, state) -> None:
    """_record_phase_monitors! (Experiment.jl:242-252): every due monitor records."""
    for m in monitors:
        if m.due(state.n):
            m.record(state)


def _next_due(monitors, n0: int, end: int) -> int:
    n = n0 + 1
    while n < end and not any(m.due(n) for m in monitors):
        n += 1
    return n


def run_pump(state, steps: Optional[int] = None, until: Optional[float] = None, monitors: Iterable = ()) -> dict:
    """run_phase!(::Pump) (Experiment.jl:294-345): EM + coupled steps, monitors after the
    steps at which they are due.  The steps between due steps are one driver call."""
    monitors = list(monitors)
    nsteps = phase_steps(steps, until, state.dt)
    n0, t0 = state.n, state.t
    end = n0 + nsteps
    while state.n < end:
        n = _next_due(monitors, state.n, end)
        state.run(n - state.n)
        record_monitors(monitors, state)
    return dict(phase="pump", steps=nsteps, n_start=n0, n_end=state.n, t_start=t0, t_end=state.t)


def run_relax(state, steps: Optional[int] = None, until: Optional[float] = None, dt: Optional[float] = None,
              subcycles: Optional[int] = None, monitors: Iterable = ()) -> dict:
    """run_phase!(::Relax) (Experiment.jl:394-417): 4TM + LLB with no heating at the relax
    step dt (default: the EM dt); n and t advance by one relax step each, so monitors see
    t = t_pump + n * dt_relax; the multiphysics window is closed at the end."""
    monitors = list(monitors)
    dt_relax = state.dt if dt is None else float(dt)
    nsteps = phase_steps(steps, until, dt_relax)
    n0, t0 = state.n, state.t
    for _ in range(nsteps):
        state.relax_step(dt_relax, subcycles=subcycles)
        state.n = state.n + 1
        state.t = state.t + dt_relax
        record_monitors(monitors, state)
    state.last_mp_n = state.n
    return dict(phase="relax", steps=nsteps, n_start=n0, n_end=state.n, t_start=t0, t_end=state.t, dt=dt_relax)


def equilibrate(state, steps: int = EQUIL_STEPS, dt: float = EQUIL_DT, subcycles: int = 1) -> None:
    """The equilibration dry-run before the pump (replicate_production.jl:433-439):
    relax steps that advance neither n nor t."""
    for _ in range(int(steps)):
        state.relax_step(dt, subcycles=subcycles)
