"""meep step functions and stop conditions for Simulation.run (2D cells).

Follows meep_torch/mp/run_control.py: the calling convention of step
functions (:199-222), the combinators at_beginning / at_end / at_time /
at_every / after_time / before_time / after_sources / during_sources
(:225-322), the stop conditions stop_when_fields_decayed and
stop_after_walltime (:325-354), FluxData (:189-191) and the hook schedule of
its _run_until loop (plan_hooks / chunk_size / fire_per_step /
fire_scheduled, :457-519).

The schedule decides how many Yee steps the HIP engine may take in one call
before a hook or a stop condition has to look at the fields: hook-free
stretches run as one engine call, an active per-step hook or any stop
condition forces one step per call.
"""
import inspect
import math
import time
from collections import namedtuple

FluxData = namedtuple("FluxData", ["E", "H"])

_EPS_T = 1e-9   # time comparisons tolerate this much rounding of k * dt


def _n_args(func):
    """Positional arguments a step function takes (self excluded for bound
    methods and callable objects)."""
    if inspect.ismethod(func):
        return func.__code__.co_argcount - 1
    if inspect.isfunction(func):
        return func.__code__.co_argcount
    return func.__call__.__func__.__code__.co_argcount - 1


def eval_step_func(sim, func, todo):
    """f(sim) runs on todo == "step" only; f(sim, todo) always runs and
    decides for itself (meep's protocol)."""
    n = _n_args(func)
    if n == 1:
        if todo == "step":
            func(sim)
    elif n == 2:
        func(sim, todo)
    else:
        raise ValueError(f"step function {getattr(func, '__name__', func)!r} requires 1 or 2 arguments")


class _Hook:
    """Base of the combinators: each is itself an f(sim, todo) step function
    wrapping the step functions it was given."""

    def _fire(self, sim, todo):
        for g in self.fns:
            eval_step_func(sim, g, todo)


class at_beginning(_Hook):
    def __init__(self, *fns):
        self.fns = fns
        self._fired = False

    def __call__(self, sim, todo="step"):
        if self._fired:
            return
        self._fire(sim, todo)
        self._fired = True


class at_end(_Hook):
    def __init__(self, *fns):
        self.fns = fns

    def __call__(self, sim, todo="step"):
        if todo != "finish":
            return
        self._fire(sim, "step")
        self._fire(sim, "finish")


class at_time(_Hook):
    def __init__(self, t, *fns):
        self.t = float(t)
        self.fns = fns
        self._fired = False

    def __call__(self, sim, todo="step"):
        if todo != "finish" and sim.round_time() < self.t:
            return
        if todo == "finish" or not self._fired:
            self._fire(sim, todo)
        if todo == "step":
            self._fired = True


class at_every(_Hook):
    def __init__(self, dt, *fns):
        self.every = float(dt)
        self.fns = fns
        self._last = 0.0

    def __call__(self, sim, todo="step"):
        t = sim.round_time()
        if todo == "finish" or t >= self._last + self.every - 0.5 * sim.grid.dt:
            self._fire(sim, todo)
            self._last = t


class after_time(_Hook):
    def __init__(self, t, *fns):
        self.t = float(t)
        self.fns = fns

    def __call__(self, sim, todo="step"):
        if todo == "finish" or sim.round_time() >= self.t:
            self._fire(sim, todo)


class before_time(_Hook):
    def __init__(self, t, *fns):
        self.t = float(t)
        self.fns = fns

    def __call__(self, sim, todo="step"):
        if todo == "finish" or sim.round_time() < self.t:
            self._fire(sim, todo)


class after_sources(_Hook):
    def __init__(self, *fns):
        self.fns = fns

    def __call__(self, sim, todo="step"):
        if sim.round_time() >= sim._source_cutoff():
            self._fire(sim, todo)


class during_sources(_Hook):
    def __init__(self, *fns):
        self.fns = fns
        self._closed = False

    def __call__(self, sim, todo="step"):
        if sim.round_time() < sim._source_cutoff():
            self._fire(sim, "step")
        elif not self._closed:
            self._fire(sim, "finish")
            self._closed = True


class stop_when_fields_decayed:
    """Fed |c(pt)|^2 every step; every ``dt`` of simulated time the maximum of
    the window just ended is compared with the largest window maximum so far,
    and the run stops once it has fallen to ``decay_by`` of it."""

    def __init__(self, dt, c, pt, decay_by):
        self.dt = float(dt)
        self.c = c
        self.pt = pt
        self.decay_by = float(decay_by)
        self._t_window = 0.0
        self._peak = 0.0
        self._window = 0.0

    def __call__(self, sim):
        return self.observe(sim.get_field_point(self.c, self.pt), sim.round_time())

    def observe(self, value, t):
        """One step's field value at the end time t of the step."""
        self._window = max(self._window, abs(value) ** 2)
        if t <= self._t_window + self.dt:
            return False
        w, self._window = self._window, 0.0
        self._t_window = t
        self._peak = max(self._peak, w)
        return self._peak > 0 and w <= self.decay_by * self._peak

    def steps_to_check(self, step, dt):
        """Steps s >= 1 from step `step` until the first one whose end time
        (step + s) * dt closes the current window (the next comparison)."""
        limit = self._t_window + self.dt
        s = max(1, int(math.floor(limit / dt)) - step)
        while (step + s) * dt <= limit:
            s += 1
        while s > 1 and (step + s - 1) * dt > limit:
            s -= 1
        return s


def stop_after_walltime(seconds):
    t0 = time.time()
    return lambda sim: (time.time() - t0) > seconds


# ---- schedule of the run loop ---------------------------------------------
def plan_hooks(step_funcs, cutoff):
    """Sort step functions into (begin, per_step, scheduled):
    begin -- fired once before the first step;
    per_step -- (f, t_on, t_off): fired after every step whose end time t has
      t >= t_on and t - dt < t_off;
    scheduled -- [t_next, period or None, fns]: fired when t reaches t_next.
    at_end goes in none of them: the run's closing "finish" pass over the
    original step functions fires it."""
    begin, per_step, scheduled = [], [], []
    for f in step_funcs:
        if isinstance(f, at_beginning):
            begin.append(f)
        elif isinstance(f, at_end):
            continue
        elif isinstance(f, at_time):
            scheduled.append([f.t, None, f.fns])
        elif isinstance(f, at_every):
            scheduled.append([0.0, f.every, f.fns])
        elif isinstance(f, after_time):
            per_step += [(g, f.t, math.inf) for g in f.fns]
        elif isinstance(f, before_time):
            per_step += [(g, 0.0, f.t) for g in f.fns]
        elif isinstance(f, after_sources):
            per_step += [(g, cutoff, math.inf) for g in f.fns]
        elif isinstance(f, during_sources):
            per_step += [(g, 0.0, cutoff) for g in f.fns]
        elif callable(f):
            per_step.append((f, 0.0, math.inf))
        else:
            raise TypeError(f"unsupported step function {f!r}")
    return begin, per_step, scheduled


def chunk_size(t, dt, steps_left, per_step, scheduled, has_conds):
    """Yee steps the engine may take from time t before a hook or a stop
    condition must run (at least 1)."""
    if has_conds:
        return 1
    k = steps_left
    for _, t_on, t_off in per_step:
        if t + dt >= t_on - _EPS_T and t < t_off - _EPS_T:
            return 1
        if t < t_on:
            k = min(k, max(1, math.ceil((t_on - t) / dt - _EPS_T)))
    for t_next, _, _ in scheduled:
        k = min(k, max(1, math.ceil((t_next - t) / dt - _EPS_T)))
    return max(1, k)


def fire_per_step(sim, per_step, t, dt):
    for g, t_on, t_off in per_step:
        if t >= t_on - _EPS_T and t - dt < t_off - _EPS_T:
            eval_step_func(sim, g, "step")


def fire_scheduled(sim, scheduled, t):
    for entry in scheduled:
        while entry[0] is not None and t >= entry[0] - _EPS_T:
            for g in entry[2]:
                eval_step_func(sim, g, "step")
            entry[0] = entry[0] + entry[1] if entry[1] else None
    scheduled[:] = [e for e in scheduled if e[0] is not None]


__all__ = ["FluxData", "at_beginning", "at_end", "at_time", "at_every", "after_time", "before_time",
           "after_sources", "during_sources", "stop_when_fields_decayed", "stop_after_walltime",
           "plan_hooks", "chunk_size", "fire_per_step", "fire_scheduled", "eval_step_func"]
