"""Host logic of the step-function protocol (meep_torch/mp/run_control.py
:199-222, :225-322, :457-519) on a stand-in simulation object: which hook
fires when, how many Yee steps one engine call may take between hooks, the
calling convention, and the stop conditions.  The expected schedules are
worked out by hand from meep's rules (fire after the step whose end time
reaches the hook time; at_every from t = 0)."""
import math

import pytest

from meep_torch_amd.mp import run_control as RC


class _Grid:
    def __init__(self, dt):
        self.dt = dt


class FakeSim:
    """Steps a clock through the same loop shape as Simulation._run_until."""

    def __init__(self, dt=0.05, cutoff=3.0, field=None):
        self.grid = _Grid(dt)
        self.t_step = 0
        self.cutoff = cutoff
        self.field = field or (lambda t: 0.0)
        self.calls = []

    def round_time(self):
        return self.t_step * self.grid.dt

    def _source_cutoff(self):
        return self.cutoff

    def get_field_point(self, c, pt):
        return self.field(self.round_time())

    def run(self, *step_funcs, n_total=None, conds=None, gate=0.0):
        dt = self.grid.dt
        begin, per_step, scheduled = RC.plan_hooks(step_funcs, self.cutoff)
        stop_at = None if n_total is None else self.t_step + n_total
        for g in begin:
            RC.eval_step_func(self, g, "step")
        RC.fire_scheduled(self, scheduled, self.t_step * dt)
        chunks = []
        while stop_at is None or self.t_step < stop_at:
            t = self.t_step * dt
            left = stop_at - self.t_step if stop_at is not None else 1 << 30
            k = RC.chunk_size(t, dt, left, per_step, scheduled, conds is not None)
            chunks.append(k)
            self.t_step += k
            t = self.t_step * dt
            RC.fire_per_step(self, per_step, t, dt)
            RC.fire_scheduled(self, scheduled, t)
            if conds is not None and t >= gate - 1e-9 and any([bool(g(self)) for g in conds]):
                break
        for f in step_funcs:
            RC.eval_step_func(self, f, "finish")
        return chunks


def test_at_every_fires_on_its_grid_and_sizes_chunks():
    sim, seen = FakeSim(), []
    chunks = sim.run(RC.at_every(5.0, lambda s: seen.append(round(s.round_time(), 9))), n_total=800)
    assert seen == [0.0, 5.0, 10.0, 15.0, 20.0, 25.0, 30.0, 35.0, 40.0]
    assert chunks == [100] * 8           # 5.0 / dt = 100 steps between hooks


def test_hook_free_run_is_one_engine_call():
    assert FakeSim().run(n_total=800) == [800]


def test_after_time_and_before_time_windows():
    sim, a, b = FakeSim(dt=0.5), [], []
    chunks = sim.run(RC.after_time(3.0, lambda s: a.append(s.t_step)),
                     RC.before_time(1.0, lambda s: b.append(s.t_step)), n_total=10)
    assert b == [1, 2]                   # t = 0.5 and 1.0: t - dt < 1.0
    assert a == [6, 7, 8, 9, 10]         # t >= 3.0
    # before_time is active for the first two steps; then one call of 4 steps
    # reaches t = 3.0, where after_time fires for the first time; from there
    # on it is active in every coming step
    assert chunks == [1, 1, 4, 1, 1, 1, 1]


def test_at_time_at_beginning_at_end_and_plain_functions():
    sim, log = FakeSim(dt=0.25), []
    sim.run(RC.at_beginning(lambda s: log.append(("begin", s.t_step))),
            RC.at_time(1.0, lambda s: log.append(("at", s.t_step))),
            RC.at_end(lambda s: log.append(("end", s.t_step))), n_total=8)
    assert log == [("begin", 0), ("at", 4), ("end", 8)]
    plain = []
    FakeSim(dt=0.25).run(lambda s: plain.append(s.t_step), n_total=3)
    assert plain == [1, 2, 3]            # a bare f(sim) runs after every step


def test_two_argument_functions_see_finish_and_bad_arity_raises():
    todo = []
    FakeSim().run(lambda s, t: todo.append(t), n_total=2)
    assert todo == ["step", "step", "finish"]
    with pytest.raises(ValueError):
        RC.eval_step_func(FakeSim(), lambda a, b, c: None, "step")


def test_sources_windows_use_the_cutoff():
    sim, during, after = FakeSim(dt=1.0, cutoff=3.0), [], []
    sim.run(RC.during_sources(lambda s: during.append(s.t_step)),
            RC.after_sources(lambda s: after.append(s.t_step)), n_total=5)
    assert during == [1, 2, 3] and after == [3, 4, 5]


def test_stop_condition_runs_single_steps():
    sim = FakeSim()
    chunks = sim.run(conds=[lambda s: s.t_step >= 7])
    assert chunks == [1] * 7 and sim.t_step == 7


def test_stop_when_fields_decayed_closed_form():
    """|f|^2 = exp(-t) sampled every 0.1: windows of 1.0-1.1 time units, each
    window's maximum is its first sample, and the first window's is
    exp(-0.1).  The run stops at the end of the first window whose first
    sample s has exp(-s) <= 1e-3 exp(-0.1), i.e. s >= 0.1 + ln 1000 = 7.008;
    the window before it started below 7.008, so s < 7.008 + 1.1, and the
    stop comes 1.0-1.1 after s: between 8.0 and 9.2."""
    sim = FakeSim(dt=0.1, field=lambda t: math.exp(-t / 2))
    sim.run(conds=[RC.stop_when_fields_decayed(1.0, "Ez", None, 1e-3)])
    assert 8.0 <= sim.round_time() <= 9.2


def test_chunk_size_never_passes_a_scheduled_time():
    sch = [[2.0, 0.5, ()]]
    assert RC.chunk_size(0.0, 0.1, 1000, [], sch, False) == 20
    assert RC.chunk_size(1.95, 0.1, 1000, [], sch, False) == 1
    assert RC.chunk_size(0.0, 0.1, 7, [], sch, False) == 7
