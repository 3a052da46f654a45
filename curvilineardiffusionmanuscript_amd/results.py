"""Per-step timing records in the formats the reference's benchmark drivers
write and its plotting notebooks read.

* CSV with columns ``time, time_per_cycle, iter_per_cycle, error``: simulation
  time at the start of each cycle, the step's wall time (``stats.time``), its
  pseudo-iterations and its relative error
  (test/integrated/2d/gaussian/benchmark_nonlinear_2d_w_source.jl:178-199,263-268).
* JSON in the layout of ``TimerOutputs.todict``: nested sections with
  ``n_calls``, ``time_ns``, ``allocated_bytes``, ``total_time_ns``,
  ``total_allocated_bytes`` and ``inner_timers`` (same file :270-275), of which
  the notebooks read ``inner_timers["nonlinear_thermal_conduction_step!"]``
  (test/integrated/2d/gaussian/plot_nonlinear_benchmarks.jl:5-16).
"""
from __future__ import annotations

import csv
import json

STEP_SECTION = "nonlinear_thermal_conduction_step!"
COLUMNS = ("time", "time_per_cycle", "iter_per_cycle", "error")


class StepLog:
    """Collects what the reference's solve loop pushes per cycle and what its
    ``@timeit`` block around the step accumulates."""

    def __init__(self):
        self.time, self.time_per_cycle, self.iter_per_cycle, self.error = [], [], [], []
        self.n_calls = 0
        self.step_time_ns = 0

    def record(self, t, stats, wall_s=None):
        """One cycle at simulation time t: stats of the step (``time``, ``niter``,
        ``rel_err``), and the wall time of the timed block (default stats.time)."""
        self.time.append(float(t))
        self.time_per_cycle.append(float(stats.time))
        self.iter_per_cycle.append(float(stats.niter))
        self.error.append(float(stats.rel_err))
        self.n_calls += 1
        self.step_time_ns += int(round((stats.time if wall_s is None else wall_s) * 1e9))

    def reset_timer(self):
        """``reset_timer!()`` (the reference's loop resets after cycle 0)."""
        self.n_calls = 0
        self.step_time_ns = 0

    def write_csv(self, path):
        write_timing_csv(path, self.time, self.time_per_cycle, self.iter_per_cycle, self.error)

    def write_timer_json(self, path):
        write_timer_json(path, {STEP_SECTION: (self.n_calls, self.step_time_ns)})


def write_timing_csv(path, time, time_per_cycle, iter_per_cycle, error):
    n = len(time)
    if not (len(time_per_cycle) == len(iter_per_cycle) == len(error) == n):
        raise ValueError("the four columns must have one value per cycle")
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(COLUMNS)
        for row in zip(time, time_per_cycle, iter_per_cycle, error):
            w.writerow([repr(float(v)) for v in row])


def read_timing_csv(path):
    with open(path, newline="") as f:
        r = csv.reader(f)
        head = next(r)
        if tuple(head) != COLUMNS:
            raise ValueError(f"unexpected columns {head}")
        cols = {c: [] for c in COLUMNS}
        for row in r:
            for c, v in zip(COLUMNS, row):
                cols[c].append(float(v))
    return cols


def timer_dict(sections):
    """{name: (n_calls, time_ns)} -> the nested dict of TimerOutputs.todict, with
    the sections as inner timers of the root (no allocations are tracked: 0)."""
    inner = {}
    for name, (n, t) in sections.items():
        inner[name] = {"n_calls": int(n), "time_ns": int(t), "allocated_bytes": 0, "total_allocated_bytes": 0,
                       "total_time_ns": int(t), "inner_timers": {}}
    total = sum(v["time_ns"] for v in inner.values())
    return {"n_calls": 0, "time_ns": 0, "allocated_bytes": 0, "total_allocated_bytes": 0, "total_time_ns": total,
            "inner_timers": inner}


def write_timer_json(path, sections):
    with open(path, "w") as f:
        json.dump(timer_dict(sections), f, indent=2)


def runtimes(path):
    """(total seconds, milliseconds per cycle, calls) of the step section, as
    plot_nonlinear_benchmarks.jl:5-16 computes them."""
    with open(path) as f:
        d = json.load(f)
    s = d["inner_timers"][STEP_SECTION]
    tot_s = s["total_time_ns"] * 1e-9
    return tot_s, tot_s / s["n_calls"] * 1e3, s["n_calls"]
