"""Run-control diagnostics at host barriers (torchfdtd/run_control.py:15-159).

The reduction (FP64 volume-weighted E/H norm, field peak, finiteness of every CPML
memory) runs on the GPU; only three scalars cross to the host per barrier.
"""
from __future__ import annotations

import math

from .planning import TAIL_OUTER, pulse_parameters


def source_end_time(project):
    """Conservative end of the last soft-source envelope (run_control.py:15-55, soft sources)."""
    cutoff = project.region.run_control.source_tail_amplitude
    ends = [0.]
    for raw in project.sources:
        s = project.resolved_source(raw)
        if not s.enabled or s.amplitude == 0:
            continue
        if s.pulse == 'continuous':
            return math.inf
        if s.pulse == 'sampled':
            ends.append(s.signal.time_s[-1])
            continue
        p = pulse_parameters(s)
        tail = TAIL_OUTER if s.eliminate_discontinuities else math.sqrt(-2*math.log(cutoff))
        ends.append(p.offset_s+tail*p.sigma_s)
    return max(ends)


class DecayDecision:
    """Backend-independent stop/divergence decision (run_control.py:58-85)."""

    def __init__(self, control, source_end_s, dt):
        self.control, self.dt = control, dt
        self.eligible_time = source_end_s+control.after_source_s
        self.peak = 0.
        self.frozen_peak = None
        self.below = 0
        self.history = []

    def update(self, step, energy, field_peak):
        c = self.control
        if not math.isfinite(energy) or not math.isfinite(field_peak):
            raise FloatingPointError(f'Non-finite full-domain or auxiliary state at step {step}.')
        if c.divergence_check and c.field_limit is not None and field_peak > c.field_limit:
            raise FloatingPointError(f'Field magnitude limit exceeded at step {step}: {field_peak:.6g}.')
        eligible = step >= c.min_steps and step*self.dt >= self.eligible_time
        self.peak = max(self.peak, energy)
        if eligible and self.frozen_peak is None:
            self.frozen_peak = self.peak
        baseline = self.frozen_peak if self.frozen_peak is not None else self.peak
        ratio = energy/baseline if baseline else 0.
        if c.divergence_check and eligible and ratio > c.growth_limit:
            raise FloatingPointError(f'Post-source state norm grew by {ratio:.6g} at step {step}.')
        self.below = self.below+1 if eligible and ratio <= c.decay_threshold else 0
        self.history.append(dict(step=step, time_s=step*self.dt, state_norm=energy,
                                 relative_norm=ratio, field_peak=field_peak, source_finished=eligible))
        return c.auto_shutoff and self.below >= c.consecutive_checks


class StateDiagnostics:
    """Sum_c vol E_c^2 / inv_eps_c + vol H_c^2 in FP64 and max |E|,|H| of one grid (HIP reduction),
    raising FloatingPointError when E, H or the CPML memory holds a non-finite value."""

    backend = 'hip'

    def __init__(self, grid):
        from .diagnostics import StateNorms
        self.grid = grid
        self.norms = StateNorms([grid])

    def measure(self):
        energy, peak, bad = self.norms.values()[0]
        if bad:
            raise FloatingPointError('Non-finite full-domain or CPML state detected.')
        return float(energy), float(peak)
