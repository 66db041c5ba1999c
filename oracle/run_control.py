"""Oracle: the run-control decision of torchfdtd/run_control.py:58-85 and the FP64 state norm of
torchfdtd/run_control.py:127-159 (uniform meshes, no dispersive or incident-line state), in plain
Python and NumPy.  Test infrastructure only.
"""
from __future__ import annotations

import math

import numpy as np


def state_norm(E, H, inverse, memories=(), materials=(), dt=None):
    """(energy, peak, finite): sum_c E_c^2/inv_eps_c + H_c^2 in FP64, plus the oscillator energies
    (Q/(dt sqrt s))^2 + (P w0/sqrt s)^2 of dispersive materials, max |E|,|H|, and whether every CPML
    memory array is finite (run_control.py:131-159; unit cell volumes)."""
    e, h, inv = (np.asarray(a, dtype=np.float64) for a in (E, H, inverse))
    energy = 0.
    for j in range(3):
        energy = energy+float((e[..., j]*e[..., j]*(1/inv[..., j if inv.shape[-1] == 3 else 0])).sum())
        energy = energy+float((h[..., j]*h[..., j]).sum())
    for state in materials:
        for j, (w0, strength, _) in enumerate(state.oscillators):
            p = np.asarray(state.P[j] if state.multiple else state.P, dtype=np.float64)
            q = np.asarray(state.Q[j] if state.multiple else state.Q, dtype=np.float64)
            energy = energy+float(((q/(dt*np.sqrt(strength)))**2).sum())
            energy = energy+float(((p*(w0/np.sqrt(strength)))**2).sum())
    peak = float(max(np.abs(e).max(), np.abs(h).max()))
    finite = all(bool(np.isfinite(m).all()) for m in memories)
    return energy, peak, finite


class Rule:
    """run_control.py:58-85: eligibility after min_steps and the end of the sources, energy baseline
    frozen at the first eligible check, growth and field limits, consecutive decayed checks."""

    def __init__(self, eligible_time, dt, decay_threshold=1e-6, growth_limit=1e6, field_limit=None, min_steps=100,
                 consecutive_checks=3, divergence_check=True, auto_shutoff=True):
        self.eligible_time, self.dt = eligible_time, dt
        self.decay_threshold, self.growth_limit, self.field_limit = decay_threshold, growth_limit, field_limit
        self.min_steps, self.consecutive_checks = min_steps, consecutive_checks
        self.divergence_check, self.auto_shutoff = divergence_check, auto_shutoff
        self.peak, self.frozen, self.below, self.history = 0., None, 0, []

    def update(self, step, energy, field_peak):
        """True when the run stops at this check; FloatingPointError on divergence."""
        if not math.isfinite(energy) or not math.isfinite(field_peak):
            raise FloatingPointError(f'Non-finite full-domain or auxiliary state at step {step}.')
        if self.divergence_check and self.field_limit is not None and field_peak > self.field_limit:
            raise FloatingPointError(f'Field magnitude limit exceeded at step {step}: {field_peak:.6g}.')
        eligible = step >= self.min_steps and step*self.dt >= self.eligible_time
        self.peak = max(self.peak, energy)
        if eligible and self.frozen is None:
            self.frozen = self.peak
        base = self.peak if self.frozen is None else self.frozen
        ratio = energy/base if base else 0.
        if self.divergence_check and eligible and ratio > self.growth_limit:
            raise FloatingPointError(f'Post-source state norm grew by {ratio:.6g} at step {step}.')
        self.below = self.below+1 if eligible and ratio <= self.decay_threshold else 0
        self.history.append(dict(step=step, time_s=step*self.dt, state_norm=energy, relative_norm=ratio,
                                 field_peak=field_peak, source_finished=eligible))
        return self.auto_shutoff and self.below >= self.consecutive_checks
