"""Run control at diagnostic barriers, decided on the GPU (torchfdtd/run_control.py:15-159).

``RunMonitor`` watches one grid or a whole cohort.  At a diagnostic barrier it enqueues, on the
current stream, the FP64 state norms of every grid (tfdtd_state_norm: energy, field peak,
non-finite flag of E, H and the CPML memory), the run-control rule on those norms
(tfdtd_decay_update, one lane per grid) and an asynchronous copy of the per-grid verdicts into
pinned host memory, marked by an event.  Nothing waits for the GPU at a barrier unless the rule
could stop the run at that very barrier -- auto-shutoff with the checks since the last verdict
that has arrived able to reach the required run of decayed checks -- so the time loop keeps the
GPU fed between barriers.  A divergence verdict raises ``FloatingPointError`` when it arrives,
naming the barrier step at which the rule found it.

The rule itself (and the texts of its errors, which callers match) is the reference's: see
tfdtd_decay_update in include/tfdtd.h.
"""
from __future__ import annotations

import ctypes
import math
from collections import deque

import torch

from . import _native as N
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


def decay_rule(project, time_step):
    """The tfdtd_decay_rule of a project's run_control settings."""
    c = project.region.run_control
    rule = N.DecayRule()
    rule.eligible_time = source_end_time(project)+c.after_source_s
    rule.time_step = time_step
    rule.decay_threshold = c.decay_threshold
    rule.growth_limit = c.growth_limit
    rule.field_limit = 0. if c.field_limit is None else c.field_limit
    rule.min_steps = c.min_steps
    rule.consecutive_checks = c.consecutive_checks
    rule.divergence_check = int(c.divergence_check)
    rule.auto_shutoff = int(c.auto_shutoff)
    return rule


def verdict_error(state, label=''):
    """The FloatingPointError of a divergence verdict (texts of run_control.py:69-83, 159)."""
    step = int(state.verdict_step)
    if state.verdict == N.RUN_NONFINITE:
        text = ('Non-finite CPML or incident-line memory detected.' if state.verdict_value == 1 else
                f'Non-finite full-domain or auxiliary state at step {step}.')
    elif state.verdict == N.RUN_FIELD_LIMIT:
        text = f'Field magnitude limit exceeded at step {step}: {state.verdict_value:.6g}.'
    else:
        text = f'Post-source state norm grew by {state.verdict_value:.6g} at step {step}.'
    return FloatingPointError(f'{label}{text} (diagnostic step {step})')


class RunMonitor:
    """Device-side run control of `grids` (one project each) over at most `checks` barriers.

    check(step) enqueues one barrier; stopped(step) says whether the rule stopped the run at
    that barrier, waiting for the GPU only when it could have; finish() waits for every
    verdict and raises the first divergence; history(g) lists grid g's evaluated checks."""

    RING = 8     # pinned verdict buffers in flight

    def __init__(self, grids, projects, checks, labels=None, energies=None):
        from .diagnostics import StateNorms
        self.grids, self.projects = list(grids), list(projects)
        count = len(self.grids)
        self.lib = N.library()
        self.norms = StateNorms(self.grids)
        device = self.grids[0].device
        rules = (N.DecayRule*count)(*[decay_rule(p, g.time_step) for p, g in zip(self.projects, self.grids)])
        self.rules = [rules[k] for k in range(count)]
        self.rule_table = torch.frombuffer(bytearray(bytes(rules)), dtype=torch.uint8).to(device)
        self.state_bytes = ctypes.sizeof(N.DecayState)*count
        self.states = torch.zeros(self.state_bytes, dtype=torch.uint8, device=device)
        self.rows = max(1, int(checks))
        self.table = torch.zeros((count, self.rows, 6), dtype=torch.float64, device=device)
        self.labels = labels or ['']*count
        # Per grid, None or a function giving further FP64 energy on the device (dispersive oscillators).
        self.energies = energies or [None]*count
        self.ring = [torch.empty(self.state_bytes, dtype=torch.uint8, pin_memory=True) for _ in range(self.RING)]
        self.pending = deque()      # (step, ring slot, event), oldest first
        self.next_slot = 0
        self.known = self._parse(bytes(self.state_bytes))       # verdicts of the newest barrier read
        self.known_step = 0
        self.failed_step = None     # barrier step of the first divergence verdict read

    @property
    def backend(self):
        return 'hip'

    def _parse(self, raw):
        array = (N.DecayState*len(self.grids)).from_buffer_copy(raw)
        return [array[k] for k in range(len(self.grids))]

    def check(self, step):
        """Enqueue the norms, the rule and the verdict copy of barrier `step`."""
        stream = N.stream_handle()
        self.norms.launch()
        for g, extra in enumerate(self.energies):
            if extra is not None:
                self.norms.out[g, 0] += extra()
        N.check(self.lib.tfdtd_decay_update(self.rule_table.data_ptr(), self.states.data_ptr(), self.norms.out.data_ptr(),
                                            len(self.grids), int(step), self.table.data_ptr(), self.rows, stream),
                'tfdtd_decay_update')
        if len(self.pending) == self.RING:
            self._read_oldest()
        slot = self.next_slot
        self.next_slot = (slot+1) % self.RING
        self.ring[slot].copy_(self.states, non_blocking=True)
        event = torch.cuda.Event()
        event.record()
        self.pending.append((step, slot, event))

    def _read_oldest(self):
        step, slot, event = self.pending.popleft()
        event.synchronize()
        self.known = self._parse(self.ring[slot].numpy().tobytes())
        self.known_step = step
        for k, state in enumerate(self.known):
            if state.verdict not in (N.RUN_GOING, N.RUN_DECAYED):
                self.failed_step = int(state.verdict_step)
                raise verdict_error(state, self.labels[k])

    def poll(self):
        """Read every verdict that has arrived, without waiting."""
        while self.pending and self.pending[0][2].query():
            self._read_oldest()

    def wait(self, step=None):
        """Read every verdict up to barrier `step` (all pending when None), waiting for them."""
        while self.pending and (step is None or self.pending[0][0] <= step):
            self._read_oldest()

    def _eligible(self, k, step):
        rule = self.rules[k]
        return step >= rule.min_steps and step*rule.time_step >= rule.eligible_time

    def could_stop(self, step):
        """Whether any grid's rule could reach a stop at barrier `step` (enqueued already) given the
        verdicts read so far: each barrier not yet read may extend the run of decayed checks."""
        unread = sum(1 for s, _, _ in self.pending if s < step)
        for k, (rule, state) in enumerate(zip(self.rules, self.known)):
            if state.verdict == N.RUN_DECAYED:
                return True
            if rule.auto_shutoff and self._eligible(k, step) and state.below+unread+1 >= rule.consecutive_checks:
                return True
        return False

    def stopped(self, step):
        """True when the rule stopped the run at barrier `step` (the first grid that decayed)."""
        self.poll()
        if not self.could_stop(step):
            return False
        self.wait(step)
        return any(s.verdict == N.RUN_DECAYED and s.verdict_step == step for s in self.known)

    def latest(self, g=0):
        """The newest history row of grid g after waiting for every barrier enqueued so far."""
        self.wait()
        rows = self.history(g)
        return rows[-1] if rows else None

    def finish(self):
        self.wait()
        self.table_host = self.table.cpu().numpy()

    def history(self, g):
        count = min(self.known[g].checks, self.rows)
        table = getattr(self, 'table_host', None)
        rows = (table if table is not None else self.table.cpu().numpy())[g, :count]
        return [dict(step=int(r[0]), time_s=float(r[1]), state_norm=float(r[2]), relative_norm=float(r[3]),
                     field_peak=float(r[4]), source_finished=bool(r[5])) for r in rows]


def check_count(project, steps=None):
    """Number of diagnostic barriers of a run (every check_interval and the last step)."""
    r = project.region
    total = r.steps if steps is None else steps
    return total//r.run_control.check_interval+1

