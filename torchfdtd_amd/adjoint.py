"""Discrete-adjoint gradients of the HIP Yee/CPML run with respect to the permittivity.

The drop-in for the reference's differentiable API (torchfdtd/differentiable.py:755-900,
DifferentiableSimulation; torchfdtd/adjoint_spectrum.py, SpectralObservation;
torchfdtd/adjoint_planes.py:154-285, DifferentiablePlaneSimulation): epsilon -> point
signals, their spectra, or plane spectra, as torch operations whose backward is the
exact transpose of the discrete forward step.

How the backward works here.  The transposed step is linear in the adjoint state and
needs no forward state except curl H of each step's E update on the cells whose gradient
is wanted (the design box).  The forward run therefore records that curl on the box
(``tfdtd_record_curl``); the reverse sweep (``tfdtd_adjoint_step``: seeds, H transpose
with the gradient, E transpose) then runs from the last step to the first.  When the
record of every step fits the record budget (HBM is large) nothing is replayed.
Otherwise the forward run keeps its full state every K ~ sqrt(T) steps, and the reverse
sweep replays one K-step segment at a time from its stored state, recording the segment's
curls, before reversing it -- two passes over the forward instead of the reference's
binomial replays (differentiable.py:614-640).

The gradient is formed on the design box only (default: every cell outside the CPML
slabs of the E update); cells outside it get zero.  Sources are constant in epsilon.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as N
from .graphs import StepGraphs, observation_schedule
from .grid import march_family
from .models import Project
from .monitors import interpolation_map, plane_plan
from .planning import index_at, periodic_axes
from .plan import resolve_plan
from .solver import Stepper

COMPONENTS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


@dataclass(frozen=True)
class AdjointOptions:
    """checkpoint_interval: steps between stored forward states (0: record every step's curl and
    replay nothing; None: 0 when the record fits `record_budget_bytes`, else an even ~sqrt(T)).
    record_budget_bytes: default a quarter of the free device memory.  design_box: cells
    ((x0, x1), (y0, y1), (z0, z1)) whose gradient is formed (default: the whole grid, CPML cells
    included, as the reference returns dJ/deps over epsilon's shape; a box records a smaller
    curl).  cuda_graph_steps: replay width of the forward and reverse graphs."""
    checkpoint_interval: int | None = None
    record_budget_bytes: int | None = None
    design_box: tuple | None = None
    cuda_graph_steps: int = 32

    def __post_init__(self):
        if self.checkpoint_interval is not None and (isinstance(self.checkpoint_interval, bool) or self.checkpoint_interval < 0):
            raise ValueError('checkpoint_interval must be a non-negative integer or None.')
        if not 1 <= self.cuda_graph_steps <= 64:
            raise ValueError('cuda_graph_steps must be an integer from 1 to 64.')


@dataclass
class DifferentiableResult:
    """Point signals (steps, monitors) and their spectra (differentiable.py:65-87)."""
    signals: torch.Tensor
    time_step: float
    monitor_components: tuple
    report: dict = field(default_factory=dict)

    def spectrum(self, frequency_hz, *, window=None):
        """dt * sum_t x(t) w(t) exp(-2 pi i f t), t = (n+1) dt (+dt/2 for H), in FP64."""
        signals = self.signals.to(torch.float64)
        f = torch.as_tensor(frequency_hz, device=signals.device, dtype=torch.float64).reshape(-1)
        if not bool(torch.isfinite(f).all()) or bool((f <= 0).any()):
            raise ValueError('Frequencies must be finite and positive.')
        if bool((f*self.time_step >= .5).any()):
            raise ValueError('Frequencies must be below temporal Nyquist.')
        t = (torch.arange(signals.shape[0], device=signals.device, dtype=torch.float64)+1)*self.time_step
        weights = torch.ones_like(t) if window is None else torch.as_tensor(window, device=t.device, dtype=t.dtype)
        values = []
        for i, comp in enumerate(self.monitor_components):
            times = t+(self.time_step/2 if comp[0] == 'H' else 0)
            kernel = torch.exp(-2j*torch.pi*f[:, None]*times[None, :])
            values.append(kernel@(signals[:, i]*weights).to(kernel.dtype)*self.time_step)
        return torch.stack(values, dim=-1)


@dataclass
class DifferentiablePlaneResult:
    """Spectral fields of one plane (adjoint_planes.py:69-139): fields (frequency, point, 6)."""
    fields: torch.Tensor
    frequency_hz: torch.Tensor
    points_um: torch.Tensor
    weights: torch.Tensor
    shape: tuple
    normal: str
    report: dict
    components: tuple = COMPONENTS
    fourier_convention: str = 'exp(+2 pi i f t)'

    def poynting(self, normal=None):
        axis = 'xyz'.index(self.normal if normal is None else normal)
        b, c = (axis+1) % 3, (axis+2) % 3
        return .5*(self.fields[..., b]*self.fields[..., 3+c].conj()-self.fields[..., c]*self.fields[..., 3+b].conj()).real

    def flux(self):
        return self.poynting()@self.weights


def _segment_widths(steps, interval, width):
    """Graph widths that cover the sweep's segments without single-step replays: the checkpoint
    interval (or the graph width) and the remainders the run leaves."""
    widths = set()
    for span in ((interval, steps-(steps-1)//interval*interval) if interval else (steps,)):
        if span > 64:
            widths.update({width, span % width})
        else:
            widths.add(span)
    return tuple(sorted(w for w in widths if w > 0))


def _even_interval(steps):
    k = max(2, math.ceil(math.sqrt(steps)))
    return k+(k % 2)


class _Engine:
    """Forward run with observers and the curl record, and its reverse sweep."""

    def __init__(self, project, epsilon, observers, options):
        r = project.region
        self.options = options
        self.steps = r.steps
        self.run = run = Stepper(project, kernel='hip', epsilon=epsilon, two_step=False)
        run.set_trace_points(observers)
        g = self.grid = run.grid
        self.lib = N.library()
        self.real = g.dtype
        self.dtype = N.F64 if g.dtype == torch.float64 else N.F32
        self.per_component = epsilon.ndim == 4
        self.box = self._box(options.design_box)
        self.box_shape = tuple(hi-lo for lo, hi in self.box)
        self.box_cells = math.prod(self.box_shape)
        item = g.E.element_size()
        every = self.box_cells*3*item*self.steps
        budget = options.record_budget_bytes
        if budget is None:
            budget = torch.cuda.mem_get_info(g.device)[0]//4
        if options.checkpoint_interval is None:
            interval = 0 if every <= budget else _even_interval(self.steps)
        else:
            # Any interval serves: every graph replay (pairs, then one single step for an odd count)
            # leaves the state in the grid's own buffers, where checkpoints are taken.
            interval = options.checkpoint_interval
        self.interval = interval if interval < self.steps else 0
        self.modulus = self.steps if self.interval == 0 else self.interval
        self.record = torch.zeros((self.modulus, self.box_cells, 3), device=g.device, dtype=self.real)
        # Each record reads the E-family CPML memory its step's E update reads: the grid's own
        # buffer for single steps and the first step of a pair, the second set for the pair's second.
        self.records = [self._record_descriptor(g.H, g.psi_buffers[False], g.psi_buffers[False])]
        if run.one_pass:
            self.records.append(self._record_descriptor(run.H_b, run.psi_e_b, g.psi_buffers[False]))
        width = min(options.cuda_graph_steps, self.steps)
        self.width = width
        self.widths = _segment_widths(self.steps, self.interval, width)
        self.forward_graphs = StepGraphs(self._step, run.states()+[self.record], self.steps, width,
                                         pair=self._pair if run.one_pass else None,
                                         prepare=run.refresh_materials if run.one_pass else None, widths=self.widths)
        self.report = dict(adjoint='discrete Yee/CPML (HIP transpose)', steps=self.steps, design_box=[list(b) for b in self.box],
                           checkpoint_interval=self.interval, record_bytes=self.record.numel()*item,
                           forward_kernel='one-pass' if run.one_pass else 'split', observers=len(observers))

    # ------------------------------------------------------------------ set-up
    def _box(self, box):
        r = self.grid.region
        if box is None:
            box = tuple((0, n) for n in r.shape)
        box = tuple((int(a), int(b)) for a, b in box)
        if len(box) != 3 or any(not 0 <= a < b <= n for (a, b), n in zip(box, r.shape)):
            raise ValueError(f'design box {box} must lie inside the grid {tuple(r.shape)}')
        return box

    def _record_descriptor(self, H, psi_e, psi_e_out):
        d = N.CurlRecord()
        d.e = march_family(self.grid, False, psi_e, psi_e_out)
        d.dtype = self.dtype
        d.nx, d.ny, d.nz = self.grid.Nx, self.grid.Ny, self.grid.Nz
        d.H = H.data_ptr()
        d.periodic_mask = sum(1 << a for a in periodic_axes(self.grid.region))
        for a in range(3):
            d.box_lo[a], d.box_hi[a] = self.box[a]
        d.record = self.record.data_ptr()
        d.modulus = self.modulus
        d.counter = self.run.counter.data_ptr()
        return d

    # ------------------------------------------------------------------ forward
    def _step(self):
        N.check(self.lib.tfdtd_record_curl(ctypes.byref(self.records[0]), N.stream_handle()), 'tfdtd_record_curl')
        self.run.step()

    def _pair(self):
        run, lib, stream = self.run, self.lib, N.stream_handle()
        for k, record, observe in ((0, self.records[0], run.observe_b), (1, self.records[1], run.observe)):
            N.check(lib.tfdtd_record_curl(ctypes.byref(record), stream), 'tfdtd_record_curl')
            run.march_step(k, stream)
            N.check(lib.tfdtd_observe_step(ctypes.byref(observe), stream), 'tfdtd_observe_step')

    def _advance(self, count):
        self.forward_graphs.advance(count)

    def _state(self):
        g = self.grid
        return [g.E, g.H, g.psi_buffers[False], g.psi_buffers[True]]

    def forward(self):
        started = time.perf_counter()
        self.checkpoints = {}
        done = 0
        marks = range(self.interval, self.steps, self.interval) if self.interval else ()
        for mark in [*marks, self.steps]:
            self._advance(mark-done)
            done = mark
            if mark < self.steps:
                self.checkpoints[mark] = [t.clone() for t in self._state()]
        torch.cuda.synchronize()
        self.report['forward_seconds'] = time.perf_counter()-started
        self.report['checkpoints'] = len(self.checkpoints)
        return self.run.traces

    # ------------------------------------------------------------------ backward
    def _adjoint_descriptors(self, seeds):
        g, run = self.grid, self.run
        self.e_bar, self.h_bar = torch.zeros_like(g.E), torch.zeros_like(g.H)
        self.psi_bar = {f: [torch.zeros_like(g.psi_buffers[f]), torch.zeros_like(g.psi_buffers[f])] for f in (False, True)}
        self.gradient = torch.zeros(self.box_cells*(3 if self.per_component else 1), device=g.device, dtype=torch.float64)
        self.reverse_counter = torch.zeros(1, device=g.device, dtype=torch.long)
        # Seeds: one target per observer (the observers are distinct elements).
        points = run.trace_points
        targets = np.array([(int(m), flat) for m, flat in points], dtype=np.int64).reshape(-1, 2)
        self.seed_targets = torch.as_tensor(targets, device=g.device)
        self.seed_offsets = torch.arange(len(points)+1, device=g.device, dtype=torch.int32)
        self.seed_columns = torch.arange(len(points), device=g.device, dtype=torch.int32)
        self.seeds = seeds
        descriptors = []
        for parity in range(2):
            d = N.Adjoint()
            d.dtype = self.dtype
            d.nx, d.ny, d.nz = g.Nx, g.Ny, g.Nz
            d.e_bar, d.h_bar = self.e_bar.data_ptr(), self.h_bar.data_ptr()
            d.inverse_permittivity = g.inverse_permittivity.data_ptr()
            d.eps_stride = g.inverse_permittivity.shape[-1]
            d.inverse_permeability = g.inverse_permeability.data_ptr()
            d.courant = g.courant_number
            d.periodic_mask = sum(1 << a for a in periodic_axes(g.region))
            d.pec_upper_mask = sum(1 << a for a in g.pec_upper)
            for a, f in g.pec_upper.items():
                d.pec_upper_factor[a] = f
            for name, forward in (('e', False), ('h', True)):
                bars = self.psi_bar[forward]
                setattr(d, name, march_family(g, forward, bars[parity], bars[1-parity]))
            d.nseed = len(points)
            if points:
                d.seed_targets, d.seed_offsets = self.seed_targets.data_ptr(), self.seed_offsets.data_ptr()
                d.seed_columns, d.seeds = self.seed_columns.data_ptr(), self.seeds.data_ptr()
                d.seed_width, d.seed_rows = len(points), self.seeds.shape[0]
            for a in range(3):
                d.box_lo[a], d.box_hi[a] = self.box[a]
            d.record, d.record_modulus = self.record.data_ptr(), self.modulus
            d.gradient, d.gradient_stride = self.gradient.data_ptr(), 3 if self.per_component else 1
            d.record_cpml = 1
            d.counter = self.reverse_counter.data_ptr()
            descriptors.append(d)
        return descriptors

    def _adjoint_step(self):
        N.check(self.lib.tfdtd_adjoint_step(ctypes.byref(self.adjoint[0]), N.stream_handle()), 'tfdtd_adjoint_step')
        for bars in self.psi_bar.values():
            bars[0].copy_(bars[1])

    def _adjoint_pair(self):
        stream = N.stream_handle()
        for d in self.adjoint:
            N.check(self.lib.tfdtd_adjoint_step(ctypes.byref(d), stream), 'tfdtd_adjoint_step')

    def _reverse(self, count):
        self.reverse_graphs.advance(count)

    def backward(self, signal_bar):
        started = time.perf_counter()
        bar = signal_bar.detach().to(device=self.grid.device)
        if bar.shape != (self.steps, len(self.run.trace_points)):
            raise ValueError('signal gradient has the wrong shape')
        # The sweep is linear in the seeds: scale them by a power of two that brings the largest to
        # about one (exact in floating point), so that FP32 adjoint fields stay far from the
        # subnormal range (objectives of DFT spectra carry dt^2 ~ 1e-32), and unscale the gradient.
        peak = float(bar.abs().max()) if bar.numel() else 0.
        self.seed_scale = 2.**-math.floor(math.log2(peak)) if peak > 0 and math.isfinite(peak) else 1.
        seeds = (bar.to(torch.float64)*self.seed_scale).to(self.real).contiguous()
        self.adjoint = self._adjoint_descriptors(seeds)
        states = [self.e_bar, self.h_bar, *self.psi_bar[False], *self.psi_bar[True], self.gradient, self.reverse_counter]
        self.reverse_graphs = StepGraphs(self._adjoint_step, states, self.steps, self.width, pair=self._adjoint_pair,
                                         widths=self.widths)
        replayed = 0
        if not self.interval:
            self.reverse_counter.fill_(self.steps-1)
            self._reverse(self.steps)
        else:
            starts = list(range(0, self.steps, self.interval))
            for start in reversed(starts):
                end = min(self.steps, start+self.interval)
                saved = self.checkpoints.get(start)
                for target, value in zip(self._state(), saved or [None]*4):
                    target.copy_(value) if value is not None else target.zero_()
                self.run.counter.fill_(start)
                self._advance(end-start)
                replayed += end-start
                self.reverse_counter.fill_(end-1)
                self._reverse(end-start)
        torch.cuda.synchronize()
        self.report['backward_seconds'] = time.perf_counter()-started
        self.report['replayed_steps'] = replayed
        return self.gradient


def _full_gradient(engine, epsilon):
    grad = torch.zeros(epsilon.shape, device=epsilon.device, dtype=torch.float64)
    box = tuple(slice(lo, hi) for lo, hi in engine.box)
    shape = engine.box_shape+((3,) if engine.per_component else ())
    grad[box] = engine.gradient.view(shape).to(epsilon.device)/engine.seed_scale
    return grad.to(epsilon.dtype)


class _Simulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, epsilon, engine):
        ctx.engine = engine
        ctx.shape_like = epsilon.detach()
        return engine.forward().clone()

    @staticmethod
    def backward(ctx, signal_bar):
        if torch.is_grad_enabled():
            raise RuntimeError('Higher-order derivatives are not supported by the discrete adjoint.')
        ctx.engine.backward(signal_bar)
        return _full_gradient(ctx.engine, ctx.shape_like), None


def _check_epsilon(epsilon, region):
    if not isinstance(epsilon, torch.Tensor) or epsilon.dtype not in (torch.float32, torch.float64):
        raise ValueError('epsilon must be a real float32 or float64 torch Tensor.')
    shape = tuple(region.shape)
    if tuple(epsilon.shape) not in (shape, shape+(3,)):
        raise ValueError(f'epsilon shape must be {shape}, optionally with three Yee components.')
    if (epsilon.dtype == torch.float64) != (region.precision == 'float64'):
        raise ValueError('epsilon dtype must match the project precision.')
    if not bool(torch.isfinite(epsilon).all()) or bool((epsilon < 1).any()):
        raise ValueError('This conservative CFL contract requires finite epsilon >= 1.')


def _internal_project(project):
    internal = Project.model_validate(project.model_dump())
    internal.monitors = []
    return internal


def _admit(project):
    r = project.region
    refusal = r.unsupported()
    if refusal:
        raise ValueError(refusal)
    if r.complex_fields:
        raise ValueError('The HIP adjoint runs real fields.')
    if r.run_control.auto_shutoff:
        raise ValueError('Differentiable simulations require a fixed number of timesteps.')
    active = {s.material for s in project.structures if s.enabled}
    if any(m.oscillators for m in project.materials if m.name in active):
        raise ValueError('The HIP adjoint runs nondispersive materials; dispersive structures are not differentiated.')
    if not torch.cuda.is_available():
        raise RuntimeError('The HIP adjoint requires a ROCm GPU.')


class DifferentiableSimulation(torch.nn.Module):
    """epsilon -> point signals with a discrete-adjoint backward (differentiable.py:755-900)."""

    def __init__(self, project: Project, options: AdjointOptions | None = None):
        super().__init__()
        self.project = Project.model_validate(project.model_dump())
        self.options = options or AdjointOptions()
        active = [m for m in self.project.monitors if m.enabled]
        if not active or any(m.kind != 'point' for m in active):
            raise ValueError('The differentiable API requires point monitors only.')
        _admit(self.project)
        r = self.project.region
        self.monitors = [self.project.resolved_monitor(m) for m in active]
        self.observers = []
        for m in self.monitors:
            loc = index_at(m.center, r, m.component)
            flat = ((loc[0]*r.shape[1]+loc[1])*r.shape[2]+loc[2])*3+'xyz'.index(m.component[1].lower())
            self.observers.append((m.component[0] == 'H', flat))

    @property
    def plan(self):
        """The resolved plan of the fixed project (plan.resolve_plan), shared with every entry point."""
        if getattr(self, '_plan', None) is None:
            self._plan = resolve_plan(self.project)
        return self._plan

    @property
    def plan_hash(self):
        return self.plan.plan_hash

    def _run(self, epsilon):
        _check_epsilon(epsilon, self.project.region)
        observers = sorted(set(self.observers))
        engine = _Engine(_internal_project(self.project), epsilon.detach(), observers, self.options)
        columns = torch.as_tensor([observers.index(o) for o in self.observers], device=engine.grid.device)
        signals = _Simulate.apply(epsilon, engine)
        return signals.index_select(1, columns), engine.report

    def forward(self, epsilon):
        signals, report = self._run(epsilon)
        return DifferentiableResult(signals, self.project.region.time_step, tuple(m.component for m in self.monitors), report)

    def spectrum(self, epsilon, frequency_hz, *, window=None):
        return self.forward(epsilon).spectrum(frequency_hz, window=window)


class DifferentiablePlaneSimulation(torch.nn.Module):
    """epsilon -> spectral fields on planes (adjoint_planes.py:154-285): every enabled monitor is
    a field plane with time_downsample 1 and no apodization; all six components are returned."""

    def __init__(self, project: Project, options: AdjointOptions | None = None):
        super().__init__()
        self.project = Project.model_validate(project.model_dump())
        self.options = options or AdjointOptions()
        _admit(self.project)
        active = [m for m in self.project.monitors if m.enabled]
        if not active or any(m.kind != 'field' for m in active):
            raise ValueError('DifferentiablePlaneSimulation requires enabled field monitors only.')
        r = self.project.region
        self.plans, observers, lookup = [], [], {}
        for raw in active:
            monitor = self.project.resolved_monitor(raw)
            if monitor.time_downsample != 1 or monitor.spectrum.apodization != 'none':
                raise ValueError('Differentiable planes require time_downsample=1 and no apodization.')
            plan = plane_plan(r, monitor)
            maps = []
            for component in COMPONENTS:
                indices, weights = interpolation_map(r, component, plan['points_um'])
                flat = indices.reshape(-1)
                numbers = np.empty(flat.shape, dtype=np.int64)
                known = lookup.setdefault(component, {})
                for k, index in enumerate(flat.tolist()):
                    number = known.get(index)
                    if number is None:
                        number = known[index] = len(observers)
                        observers.append((component[0] == 'H', index))
                    numbers[k] = number
                maps.append((numbers.reshape(indices.shape), weights))
            self.plans.append((monitor.id, monitor.normal, plan, maps))
        self.observers = observers
        self.components = tuple(('H' if m else 'E') for m, _ in observers)

    @property
    def plan(self):
        """The resolved plan of the fixed project (plan.resolve_plan), shared with every entry point."""
        if getattr(self, '_plan', None) is None:
            self._plan = resolve_plan(self.project)
        return self._plan

    @property
    def plan_hash(self):
        return self.plan.plan_hash

    def forward(self, epsilon, frequency_hz):
        _check_epsilon(epsilon, self.project.region)
        r = self.project.region
        engine = _Engine(_internal_project(self.project), epsilon.detach(), self.observers, self.options)
        signals = _Simulate.apply(epsilon, engine).to(torch.float64)
        device = signals.device
        f = torch.as_tensor(frequency_hz, device=device, dtype=torch.float64).reshape(-1)
        if not bool(torch.isfinite(f).all()) or bool((f <= 0).any()) or bool((f*r.time_step >= .5).any()):
            raise ValueError('Frequencies must be finite, positive and below temporal Nyquist.')
        t = (torch.arange(signals.shape[0], device=device, dtype=torch.float64)+1)*r.time_step
        magnetic = torch.as_tensor([c == 'H' for c in self.components], device=device)
        spectra = torch.zeros((f.numel(), signals.shape[1]), device=device, dtype=torch.complex128)
        for family in (False, True):
            columns = torch.nonzero(magnetic == family).flatten()
            if columns.numel():
                kernel = torch.exp(2j*torch.pi*f[:, None]*(t+(r.time_step/2 if family else 0))[None, :])*r.time_step
                spectra = spectra.index_copy(1, columns, kernel@signals.index_select(1, columns).to(torch.complex128))
        output = OrderedDict()
        for identifier, normal, plan, maps in self.plans:
            fields = []
            for numbers, weights in maps:
                index = torch.as_tensor(numbers, device=device)
                weight = torch.as_tensor(weights, device=device, dtype=torch.float64)
                fields.append((spectra[:, index]*weight[None]).sum(dim=1))
            output[identifier] = DifferentiablePlaneResult(
                torch.stack(fields, dim=-1), f.clone(), torch.as_tensor(plan['points_um'], device=device),
                torch.as_tensor(plan['weights'], device=device), plan['shape'], normal, engine.report)
        return output
