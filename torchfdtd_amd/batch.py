"""Uneven batches: many projects of any shape advance together on one GPU.

The reference batches only projects of identical topology -- shape, nodes, steps and
boundaries (torchfdtd/tensor_batch.py:32-41, 136-141; cuda_batch.py:34-35) -- and runs
anything else as one cohort per topology, one after another (grouped_batch.py:38-77).
Here a cohort only needs a common step count and precision: every grid keeps its own
shape, CPML layout, sources, traces and planes, and each kernel launch covers every
grid of the cohort through a device table of per-grid descriptors.

Per step (split kernels): one E launch, one H launch, one plane-DFT launch and one
trace/counter launch for the whole cohort.  Per pair of steps (one-pass kernels):
three tile launches (interior, z-face, boundary) for each step, plus the DFT and
trace launches.  The step loop is captured in HIP graphs exactly as for one run.

Entry points mirror the reference: run_tensor_batch (tensor_batch.py:44-113),
plan_grouped_batch / run_grouped_batch (grouped_batch.py:80-146) and
tune_tensor_batch (tuning.py:63-160), plus BatchCase / BatchItem / BatchReport
(batch.py:31-80).  tune_tensor_batch selects the launch configuration (split or
one-pass kernels, tile height, x chunk) by timing every candidate on the real
cohort and refuses candidates whose states differ in a single bit.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import math
import os
import statistics
import time
from collections import OrderedDict, defaultdict
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from . import _native as N
from .graphs import StepGraphs, observation_schedule, validate_graph_steps
from .planning import periodic_axes
from .models import Project
from .diagnostics import StateNorms
from .run_control import RunMonitor, check_count
from .solver import ENGINE_LOCK, DisplaySlice, Result, Stepper, pack_table as _pack, run_signature, run_summary


@dataclass
class BatchCase:
    """One project of a batch (torchfdtd/batch.py:31-42).

    `epsilon`, when given, is the case's relative permittivity on its grid (shape (Nx,Ny,Nz)
    or (Nx,Ny,Nz,3)), e.g. a design block already mapped to cells; it replaces the
    voxelized structures.
    """
    id: str
    project: Project
    parameters: dict = field(default_factory=dict)
    epsilon: np.ndarray | None = field(default=None, repr=False)

    def __post_init__(self):
        if not isinstance(self.id, str) or not self.id or any(c in self.id for c in '/\\\0'):
            raise ValueError('Case ids must be nonempty file-name-safe strings.')


@dataclass
class BatchItem:
    """Outcome of one case (torchfdtd/batch.py:44-67)."""
    id: str
    status: str
    parameters: dict
    summary: dict = field(default_factory=dict)
    metrics: dict = field(default_factory=dict)
    output: str | None = None
    error: str | None = None
    pid: int | None = None
    device: str | None = None
    started: float | None = None
    finished: float | None = None
    resumed: bool = False
    result: Result | None = field(default=None, repr=False)

    def load(self):
        if self.result is not None:
            return self.result
        if self.output is None:
            raise ValueError('This item kept no result.')
        return Result.load(self.output)

    def as_dict(self):
        return {k: v for k, v in self.__dict__.items() if k != 'result'}


@dataclass
class BatchReport:
    """All items of a batch and its plan (torchfdtd/batch.py:69-80)."""
    items: list
    seconds: float
    plan: dict

    @property
    def successful(self):
        return all(item.status == 'completed' for item in self.items)

    def raise_for_errors(self):
        failed = [item for item in self.items if item.status != 'completed']
        if failed:
            raise RuntimeError('; '.join(f'{item.id}: {item.status} {item.error or ""}'.strip() for item in failed))
        return self

    def as_dict(self):
        return dict(items=[item.as_dict() for item in self.items], seconds=self.seconds, plan=self.plan)


@dataclass(frozen=True)
class Launch:
    """Kernel configuration of a cohort: 'split' (E and H launches) or 'march' (one-pass tiles)."""
    kernel: str = 'march'
    tile_y: int = 8
    chunk_x: int = 6

    def __post_init__(self):
        if self.kernel not in ('split', 'march'):
            raise ValueError("launch kernel must be 'split' or 'march'")
        if self.tile_y not in (8, 16):
            raise ValueError('tile_y must be 8 or 16')
        if isinstance(self.chunk_x, bool) or not isinstance(self.chunk_x, int) or self.chunk_x < 0:
            raise ValueError('chunk_x must be a non-negative integer (0: the whole x axis)')

    def as_dict(self):
        return dict(kernel=self.kernel, tile_y=self.tile_y, chunk_x=self.chunk_x)


DEFAULT_CANDIDATES = (Launch('split'), Launch('march', 8, 2), Launch('march', 8, 3), Launch('march', 8, 4),
                      Launch('march', 8, 6), Launch('march', 8, 8), Launch('march', 8, 16), Launch('march', 8, 0),
                      Launch('march', 16, 3), Launch('march', 16, 6), Launch('march', 16, 16))


def _cases(cases):
    cases = [BatchCase(f'case-{i:04d}', c) if isinstance(c, Project) else c for i, c in enumerate(cases)]
    if any(not isinstance(c, BatchCase) for c in cases):
        raise TypeError('Supply BatchCase or Project instances.')
    if len({c.id for c in cases}) != len(cases):
        raise ValueError('Duplicate case ids.')
    return cases


def _admit(project):
    """The refusals of tensor_batch.py:77-86 that apply to this engine."""
    r = project.region
    refusal = r.unsupported()
    if refusal:
        raise ValueError(refusal)
    if r.backend == 'cpu':
        raise ValueError('Tensor batch cannot execute a CPU project. Set backend="cuda" or "auto" explicitly.')
    if r.complex_fields:
        raise ValueError('Tensor batch currently requires real fields.')
    if r.run_control.auto_shutoff:
        raise ValueError('Tensor batch requires fixed-duration runs. Set auto_shutoff=False.')


class Cohort:
    """Device state of a batch of runs and its batched step.

    Every run is a Stepper (the single-run device state) sharing one step counter; the
    cohort adds device tables of the runs' descriptors, one trace buffer for every point
    monitor of the batch, and the batched launches that advance all runs at once.
    """

    def __init__(self, projects, launch=Launch(), steps=None, epsilons=None):
        self.projects = list(projects)
        epsilons = [None]*len(self.projects) if epsilons is None else list(epsilons)
        if not self.projects:
            raise ValueError('A cohort needs at least one project.')
        first = self.projects[0].region
        self.steps = first.steps if steps is None else steps
        for p in self.projects:
            if p.region.steps != first.steps or p.region.precision != first.precision:
                raise ValueError('Every project of a cohort needs the same steps and precision.')
            if p.region.run_control.check_interval != first.run_control.check_interval:
                raise ValueError('Every project of a cohort needs the same run_control.check_interval.')
        self.lib = lib = N.library()
        device = torch.device('cuda', torch.cuda.current_device())
        self.counter = torch.zeros(1, device=device, dtype=torch.long)
        one_pass = launch.kernel == 'march'
        self.runs = [Stepper(p, steps=self.steps, kernel='hip', one_pass=False, tile_y=launch.tile_y,
                             chunk_x=launch.chunk_x, counter=self.counter, epsilon=eps, two_step=False)
                     for p, eps in zip(self.projects, epsilons)]
        if any(run.dispersive for run in self.runs):
            raise ValueError('Dispersive materials run through Simulation.run (one grid), not in cohorts.')
        if one_pass:
            # A cohort's many grids fill the GPU together: the one-pass kernels serve grids of any size.
            if not all(run.one_pass_supported for run in self.runs):
                raise ValueError('The one-pass kernels need 3D grids without periodic axes and at most 8 sources per family.')
            for run in self.runs:
                run.set_path('one_pass')
        if one_pass and len({run.grid.inverse_permittivity.shape[-1] for run in self.runs}) > 1:
            # One launch serves one permittivity layout: widen the per-cell grids to per component
            # (the same value three times, so every update is unchanged).
            for run in self.runs:
                g = run.grid
                if g.inverse_permittivity.shape[-1] == 1:
                    g.inverse_permittivity = g.inverse_permittivity.expand(*g.inverse_permittivity.shape[:-1], 3).contiguous()
                    run._describe()
        self.launch = launch
        self.one_pass = one_pass
        self.grids = [run.grid for run in self.runs]
        self.dtype = N.F64 if first.precision == 'float64' else N.F32
        count = len(self.runs)
        self.e_table, self.e_prefix, e_host = _pack(lambda n: lib.tfdtd_batch_table_bytes(n, self.dtype), lib.tfdtd_batch_pack,
                                                    [run.e_step for run in self.runs], 1, device)
        self.h_table, self.h_prefix, h_host = _pack(lambda n: lib.tfdtd_batch_table_bytes(n, self.dtype), lib.tfdtd_batch_pack,
                                                    [run.h_step for run in self.runs], 1, device)
        self.e_blocks, self.h_blocks = int(e_host[0, count]), int(h_host[0, count])
        # Point traces of the whole cohort: columns [offset[i], offset[i+1]) belong to run i.
        self.trace_offsets = np.cumsum([0]+[len(run.trace_points) for run in self.runs])
        self.traces = torch.zeros((self.steps, int(self.trace_offsets[-1])), device=device,
                                  dtype=self.grids[0].dtype)
        self.trace_table = self._cohort_traces([run.trace_table for run in self.runs])
        self.observe = self._observe(self.trace_table)
        self.planes = self._pack_planes([run.plane_table[k] for run in self.runs for k in range(len(run.planes))])
        if one_pass:
            self.march = []
            for k in range(2):
                table, prefix, host = _pack(lambda n: lib.tfdtd_batch_march_bytes(n, self.dtype), lib.tfdtd_batch_march_pack,
                                            [run.march[k] for run in self.runs], 3, device)
                self.march.append((table, prefix, (ctypes.c_int32*3)(*host[:, count].tolist())))
            self.trace_table_b = self._cohort_traces([run.trace_table_b for run in self.runs])
            self.observe_b = self._observe(self.trace_table_b)
            self.planes_b = self._pack_planes([run.plane_table_b[k] for run in self.runs for k in range(len(run.planes))])

    def _cohort_traces(self, tables):
        tables = [t for t in tables if t.numel()]
        if not tables:
            return torch.zeros((0, 2), dtype=torch.int64, device=self.counter.device)
        return torch.cat(tables)

    def _observe(self, table):
        o = N.Observe()
        o.dtype = self.dtype
        o.count = table.shape[0]
        o.table = table.data_ptr() if o.count else None
        o.traces = self.traces.data_ptr() if o.count else None
        o.steps = self.steps
        o.counter = self.counter.data_ptr()
        return o

    def _pack_planes(self, descriptors):
        """One batched DFT launch per (field dtype, accumulator precision) group of planes."""
        groups = defaultdict(list)
        for d in descriptors:
            groups[(d.dtype, d.spectral)].append(d)
        packed = []
        lib = self.lib
        for (dtype, spectral), members in groups.items():
            array = (N.Plane*len(members))(*members)
            nbytes = lib.tfdtd_batch_planes_bytes(len(members), dtype, spectral)
            host = (ctypes.c_ubyte*nbytes)()
            blocks, frequencies = ctypes.c_int32(), ctypes.c_int32()
            N.check(lib.tfdtd_batch_planes_pack(array, len(members), host, nbytes, ctypes.byref(blocks),
                                                ctypes.byref(frequencies)), 'tfdtd_batch_planes_pack')
            table = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(self.counter.device)
            packed.append((table, len(members), blocks.value, frequencies.value, dtype, spectral))
        return packed

    def reset(self, epsilons):
        """Start a fresh run of the same cohort with new permittivities (same shapes and layouts): every
        state returns to zero and the descriptor tables, planes and captured graphs stay valid."""
        epsilons = list(epsilons)
        if len(epsilons) != len(self.runs):
            raise ValueError('one permittivity per grid')
        for run, eps in zip(self.runs, epsilons):
            layout = run.grid.inverse_permittivity
            run.grid.set_permittivity(np.asarray(eps))
            if run.grid.inverse_permittivity is not layout:
                raise ValueError("a reset keeps each grid's permittivity layout (per-cell or per-component)")
        self.refresh_materials()
        for value in self.states():
            value.zero_()

    def refresh_materials(self):
        """Carry permittivity changes of any grid into its material index and into the cohort's
        one-pass tables (rewritten in place: captured graphs stay valid).  Called before every
        eager one-pass step and graph replay; a version check when nothing changed."""
        if not self.one_pass:
            return False
        if not any([run.refresh_materials() for run in self.runs]):
            return False
        for k, (table, prefix, totals) in enumerate(self.march):
            fresh, _, _ = _pack(lambda n: self.lib.tfdtd_batch_march_bytes(n, self.dtype), self.lib.tfdtd_batch_march_pack,
                                [run.march[k] for run in self.runs], 3, self.counter.device)
            table.copy_(fresh)
        return True

    def states(self):
        states = [self.counter, self.traces]
        for run in self.runs:
            states.extend(s for s in run.states() if s is not run.counter)
        return states

    def _monitors(self, planes, observe, stream):
        lib = self.lib
        for table, count, blocks, frequencies, dtype, spectral in planes:
            N.check(lib.tfdtd_batch_dft_update(table.data_ptr(), count, blocks, frequencies, dtype, spectral, stream),
                    'tfdtd_batch_dft_update')
        N.check(lib.tfdtd_observe_step(ctypes.byref(observe), stream), 'tfdtd_observe_step')

    def step(self):
        """One step of every grid with the split kernels (state in the grids' own buffers)."""
        stream = N.stream_handle()
        lib, count = self.lib, len(self.runs)
        N.check(lib.tfdtd_batch_half_step(self.e_table.data_ptr(), self.e_prefix.data_ptr(), count, self.e_blocks,
                                          self.dtype, N.FAMILY_E, stream), 'tfdtd_batch_half_step E')
        N.check(lib.tfdtd_batch_half_step(self.h_table.data_ptr(), self.h_prefix.data_ptr(), count, self.h_blocks,
                                          self.dtype, N.FAMILY_H, stream), 'tfdtd_batch_half_step H')
        self._monitors(self.planes, self.observe, stream)

    def step_pair(self):
        """Two steps with the one-pass kernels: grid buffers -> second set -> grid buffers."""
        self.refresh_materials()
        stream = N.stream_handle()
        lib, count, launch = self.lib, len(self.runs), self.launch
        eps_stride = self.grids[0].inverse_permittivity.shape[-1]
        for (table, prefix, totals), planes, observe in ((self.march[0], self.planes_b, self.observe_b),
                                                          (self.march[1], self.planes, self.observe)):
            N.check(lib.tfdtd_batch_march(table.data_ptr(), prefix.data_ptr(), count, self.dtype, launch.tile_y,
                                          eps_stride, totals, stream), 'tfdtd_batch_march')
            self._monitors(planes, observe, stream)

    def advance(self, count):
        if self.one_pass:
            for _ in range(count // 2):
                self.step_pair()
            if count % 2:
                self.step()
        else:
            for _ in range(count):
                self.step()

    def graphs(self, width=1):
        return StepGraphs(self.step, self.states(), self.steps, width, pair=self.step_pair if self.one_pass else None,
                          prepare=self.refresh_materials if self.one_pass else None)

    def traces_of(self, i, completed):
        return self.traces[:completed, int(self.trace_offsets[i]):int(self.trace_offsets[i+1])].cpu().numpy().copy()

    def digest(self):
        """SHA-256 of every state array of the cohort (fields, CPML memory, planes, traces)."""
        h = hashlib.sha256()
        torch.cuda.synchronize()
        for value in [self.traces, *(s for run in self.runs for s in (run.grid.E, run.grid.H, *run.grid.memory_states))]:
            h.update(value.detach().contiguous().cpu().numpy().tobytes())
        return h.hexdigest()


def resolve_launch(projects, launch):
    """Launch for a cohort: an explicit Launch, a tuning report, or 'auto' (one-pass where it applies)."""
    if isinstance(launch, TensorBatchTuning):
        return launch.launch
    if isinstance(launch, Launch):
        return launch
    if launch not in (None, 'auto'):
        raise ValueError("launch must be a Launch, a TensorBatchTuning or 'auto'")
    one_pass = all(p.region.dimension == '3d' and not periodic_axes(p.region) and min(p.region.shape) >= 2
                   for p in projects)
    return Launch('march') if one_pass else Launch('split')


def run_tensor_batch(cases, *, objective=None, output_dir=None, keep_results=True, device=0, memory_fraction=.6,
                     cuda_graph=True, cancel=None, progress=None, cohort_size=None, cuda_graph_steps=1, launch='auto'):
    """Run independent projects in batched HIP cohorts of 1--64 cases (tensor_batch.py:44-113).

    Unlike the reference, the projects of a cohort may differ in shape, CPML layout,
    sources and monitors; they must share steps, precision and check interval.  `launch` selects the
    kernels ('auto', a Launch, or the report of tune_tensor_batch).
    """
    validate_graph_steps(cuda_graph_steps, cuda_graph)
    cases = _cases(cases)
    if not cases:
        raise ValueError('Tensor batch requires at least one case.')
    if cohort_size is None:
        cohort_size = len(cases)
    if isinstance(cohort_size, bool) or not isinstance(cohort_size, int) or not 1 <= cohort_size <= 64:
        raise ValueError('cohort_size must be an integer from 1 to 64. Split larger ensembles explicitly.')
    root = Path(output_dir) if output_dir is not None else None
    if root and ((root/'tensor-batch.json').exists() or any((root/(c.id+'.npz')).exists() for c in cases)):
        raise FileExistsError('Tensor batch output exists. Use a new directory.')
    if not 0 < memory_fraction <= .9:
        raise ValueError('memory_fraction must be positive and at most .9.')
    projects = [Project.model_validate(c.project.model_dump()) for c in cases]
    for p in projects:
        _admit(p)
    if not torch.cuda.is_available():
        raise RuntimeError('Tensor batch requires a ROCm GPU.')
    if not isinstance(device, int) or not 0 <= device < torch.cuda.device_count():
        raise ValueError('Invalid GPU device.')
    with ENGINE_LOCK, torch.cuda.device(device), torch.no_grad():
        start = time.perf_counter()
        items, plans = [], []
        for offset in range(0, len(cases), cohort_size):
            if cancel is not None and cancel.is_set() and offset:
                items.extend(BatchItem(c.id, 'cancelled', c.parameters) for c in cases[offset:])
                break
            members = projects[offset:offset+cohort_size]
            report = _run(cases[offset:offset+cohort_size], members, resolve_launch(members, launch), objective, root,
                          keep_results, memory_fraction, cuda_graph, cancel, progress, cuda_graph_steps, offset, len(cases))
            items.extend(report.items)
            plans.append(report.plan)
        plan = plans[0] if len(plans) == 1 else dict(execution='batched_hip_cohorts', cohorts=plans,
                                                      batch_size=len(cases),
                                                      loop_seconds=sum(p['loop_seconds'] for p in plans),
                                                      setup_seconds=sum(p['setup_seconds'] for p in plans))
        plan = dict(plan, total_cases=len(cases), cohort_size=cohort_size, cohorts_completed=len(plans))
        report = BatchReport(items, time.perf_counter()-start, plan)
        if root:
            root.mkdir(parents=True, exist_ok=True)
            (root/'tensor-batch.json').write_text(json.dumps(report.as_dict(), indent=2, default=str))
        return report


def _estimate_bytes(project):
    """Device bytes of one run: fields, inverse permittivity, CPML memory and the one-pass second set."""
    r = project.region
    cells = math.prod(r.shape)
    real = 8 if r.precision == 'float64' else 4
    eps = 3 if r.material_sampling == 'yee' else 1
    return cells*real*(6+eps+6+6)


def _fixed_steps(project):
    """The project with auto-shutoff off: a cohort runs every step, and each grid's rule keeps
    judging divergence (and recording its history) to the end."""
    fixed = project.model_copy(deep=True)
    fixed.region.run_control.auto_shutoff = False
    return fixed


def _run(cases, projects, launch, objective, root, keep_results, memory_fraction, cuda_graph, cancel, progress,
         cuda_graph_steps, offset, total_cases):
    started, wall_started = time.perf_counter(), time.time()
    estimate = sum(_estimate_bytes(p) for p in projects)
    free, _ = torch.cuda.mem_get_info()
    if estimate > free*memory_fraction:
        raise ValueError('Tensor cohort exceeds its GPU memory allowance. Split it into smaller cohorts.')
    cohort = Cohort(projects, launch, epsilons=[c.epsilon for c in cases])
    first = projects[0].region
    checked = [p.region.run_control.divergence_check for p in projects]
    # Every grid of the cohort is measured and judged on the device in one set of launches per
    # diagnostic barrier; the cohort runs every step (no auto-shutoff, as tensor_batch.py:247-262).
    monitor = (RunMonitor(cohort.grids, [_fixed_steps(p) for p in projects], check_count(projects[0]),
                          labels=[f'Case {c.id}, ' for c in cases]) if any(checked) else None)
    graph = cohort.graphs(cuda_graph_steps) if cuda_graph else None
    views = [DisplaySlice(p) for p in projects]
    frames, frame_steps = [[] for _ in cases], [[] for _ in cases]
    # Host observations only where a case has one due (tensor_batch.py:247-262).
    events, intervals = defaultdict(list), []
    for i, p in enumerate(projects):
        r = p.region
        interval = max(r.snapshot_interval, math.ceil(r.steps/100))
        intervals.append(interval)
        due = set(range(interval, r.steps+1, interval)) | {r.steps}
        if checked[i]:
            due.update(range(r.run_control.check_interval, r.steps+1, r.run_control.check_interval))
        for q in due:
            events[q].append(i)
    barriers = set(events)
    if progress:
        barriers.update(range(max(1, first.steps//100), first.steps+1, max(1, first.steps//100)))
    width = graph.width if graph is not None else 1
    torch.cuda.synchronize()
    setup_seconds = time.perf_counter()-started
    loop_start = time.perf_counter()
    completed, cancelled, replays, eager_steps = 0, False, 0, 0
    diagnostic_seconds = 0.
    check = first.run_control.check_interval
    for target, advance in observation_schedule(first.steps, width, barriers):
        if cancel is not None and cancel.is_set():
            cancelled = True
            break
        if graph is not None:
            graph.replay(advance)
            replays += 1
        else:
            cohort.advance(advance)
            eager_steps += advance
        completed = target
        due = events.get(completed, ())
        if monitor is not None and (completed % check == 0 or completed == first.steps):
            tick = time.perf_counter()
            monitor.check(completed)    # every grid of the cohort, enqueued; verdicts read as they arrive
            monitor.poll()
            diagnostic_seconds += time.perf_counter()-tick
        for i in due:
            r = projects[i].region
            if completed % intervals[i] == 0 or completed == r.steps:
                frames[i].append(views[i].device_frame(cohort.grids[i]))
                frame_steps[i].append(completed)
        if progress and (completed % max(1, first.steps//100) == 0 or completed == first.steps):
            progress(dict(step=completed, total=first.steps, cases=len(cases), cohort_offset=offset,
                          total_cases=total_cases, elapsed=time.perf_counter()-loop_start))
    torch.cuda.synchronize()
    seconds = time.perf_counter()-loop_start
    if monitor is not None:
        monitor.finish()
    for i, case in enumerate(cases):
        frames[i] = np.array([f.cpu().numpy() for f in frames[i]])
        if not np.isfinite(frames[i]).all():
            raise FloatingPointError(f'Case {case.id}: non-finite displayed field.')
    kernel_steps = dict(graph.kernel_steps) if graph is not None else dict(two_step=0, one_pass=0, split=eager_steps)
    if root:
        root.mkdir(parents=True, exist_ok=True)
    items = []
    for i, (case, p, run) in enumerate(zip(cases, projects, cohort.runs)):
        g = run.grid
        e, h = g.E.cpu().numpy().copy(), g.H.cpu().numpy().copy()
        if not np.isfinite(e).all() or not np.isfinite(h).all():
            raise FloatingPointError(f'Case {case.id}: non-finite final fields.')
        termination = 'cancelled' if cancelled else 'max_steps'
        summary = run_summary(p, run, e, completed, termination, seconds, setup_seconds,
                              cuda_graph=graph is not None, cuda_graph_steps=width, cuda_graph_replays=replays,
                              cuda_kernel='batched_'+launch.kernel, launch=launch.as_dict(), batch_size=len(cases),
                              batch_index=i, timing_scope='whole_cohort', kernel_steps=kernel_steps,
                              diagnostics=monitor.history(i) if monitor is not None and checked[i] else [],
                              diagnostic_seconds=diagnostic_seconds, diagnostic_scope='whole_cohort',
                              diagnostic_backend='hip' if checked[i] else None,
                              slice_index=views[i].index)
        result = Result(p, summary, frames[i], np.array(frame_steps[i]), views[i].epsilon(run.eps),
                        cohort.traces_of(i, completed), np.arange(1, completed+1)*g.time_step, e, h,
                        [dict(m.result(), run_signature=run_signature(p, completed)) for m in run.planes])
        item = BatchItem(case.id, 'cancelled' if cancelled else 'completed', case.parameters, summary=summary,
                         pid=os.getpid(), device=str(g.device), started=wall_started)
        if root:
            path = root/(case.id+'.npz')
            result.save(path)
            item.output = str(path)
        if not cancelled and objective is not None:
            try:
                value = objective(result)
                if not isinstance(value, dict):
                    value = {'objective': value}
                metrics = {str(k): float(v) for k, v in value.items()}
                if not metrics or not all(math.isfinite(v) for v in metrics.values()):
                    raise ValueError('Objective must return finite nonempty scalar metrics.')
                item.metrics = metrics
            except Exception as exc:  # an invalid objective marks only its item failed
                item.status, item.error = 'failed', f'{type(exc).__name__}: {exc}'
        if keep_results:
            item.result = result
        item.finished = time.time()
        items.append(item)
    plan = dict(execution='batched_hip', batch_size=len(cases), device=str(cohort.grids[0].device),
                estimated_memory_mb=estimate/2**20, memory_fraction=memory_fraction, loop_seconds=seconds,
                setup_seconds=setup_seconds, graph=cuda_graph, cuda_graph_steps=width, cuda_graph_replays=replays,
                launch=launch.as_dict(), cells=int(sum(math.prod(p.region.shape) for p in projects)))
    return BatchReport(items, time.perf_counter()-started, plan)


# ------------------------------------------------------------------ grouped batches (grouped_batch.py)

def plan_grouped_batch(cases, *, cohort_size=64, memory_limit_bytes=None):
    """Cohorts of at most `cohort_size` cases and `memory_limit_bytes` estimated device bytes.

    Any mix of shapes may share a cohort; cases are grouped by (steps, precision, diagnostic
    check interval), the only common properties a cohort needs, largest first so cohorts fill
    evenly.
    """
    cases = _cases(cases)
    if isinstance(cohort_size, bool) or not isinstance(cohort_size, int) or not 1 <= cohort_size <= 64:
        raise ValueError('cohort_size must be an integer from 1 to 64.')
    groups = defaultdict(list)
    for index, case in enumerate(cases):
        r = case.project.region
        groups[(r.steps, r.precision, r.run_control.check_interval)].append(index)
    cohorts = []
    for key, members in groups.items():
        members.sort(key=lambda i: -_estimate_bytes(cases[i].project))
        current, used = [], 0
        for i in members:
            size = _estimate_bytes(cases[i].project)
            if current and (len(current) == cohort_size or (memory_limit_bytes and used+size > memory_limit_bytes)):
                cohorts.append(current)
                current, used = [], 0
            current.append(i)
            used += size
        if current:
            cohorts.append(current)
    return dict(cohorts=cohorts, cohort_count=len(cohorts), cases=len(cases),
                estimated_bytes=[sum(_estimate_bytes(cases[i].project) for i in c) for c in cohorts])


def run_grouped_batch(cases, *, cohort_size=64, objective=None, output_dir=None, keep_results=True, device=0,
                      memory_fraction=.6, memory_limit_bytes=None, cuda_graph=True, cancel=None, progress=None,
                      cuda_graph_steps=1, launch='auto'):
    """Run every case in planned cohorts and restore the input order (grouped_batch.py:90-146)."""
    cases = _cases(cases)
    plan = plan_grouped_batch(cases, cohort_size=cohort_size, memory_limit_bytes=memory_limit_bytes)
    started = time.perf_counter()
    items, plans = [None]*len(cases), []
    for cohort in plan['cohorts']:
        if cancel is not None and cancel.is_set():
            for i in cohort:
                items[i] = BatchItem(cases[i].id, 'cancelled', cases[i].parameters)
            continue
        report = run_tensor_batch([cases[i] for i in cohort], objective=objective, output_dir=output_dir,
                                  keep_results=keep_results, device=device, memory_fraction=memory_fraction,
                                  cuda_graph=cuda_graph, cancel=cancel, progress=progress, cuda_graph_steps=cuda_graph_steps,
                                  launch=launch)
        for i, item in zip(cohort, report.items):
            items[i] = item
        plans.append(report.plan)
    return BatchReport(items, time.perf_counter()-started, dict(execution='grouped_hip_cohorts', plan=plan, cohorts=plans))


# ------------------------------------------------------------------ autotuner (tuning.py, reversible_cuda_tuning.py)

@dataclass
class TensorBatchTuning:
    """The measured launch selection (tuning.py:26-45)."""
    launch: Launch
    timings: dict
    rejected: dict
    digest: str
    fingerprint: str
    trial_steps: int
    seconds: float

    def as_dict(self):
        return dict(launch=self.launch.as_dict(), timings=self.timings, rejected=self.rejected, digest=self.digest,
                    fingerprint=self.fingerprint, trial_steps=self.trial_steps, seconds=self.seconds)

    def save(self, path):
        Path(path).write_text(json.dumps(self.as_dict(), indent=2))


_TUNING_CACHE: OrderedDict = OrderedDict()
_TUNING_CACHE_SIZE = 32


def _fingerprint(projects, trial_steps, candidates):
    payload = json.dumps([[p.model_dump(mode='json') for p in projects], trial_steps, [c.as_dict() for c in candidates],
                          torch.cuda.get_device_name()], sort_keys=True, default=str)
    return hashlib.sha256(payload.encode()).hexdigest()


def tune_tensor_batch(cases, *, candidates=DEFAULT_CANDIDATES, repeats=3, trial_steps=200, device=0, cache=True):
    """Time every launch candidate on the actual cohort and keep the fastest (tuning.py:63-160).

    Each candidate builds the cohort, runs `trial_steps` steps once untimed (warm-up and
    graph-free compile), then `repeats` timed trials from a zeroed state, alternating the
    candidate order between repeats.  Every candidate must reproduce the first one's state
    bit for bit (the physics may not depend on the launch); a candidate that does not is an
    error.  Candidates the cohort cannot use (the one-pass kernels on 2D or periodic grids)
    are recorded in `rejected`.  Results are cached in-process by SHA-256 of the projects,
    trial length, candidates and device (reversible_cuda_tuning.py:29-60).
    """
    started = time.perf_counter()
    cases = _cases(cases)
    if not cases:
        raise ValueError('Tuning requires at least one case.')
    if isinstance(repeats, bool) or not isinstance(repeats, int) or repeats < 1:
        raise ValueError('repeats must be a positive integer.')
    candidates = [c if isinstance(c, Launch) else Launch(*c) for c in candidates]
    if not candidates:
        raise ValueError('At least one candidate is required.')
    projects = [Project.model_validate(c.project.model_dump()) for c in cases]
    for p in projects:
        _admit(p)
    trial_steps = min(int(trial_steps), projects[0].region.steps)
    if not torch.cuda.is_available():
        raise RuntimeError('Tuning requires a ROCm GPU.')
    fingerprint = _fingerprint(projects, trial_steps, candidates)
    if cache and fingerprint in _TUNING_CACHE:
        _TUNING_CACHE.move_to_end(fingerprint)
        return _TUNING_CACHE[fingerprint]
    samples, rejected, cohorts = defaultdict(list), {}, {}
    digest = None
    with ENGINE_LOCK, torch.cuda.device(device), torch.no_grad():
        for c in candidates:
            try:
                cohorts[c] = Cohort(projects, c, steps=trial_steps, epsilons=[case.epsilon for case in cases])
            except ValueError as exc:
                rejected[json.dumps(c.as_dict())] = str(exc)
        if not cohorts:
            raise ValueError('No launch candidate applies to this cohort: ' + '; '.join(rejected.values()))
        order = list(cohorts)
        for c in order:  # warm-up, and the bitwise check of every candidate
            cohort = cohorts[c]
            cohort.advance(trial_steps)
            d = cohort.digest()
            if digest is None:
                digest = d
            elif d != digest:
                raise RuntimeError(f'Launch {c.as_dict()} changes the results; every candidate must be bitwise identical.')
        for r in range(repeats):
            for c in (order if r % 2 == 0 else order[::-1]):
                cohort = cohorts[c]
                for value in cohort.states():
                    value.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                cohort.advance(trial_steps)
                torch.cuda.synchronize()
                samples[c].append(time.perf_counter()-t0)
        del cohorts
        torch.cuda.empty_cache()
    timings = {json.dumps(c.as_dict()): dict(median_seconds=statistics.median(v), samples=v) for c, v in samples.items()}
    best = min(samples, key=lambda c: statistics.median(samples[c]))
    report = TensorBatchTuning(best, timings, rejected, digest, fingerprint, trial_steps, time.perf_counter()-started)
    if cache:
        _TUNING_CACHE[fingerprint] = report
        while len(_TUNING_CACHE) > _TUNING_CACHE_SIZE:
            _TUNING_CACHE.popitem(last=False)
    return report
