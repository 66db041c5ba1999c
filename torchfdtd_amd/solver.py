"""``Simulation(project).run()`` on the gfx950 HIP engine (torchfdtd/solver.py:642-908).

The step is the reference's closure (solver.py:785-797): E update, E sources, H
update, H sources, plane DFTs, point traces, counter += 1.  With the HIP kernel the
sources are folded into the update kernels and one C call issues the whole step;
HIP graphs replay 1 or `cuda_graph_steps` steps between host barriers (frames,
run-control diagnostics, cancellation), exactly at the reference's barrier steps.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import math
import threading
import time
from dataclasses import dataclass, field
from functools import cached_property
from pathlib import Path

import numpy as np
import torch

from . import _native as N
from .graphs import StepGraphs, observation_schedule, validate_graph_steps
from .grid import (MaterialIndex, YeeGrid, configure_cuda_kernel, half_step_descriptor, march_descriptor, march_supported,
                   resolve_cuda_kernel)
from .models import Project
from .monitors import FrequencyPlane, apodization_window, point_spectrum
from .plan import resolve_plan
from .planning import field_axes, index_at, source_terms, voxelize
from .run_control import RunMonitor, check_count

C0 = 299792458.0
ENGINE_LOCK = threading.Lock()
# Default step path by grid size (measured on MI355X, per step, tiles at their best shapes): the split
# step beats the one-pass tiles by 11-48% on 0.8-7.1 M-cell grids, by 9% at 256^3 and ties them at
# 320^3 (32.8 M cells); the two-step pass ties the split step at 256^3 and beats it by 21% at 320^3.
# So float32 grids without plane monitors or traces take the two-step pass from TWO_STEP_MIN_CELLS,
# other grids the one-pass step from ONE_PASS_MIN_CELLS, and smaller grids the split step.
# tuning.tune_launch measures every path on the grid itself.
TWO_STEP_MIN_CELLS = 20_000_000
ONE_PASS_MIN_CELLS = 48_000_000
# The two-step pass serves plane monitors and traces by writing step n+1 in the cells they read to a
# snapshot of at most SNAPSHOT_BOXES boxes; grids whose observers need more than an eighth of the
# cells keep the one-step paths (the snapshot would cost more than the pass saves).
SNAPSHOT_MAX_FRACTION = 8
SNAPSHOT_BOXES = 4


def _cells(box):
    lo, hi = box
    return math.prod(b-a for a, b in zip(lo, hi))


def _union(a, b):
    return tuple(map(min, a[0], b[0])), tuple(map(max, a[1], b[1]))


def merge_boxes(boxes, most=SNAPSHOT_BOXES):
    """Disjoint boxes (lo, hi cell triples, half-open) covering `boxes`: boxes that overlap or touch are
    replaced by their bounding box, then the pair whose bounding box adds the fewest cells, until at
    most `most` remain; sorted by lo."""
    out = [(tuple(lo), tuple(hi)) for lo, hi in boxes]
    merged = True
    while merged:
        merged = False
        for i in range(len(out)):
            for j in range(i+1, len(out)):
                a, b = out[i], out[j]
                if all(a[0][k] <= b[1][k] and b[0][k] <= a[1][k] for k in range(3)):
                    out[i] = _union(a, b)
                    del out[j]
                    merged = True
                    break
            if merged:
                break
    while len(out) > most:
        pairs = [(i, j) for i in range(len(out)) for j in range(i+1, len(out))]
        i, j = min(pairs, key=lambda ij: _cells(_union(out[ij[0]], out[ij[1]])) - _cells(out[ij[0]]) - _cells(out[ij[1]]))
        out[i] = _union(out[i], out[j])
        del out[j]
        out = merge_boxes(out, most=len(out))
    return sorted(out)


def default_step_path(cells, two_step_eligible, one_pass_supported):
    """The step path a grid of `cells` takes by default: 'two_step' (float32, no plane monitors or
    traces: `two_step_eligible`), 'one_pass' or 'split'."""
    if not one_pass_supported:
        return 'split'
    if two_step_eligible and cells >= TWO_STEP_MIN_CELLS:
        return 'two_step'
    return 'one_pass' if cells >= ONE_PASS_MIN_CELLS else 'split'


@dataclass
class Result:
    """The reference's Result record (solver.py:511-639) and its NPZ layout."""
    project: Project
    summary: dict
    frames: np.ndarray
    frame_steps: np.ndarray
    epsilon: np.ndarray
    signals: np.ndarray
    times: np.ndarray
    electric: np.ndarray
    magnetic: np.ndarray
    frequency_fields: list = field(default_factory=list)

    @property
    def point_monitors(self):
        return [self.project.resolved_monitor(m) for m in self.project.monitors if m.enabled and m.kind == 'point']

    @cached_property
    def spectra(self):
        # H is sampled after its half-step update, E after its full-step update.
        dt = self.times[1]-self.times[0] if len(self.times) > 1 else 0
        return [point_spectrum((self.times+(dt/2 if m.component.startswith('H') else 0))[::m.time_downsample],
                               self.signals[::m.time_downsample, k], m.spectrum)
                for k, m in enumerate(self.point_monitors)]

    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays, metadata = {}, []
        for k, (m, spec) in enumerate(zip(self.point_monitors, self.spectra)):
            arrays[f'monitor_{k}_frequency_hz'] = spec['frequency_hz']
            arrays[f'monitor_{k}_spectrum'] = spec['value']
            metadata.append({'id': m.id, 'name': m.name, 'component': m.component,
                             'settings': m.spectrum.model_dump(), 'units': spec['units'], 'transform': spec['transform']})
        planes = []
        for k, plane in enumerate(self.frequency_fields):
            planes.append({key: v for key, v in plane.items() if not isinstance(v, np.ndarray)})
            for key, v in plane.items():
                if isinstance(v, np.ndarray):
                    arrays[f'field_monitor_{k}_{key}'] = v
        np.savez_compressed(path, project=self.project.model_dump_json(), summary=json.dumps(self.summary),
                            field_monitors=json.dumps(planes, default=_json_default),
                            **{f'mesh_{a}_um': nodes for a, nodes in zip('xyz', self.project.region.mesh_nodes)},
                            frames=self.frames, frame_steps=self.frame_steps, epsilon=self.epsilon,
                            signals=self.signals, times=self.times, E=self.electric, H=self.magnetic,
                            monitor_spectra=json.dumps(metadata), **arrays)

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as data:
            planes = json.loads(str(data['field_monitors'])) if 'field_monitors' in data else []
            for k, plane in enumerate(planes):
                prefix = f'field_monitor_{k}_'
                plane.update({key[len(prefix):]: data[key].copy() for key in data.files if key.startswith(prefix)})
            return cls(Project.model_validate_json(str(data['project'])), json.loads(str(data['summary'])),
                       *(data[key].copy() for key in ('frames', 'frame_steps', 'epsilon', 'signals', 'times', 'E', 'H')), planes)

    def field_monitor(self, name_or_id):
        matches = [m for m in self.frequency_fields if name_or_id in (m['id'], m['name'])]
        if len(matches) != 1:
            raise ValueError('Expected one enabled frequency plane with this name or id.')
        return matches[0]


def _json_default(value):
    if isinstance(value, (np.integer,)):
        return int(value)
    if isinstance(value, (np.floating,)):
        return float(value)
    if isinstance(value, tuple):
        return list(value)
    raise TypeError(type(value))


def _check_epsilon(eps, region):
    if isinstance(eps, torch.Tensor):
        eps = eps.detach().cpu().numpy()
    eps = np.asarray(eps)
    shape = tuple(region.shape)
    if eps.shape not in (shape, shape+(3,)):
        raise ValueError(f'epsilon must have shape {shape} or {shape+(3,)}, not {eps.shape}')
    if not np.isfinite(eps).all() or (eps <= 0).any():
        raise ValueError('epsilon must be finite and positive')
    return eps


def pack_table(count_bytes, pack, descriptors, prefix_rows, device):
    """Host-pack descriptors with the library, then copy the table and block prefix to the device."""
    count = len(descriptors)
    array = (type(descriptors[0])*count)(*descriptors)
    nbytes = count_bytes(count)
    if nbytes == 0:
        raise ValueError('unsupported descriptor dtype')
    host = (ctypes.c_ubyte*nbytes)()
    prefix = (ctypes.c_int32*(prefix_rows*(count+1)))()
    N.check(pack(array, count, host, nbytes, prefix), 'batch pack')
    table = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
    prefix_host = np.array(prefix, dtype=np.int32).reshape(prefix_rows, count+1)
    return table, torch.as_tensor(prefix_host, device=device), prefix_host


class DispersiveMaterial:
    """ADE state of the samples one Drude/Lorentz material owns (materials.py:29-87, 'ade' PML mode:
    every owned sample, PML cells included): P and Q per pole, the saved E, and the tfdtd_ade
    descriptor whose coefficients a, d, k are formed in double as the reference forms them."""

    def __init__(self, grid, material, indices, per_cell):
        poles = material.oscillators
        if not 1 <= len(poles) <= N.ADE_MAX_POLES:
            raise ValueError(f'{material.name}: 1 to {N.ADE_MAX_POLES} poles per material')
        device, real = grid.E.device, grid.dtype
        self.material, self.poles = material, poles
        self.index = torch.as_tensor(np.asarray(indices, dtype=np.int64), device=device)
        samples = len(indices)*(3 if per_cell else 1)
        self.P = torch.zeros((len(poles), samples), device=device, dtype=real)
        self.Q = torch.zeros_like(self.P)
        self.old = torch.zeros(samples, device=device, dtype=real)
        grid.memory_states.extend((self.P, self.Q))
        dt = grid.time_step
        d = N.Ade()
        d.dtype = N.F64 if real == torch.float64 else N.F32
        d.poles, d.samples, d.per_cell = len(poles), samples, int(per_cell)
        d.index, d.field, d.old = self.index.data_ptr(), grid.E.data_ptr(), self.old.data_ptr()
        d.P, d.Q = self.P.data_ptr(), self.Q.data_ptr()
        for j, (w0, strength, gamma) in enumerate(poles):
            denominator = 1+.5*gamma*dt+.25*(w0*dt)**2
            d.a[j], d.d[j], d.k[j] = .5*(w0*dt)**2, denominator, strength*dt*dt/(4*denominator)
        d.eps_inf = material.epsilon_inf
        self.descriptor = d
        self.dt = dt

    def energy(self):
        """FP64 oscillator energy (run_control.py:144-151): sum_j (Q_j/(dt sqrt s_j))^2 + (P_j w0_j/sqrt s_j)^2."""
        total = torch.zeros((), device=self.P.device, dtype=torch.float64)
        for j, (w0, strength, _) in enumerate(self.poles):
            q = self.Q[j].to(torch.float64)/(self.dt*math.sqrt(strength))
            p = self.P[j].to(torch.float64)*(w0/math.sqrt(strength))
            total = total+(q*q).sum()+(p*p).sum()
        return total


def dispersive_materials(project, grid, ownership):
    """One DispersiveMaterial per Drude/Lorentz material that owns samples."""
    if ownership is None:
        return []
    states = []
    for i, m in enumerate(project.materials):
        if m.oscillators and np.any(ownership == i):
            states.append(DispersiveMaterial(grid, m, np.flatnonzero(ownership.reshape(-1) == i), ownership.ndim == 3))
    if states and project.region.pml_dispersion != 'ade':
        raise ValueError(f'pml_dispersion="{project.region.pml_dispersion}" is not run by this engine; '
                         'dispersive materials run with pml_dispersion="ade"')
    return states


class Stepper:
    """Device state of one resident run and its step (the reference's closure, solver.py:776-797)."""

    def __init__(self, project, steps=None, kernel=None, one_pass=None, tile_y=8, chunk_x=6, counter=None, epsilon=None,
                 two_step=None, chunk_x2=20, waves2=12, launch=None):
        p, r = project, project.region
        self.project = p
        self.steps = r.steps if steps is None else steps
        self.kernel = kernel or resolve_cuda_kernel(r)
        # An explicit relative permittivity (a design already mapped to the grid) replaces the voxelized structures.
        ownership = None
        if epsilon is None:
            self.eps, ownership, _ = voxelize(p, with_ownership=True)
        else:
            self.eps = _check_epsilon(epsilon, r)
        self.grid = g = YeeGrid(r)
        g.set_permittivity(self.eps)
        self.lib = N.library()
        # Drude/Lorentz materials: their ADE wraps the E update of the split step.
        self.dispersive = dispersive_materials(p, g, ownership)
        if self.dispersive and self.kernel != 'hip':
            raise ValueError('dispersive materials run on the HIP kernel (cuda_kernel "hip" or "auto")')
        device, real = g.E.device, g.dtype
        # A batch shares one counter among its grids.
        self.counter = torch.zeros(1, device=device, dtype=torch.long) if counter is None else counter
        terms = source_terms(p, self.steps)
        self.sources = {'E': [], 'H': []}
        for t in terms:
            self.sources[t.family].append((t.comp, t.box, torch.as_tensor(t.values, device=device)))
        self.point_monitors = [p.resolved_monitor(m) for m in p.monitors if m.enabled and m.kind == 'point']
        self.trace_points = []
        for m in self.point_monitors:
            loc = index_at(m.center, r, m.component)
            flat = ((loc[0]*r.shape[1]+loc[1])*r.shape[2]+loc[2])*3+'xyz'.index(m.component[1].lower())
            self.trace_points.append((m.component[0] == 'H', flat))
        self.traces = torch.zeros((self.steps, len(self.trace_points)), device=device, dtype=real)
        self.planes = [FrequencyPlane(g, p.resolved_monitor(m)) for m in p.monitors if m.enabled and m.kind == 'field']
        if self.kernel == 'torch':
            configure_cuda_kernel(g, 'torch')
        self.one_pass_supported = supported = (self.kernel == 'hip' and march_supported(g, self.sources)
                                               and not self.dispersive)
        # Default path by size (the one-pass tiles of a small grid are mostly CPML boundary tiles and
        # too few to fill the GPU): the two-step pass for large grids that allow it, the one-pass step
        # for larger grids that do not, the split step otherwise.
        eligible = self._two_step_eligible()
        large = default_step_path(math.prod(r.shape), eligible, supported) != 'split'
        self.one_pass = supported and (large or bool(two_step)) if one_pass is None else bool(one_pass)
        if self.one_pass and not supported:
            raise ValueError('the one-pass kernel needs a 3D grid without periodic axes, dispersive materials or '
                             'more than 8 sources per family')
        self.tile_y, self.chunk_x = tile_y, chunk_x
        self.chunk_x2, self.waves2 = chunk_x2, waves2      # two-step pass: planes per chunk, wavefronts per tile
        # Two steps per pass (float32 one-pass grids; plane monitors and traces sample step n+1 from a
        # snapshot of the planes they read).
        self.two_step = self._two_step_allowed() if two_step is None else bool(two_step)
        if self.two_step and not self._two_step_allowed():
            raise ValueError('the two-step kernel needs a float32 one-pass grid whose plane monitors and point traces '
                             f'read at most 1/{SNAPSHOT_MAX_FRACTION} of the cells in at most {SNAPSHOT_BOXES} boxes')
        # One-byte material index of 1/eps for the one-pass step (grids with few distinct values).
        self.materials = MaterialIndex(g) if self.one_pass else None
        self._describe()
        # launch='auto': measure the launch shapes of this grid's step path and keep the fastest
        # (tuning.tune_launch; the record lands in self.launch_record).
        self.launch_record = None
        if launch == 'auto':
            from .tuning import tune_launch
            self.launch_record = tune_launch(self)
        elif launch is not None:
            raise ValueError("launch is None (default shapes) or 'auto' (measured)")

    def _observer_boxes(self):
        """The cell box each trace and each plane monitor reads (traces first, in order)."""
        g = self.grid
        shape = (g.Nx, g.Ny, g.Nz)
        boxes = []
        for _, flat in self.trace_points:
            cell = np.unravel_index(flat // 3, shape)
            boxes.append((tuple(int(c) for c in cell), tuple(int(c)+1 for c in cell)))
        for plane in self.planes:
            cells = torch.cat([(indices // 3).reshape(-1) for indices, _ in plane.maps])
            coords = [cells // (g.Ny*g.Nz), (cells // g.Nz) % g.Ny, cells % g.Nz]
            boxes.append((tuple(int(c.min()) for c in coords), tuple(int(c.max())+1 for c in coords)))
        return boxes

    def observed_boxes(self):
        """The boxes of cells a step n+1 snapshot covers: every cell the observers read, merged into at
        most SNAPSHOT_BOXES disjoint boxes ([] without observers)."""
        return merge_boxes(self._observer_boxes())

    def _two_step_eligible(self):
        if self.grid.dtype != torch.float32:
            return False
        g = self.grid
        cells = sum(_cells(b) for b in self.observed_boxes())
        return cells <= max(2*g.Ny*g.Nz, g.Nx*g.Ny*g.Nz // SNAPSHOT_MAX_FRACTION)

    def _snapshot_index(self, indices, boxes):
        """Element indices of the fields -> the same elements in the snapshot (box after box)."""
        g = self.grid
        cells, comp = indices // 3, indices % 3
        x, y, z = cells // (g.Ny*g.Nz), (cells // g.Nz) % g.Ny, cells % g.Nz
        out = torch.full_like(indices, -1)
        base = 0
        for lo, hi in boxes:
            inside = ((x >= lo[0]) & (x < hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2]))
            local = ((x-lo[0])*(hi[1]-lo[1]) + (y-lo[1]))*(hi[2]-lo[2]) + (z-lo[2])
            out = torch.where(inside, (base+local)*3 + comp, out)
            base += _cells((lo, hi))
        if bool((out < 0).any()):
            raise RuntimeError('an observer reads a cell outside the step n+1 snapshot')
        return out.contiguous()

    def _two_step_allowed(self):
        return self.one_pass and self._two_step_eligible()

    def paths(self):
        """The step paths this grid can take: 'split' always; 'one_pass' on one-pass grids; 'two_step'
        on float32 one-pass grids whose observers read at most 1/SNAPSHOT_MAX_FRACTION of the cells (in
        at most SNAPSHOT_BOXES merged boxes)."""
        out = ['split']
        if self.one_pass_supported:
            out.append('one_pass')
            if self._two_step_eligible():
                out.append('two_step')
        return out

    @property
    def path(self):
        return 'two_step' if self.two_step else 'one_pass' if self.one_pass else 'split'

    def set_path(self, path):
        """Switch the step path between whole units (the state in the grid's own buffers, as after any
        step_pair, quad or single step of the split path); graphs captured before must be captured again."""
        if path not in self.paths():
            raise ValueError(f'path {path!r} is not available on this grid (available: {self.paths()})')
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError('the step path changes outside graph capture only')
        self.one_pass = path != 'split'
        self.two_step = path == 'two_step'
        if self.one_pass and self.materials is None:
            self.materials = MaterialIndex(self.grid)
        self._describe()

    def set_trace_points(self, points):
        """Replace the recorded traces by `points`: (magnetic, flat element index) pairs."""
        self.trace_points = [(bool(m), int(flat)) for m, flat in points]
        self.traces = torch.zeros((self.steps, len(self.trace_points)), device=self.grid.device, dtype=self.grid.dtype)
        self.two_step = self.two_step and self._two_step_allowed()   # the traces may widen the snapshot
        self._describe()

    def _trace_table(self, E, H):
        """int64 (address, element index) of every point trace in the fields E, H."""
        rows = [((H if magnetic else E).data_ptr(), flat) for magnetic, flat in self.trace_points]
        return torch.as_tensor(np.array(rows, dtype=np.int64).reshape(-1, 2), device=E.device)

    def _observe(self, table):
        g = self.grid
        o = N.Observe()
        o.dtype = N.F64 if g.dtype == torch.float64 else N.F32
        o.count = len(self.trace_points)
        o.table = table.data_ptr() if o.count else None
        o.traces = self.traces.data_ptr() if o.count else None
        o.steps = self.steps
        o.counter = self.counter.data_ptr()
        return o

    def _planes(self, E, H):
        table = (N.Plane*max(1, len(self.planes)))()
        for k, plane in enumerate(self.planes):
            table[k] = plane.descriptor(self.counter, E, H)
        return table

    def _describe(self):
        g = self.grid
        dtype = N.F64 if g.dtype == torch.float64 else N.F32
        e_sources = self.sources['E'] if self.kernel == 'hip' else ()
        if self.dispersive:
            # The E sources follow the ADE correction: a source-only launch after it.
            self.e_inject = half_step_descriptor(g, False, e_sources, self.counter)
            e_sources = ()
        self.e_step = half_step_descriptor(g, False, e_sources, self.counter)
        self.h_step = half_step_descriptor(g, True, self.sources['H'] if self.kernel == 'hip' else (), self.counter)
        self.plane_table = self._planes(g.E, g.H)
        self.trace_table = self._trace_table(g.E, g.H)
        self.observe = self._observe(self.trace_table)
        if self.one_pass:
            if self.materials is not None:
                self.materials.refresh()     # the permittivity may have been replaced (a widened layout)
            # Second buffer set of the one-pass step: time n is read from one set, n+1 written to the other
            # (allocated once; a retile keeps it).
            if getattr(self, 'E_b', None) is None:
                self.E_b, self.H_b = torch.empty_like(g.E), torch.empty_like(g.H)
                self.psi_e_b = torch.zeros_like(g.psi_buffers[False])
            psi_a = g.psi_buffers[False]
            self.march = [march_descriptor(g, g.E, g.H, self.E_b, self.H_b, psi_a, self.psi_e_b, self.sources, self.counter,
                                           self.tile_y, self.chunk_x, self.materials),
                          march_descriptor(g, self.E_b, self.H_b, g.E, g.H, self.psi_e_b, psi_a, self.sources, self.counter,
                                           self.tile_y, self.chunk_x, self.materials)]
            # The step reads its descriptor from device memory (a batch of one grid): the kernel then
            # loads tile parameters on demand instead of holding the whole descriptor in registers.
            self.march_tables = []
            for d in self.march:
                table, prefix, host = pack_table(lambda n: self.lib.tfdtd_batch_march_bytes(n, dtype),
                                                 self.lib.tfdtd_batch_march_pack, [d], 3, g.device)
                self.march_tables.append((table, prefix, (ctypes.c_int32*3)(*host[:, 1].tolist())))
            if self.two_step:
                # The two-step pass double-buffers the H-family CPML memory as well.
                if getattr(self, 'psi_h_b', None) is None:
                    self.psi_h_b = torch.zeros_like(g.psi_buffers[True])
                psi_h = g.psi_buffers[True]
                # Observers sample step n+1 from a snapshot of the cells they read, which the pass
                # writes box after box; the snapshot's plane and trace tables index into it.
                self.snapshot_boxes = boxes = self.observed_boxes()
                snapshot = None
                if boxes:
                    total = sum(_cells(b) for b in boxes)
                    if getattr(self, 'snap_e', None) is None or self.snap_e.shape[0] != total:
                        self.snap_e = torch.zeros((total, 3), device=g.E.device, dtype=g.dtype)
                        self.snap_h = torch.zeros((total, 3), device=g.E.device, dtype=g.dtype)
                    snapshot = (self.snap_e, self.snap_h, boxes)
                    self.plane_table_s = (N.Plane*max(1, len(self.planes)))()
                    self.snap_indices = [[self._snapshot_index(indices, boxes) for indices, _ in plane.maps]
                                         for plane in self.planes]
                    for k, plane in enumerate(self.planes):
                        self.plane_table_s[k] = plane.descriptor(self.counter, self.snap_e, self.snap_h,
                                                                 indices=self.snap_indices[k])
                    flats = self._snapshot_index(torch.as_tensor([flat for _, flat in self.trace_points], dtype=torch.long,
                                                                 device=g.E.device), boxes).tolist()
                    rows = [((self.snap_h if magnetic else self.snap_e).data_ptr(), f)
                            for (magnetic, _), f in zip(self.trace_points, flats)]
                    self.trace_table_s = torch.as_tensor(np.array(rows, dtype=np.int64).reshape(-1, 2), device=g.E.device)
                    self.observe_s = self._observe(self.trace_table_s)
                self.march2 = [march_descriptor(g, g.E, g.H, self.E_b, self.H_b, psi_a, self.psi_e_b, self.sources,
                                                self.counter, self.waves2, self.chunk_x2, self.materials,
                                                psi_h=(psi_h, self.psi_h_b), snapshot=snapshot),
                               march_descriptor(g, self.E_b, self.H_b, g.E, g.H, self.psi_e_b, psi_a, self.sources,
                                                self.counter, self.waves2, self.chunk_x2, self.materials,
                                                psi_h=(self.psi_h_b, psi_h), snapshot=snapshot)]
                self.march2_tables = [self._march2_table(d) for d in self.march2]
            self.plane_table_b = self._planes(self.E_b, self.H_b)
            self.trace_table_b = self._trace_table(self.E_b, self.H_b)
            self.observe_b = self._observe(self.trace_table_b)

    def oscillator_energy(self):
        """FP64 energy of every dispersive material's oscillators (a device scalar)."""
        return sum(m.energy() for m in self.dispersive)

    def retile(self, tile_y=None, chunk_x=None, waves2=None, chunk_x2=None):
        """Rebuild the step's descriptors for other launch shapes (rows per one-pass tile, planes per
        chunk, wavefronts per two-step tile, planes per two-step chunk); the state is kept.  Graphs
        captured before hold the old descriptors and must be captured again."""
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError('launch shapes change outside graph capture only')
        self.tile_y = self.tile_y if tile_y is None else tile_y
        self.chunk_x = self.chunk_x if chunk_x is None else chunk_x
        self.waves2 = self.waves2 if waves2 is None else waves2
        self.chunk_x2 = self.chunk_x2 if chunk_x2 is None else chunk_x2
        self._describe()

    def launch_shape(self):
        return dict(tile_y=self.tile_y, chunk_x=self.chunk_x, waves2=self.waves2, chunk_x2=self.chunk_x2)

    def _march2_table(self, d):
        table, prefix, host = pack_table(lambda n: self.lib.tfdtd_batch_march_bytes(n, N.F32),
                                         self.lib.tfdtd_batch_march2_pack, [d], 3, self.grid.device)
        return table, prefix, (ctypes.c_int32*3)(*host[:, 1].tolist())

    def states(self):
        extra = [self.E_b, self.H_b, self.psi_e_b] if self.one_pass else []
        if self.two_step:
            extra.append(self.psi_h_b)
        return [self.grid.E, self.grid.H, self.counter, self.traces, *self.grid.memory_states, *extra]

    def refresh_materials(self):
        """Carry a change of the grid's 1/eps into the one-pass step's material index: rebuilds the
        index and rewrites the device descriptors in place (captured graphs stay valid).  A version
        check when nothing changed; called before every eager one-pass step and graph replay."""
        m = self.materials
        if m is None or not m.stale():
            return False
        inv = self.grid.inverse_permittivity
        if any(d.inverse_permittivity != inv.data_ptr() or d.eps_stride != inv.shape[-1] for d in self.march):
            raise RuntimeError('the inverse permittivity tensor was replaced after the step was built; '
                               'build a new Stepper (in-place changes are followed)')
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError('the permittivity changed while the step was being captured')
        m.refresh()
        dtype = self.march[0].dtype
        for d, (table, prefix, totals) in zip(self.march, self.march_tables):
            m.describe(d)
            fresh, _, _ = pack_table(lambda n: self.lib.tfdtd_batch_march_bytes(n, dtype), self.lib.tfdtd_batch_march_pack,
                                     [d], 3, self.grid.device)
            table.copy_(fresh)
        for d, (table, _, _) in zip(self.march2, self.march2_tables) if self.two_step else ():
            m.describe(d)
            table.copy_(self._march2_table(d)[0])
        return True

    def march_step(self, k, stream):
        """One one-pass step from buffer set k (0: grid buffers -> second set, 1: back)."""
        self.refresh_materials()
        table, prefix, totals = self.march_tables[k]
        d = self.march[k]
        N.check(self.lib.tfdtd_batch_march(table.data_ptr(), prefix.data_ptr(), 1, d.dtype, self.tile_y, d.eps_stride,
                                           totals, stream), 'tfdtd_batch_march')

    def march2_step(self, k, stream):
        """Two steps in one pass from buffer set k (0: grid buffers -> second set, 1: back); the
        counter is not advanced."""
        self.refresh_materials()
        table, prefix, totals = self.march2_tables[k]
        d = self.march2[k]
        N.check(self.lib.tfdtd_batch_march2(table.data_ptr(), prefix.data_ptr(), 1, d.tile_y, d.eps_stride,
                                            int(d.snap_count > 0), totals, stream), 'tfdtd_batch_march2')

    def quad(self):
        """Four steps with the two-step kernel: grid buffers -> second set -> grid buffers.  After each
        pass the observers sample step n+1 from the snapshot, then step n+2 from the output buffers (the
        launches of two one-pass steps, in their order); without observers the observe launches only
        advance the counter."""
        stream = N.stream_handle(self.grid.device)
        lib = self.lib
        observed = bool(self.snapshot_boxes)
        for k, planes, observe in ((0, self.plane_table_b, self.observe_b), (1, self.plane_table, self.observe)):
            self.march2_step(k, stream)
            for table, obs in (((self.plane_table_s, self.observe_s) if observed else (planes, observe)), (planes, observe)):
                if self.planes:
                    N.check(lib.tfdtd_dft_update(table, len(self.planes), stream), 'tfdtd_dft_update')
                N.check(lib.tfdtd_observe_step(ctypes.byref(obs), stream), 'tfdtd_observe_step')

    def step_pair(self):
        """Two steps with the one-pass kernel: grid buffers -> second set -> grid buffers."""
        stream = N.stream_handle(self.grid.device)
        lib = self.lib
        for k, planes, observe in ((0, self.plane_table_b, self.observe_b), (1, self.plane_table, self.observe)):
            self.march_step(k, stream)
            if self.planes:
                N.check(lib.tfdtd_dft_update(planes, len(self.planes), stream), 'tfdtd_dft_update')
            N.check(lib.tfdtd_observe_step(ctypes.byref(observe), stream), 'tfdtd_observe_step')

    def advance(self, count):
        """`count` steps, eagerly: fours through the two-step kernel and pairs through the one-pass kernel
        where available, else single steps."""
        if self.two_step:
            for _ in range(count // 4):
                self.quad()
            count %= 4
        if self.one_pass:
            for _ in range(count // 2):
                self.step_pair()
            if count % 2:
                self.step()
        else:
            for _ in range(count):
                self.step()

    def step(self):
        stream = N.stream_handle(self.grid.device)
        if self.dispersive:
            self._dispersive_step(stream)
            return
        if self.kernel == 'hip':
            N.check(self.lib.tfdtd_step_fused(ctypes.byref(self.e_step), ctypes.byref(self.h_step), self.plane_table,
                                              len(self.planes), ctypes.byref(self.observe), 1, stream), 'tfdtd_step_fused')
            return
        g = self.grid
        g.update_E()
        self._inject(g.E, 'E')
        g.update_H()
        self._inject(g.H, 'H')
        if self.planes:
            N.check(self.lib.tfdtd_dft_update(self.plane_table, len(self.planes), stream), 'tfdtd_dft_update')
        N.check(self.lib.tfdtd_observe_step(ctypes.byref(self.observe), stream), 'tfdtd_observe_step')

    def _dispersive_step(self, stream):
        """The reference's order with MaterialADE (cuda_kernels.py:355-362, solver.py:785-797): save E at
        the owned samples, E update, ADE correction, E sources, H update and sources, planes, traces."""
        lib = self.lib
        for m in self.dispersive:
            N.check(lib.tfdtd_ade_save(ctypes.byref(m.descriptor), stream), 'tfdtd_ade_save')
        N.check(lib.tfdtd_yee_half_step(ctypes.byref(self.e_step), stream), 'tfdtd_yee_half_step')
        for m in self.dispersive:
            N.check(lib.tfdtd_ade_correct(ctypes.byref(m.descriptor), stream), 'tfdtd_ade_correct')
        N.check(lib.tfdtd_inject(ctypes.byref(self.e_inject), stream), 'tfdtd_inject')
        N.check(lib.tfdtd_yee_half_step(ctypes.byref(self.h_step), stream), 'tfdtd_yee_half_step')
        if self.planes:
            N.check(lib.tfdtd_dft_update(self.plane_table, len(self.planes), stream), 'tfdtd_dft_update')
        N.check(lib.tfdtd_observe_step(ctypes.byref(self.observe), stream), 'tfdtd_observe_step')

    def _inject(self, field, family):
        for comp, box, values in self.sources[family]:
            field[box+(comp,)] += values.index_select(0, self.counter)[0]


class Simulation:
    """Drop-in for the reference's Simulation (solver.py:642-670) on the HIP engine."""

    def __init__(self, project: Project, epsilon=None):
        self.project = Project.model_validate(project.model_dump())
        self.epsilon = epsilon

    def run(self, progress=None, cancel=None, cuda_graph=True, cuda_graph_steps=1, *, launch=None):
        """launch='auto' measures the step's launch shapes on this grid first (tuning.tune_launch)."""
        with ENGINE_LOCK, torch.no_grad():
            return self._run(progress, cancel, cuda_graph, cuda_graph_steps, launch)

    def _run(self, progress, cancel, cuda_graph, cuda_graph_steps, launch=None):
        p, r = self.project, self.project.region
        refusal = r.unsupported()
        if refusal:
            raise ValueError(refusal)
        if r.complex_fields:
            raise ValueError('Bloch (complex) fields are not run by this engine.')
        if r.backend == 'cpu':
            raise ValueError('This engine runs on the GPU (HIP, gfx950); backend="cpu" is the reference NumPy path.')
        if not torch.cuda.is_available():
            raise RuntimeError('The HIP engine requires a ROCm GPU; none is visible.')
        validate_graph_steps(cuda_graph_steps, cuda_graph)
        started = time.perf_counter()
        run = Stepper(p, epsilon=self.epsilon, launch=launch)
        g = run.grid
        control = r.run_control
        monitor = (RunMonitor([g], [p], check_count(p), energies=[run.oscillator_energy if run.dispersive else None])
                   if control.auto_shutoff or control.divergence_check else None)
        graph = (StepGraphs(run.step, run.states(), r.steps, cuda_graph_steps, pair=run.step_pair if run.one_pass else None,
                            prepare=run.refresh_materials if run.one_pass else None,
                            quad=run.quad if run.two_step else None)
                 if cuda_graph else None)
        view = DisplaySlice(p)
        plane_index = view.index
        eps_plane = view.epsilon(run.eps)
        interval = max(r.snapshot_interval, math.ceil(r.steps/100))
        frames, frame_steps = [], []
        torch.cuda.synchronize()
        setup_seconds = time.perf_counter()-started
        compute_start = time.perf_counter()
        completed, termination, diagnostic_seconds, replays = 0, 'max_steps', 0., 0
        barriers = set(range(interval, r.steps+1, interval))
        if monitor is not None:
            barriers.update(range(control.check_interval, r.steps+1, control.check_interval))
        width = graph.width if graph is not None else 1
        eager_steps = 0
        # Barriers enqueue the diagnostics and the frame copies on the GPU; the host waits only
        # where the run-control rule could stop the run at that barrier or a progress report
        # needs the frame.
        for target, advance in observation_schedule(r.steps, width, barriers):
            if cancel is not None and cancel.is_set():
                termination = 'cancelled'
                break
            if graph is not None:
                graph.replay(advance)
                replays += 1
            else:
                run.step()
                eager_steps += 1
            completed = target
            stop = False
            if monitor is not None and (completed % control.check_interval == 0 or completed == r.steps):
                t0 = time.perf_counter()
                monitor.check(completed)
                stop = monitor.stopped(completed)
                diagnostic_seconds += time.perf_counter()-t0
            if stop:
                termination = 'decayed'
            if completed % interval == 0 or completed == r.steps or stop:
                frames.append(view.device_frame(g))
                frame_steps.append(completed)
                if progress:
                    latest = monitor.latest() if monitor is not None else None
                    frame = frames[-1].cpu().numpy()
                    if not np.isfinite(frame).all():
                        raise FloatingPointError('Fields diverged. Check mesh, sources and precision.')
                    progress({'step': completed, 'total': r.steps, 'frame': frame.tolist(),
                              'elapsed': time.perf_counter()-compute_start,
                              'termination_reason': termination if stop else None, 'diagnostics': latest})
            if stop:
                break
        torch.cuda.synchronize()
        seconds = time.perf_counter()-compute_start
        frames = check_frames(monitor, frames, frame_steps)
        e, h = g.E.cpu().numpy().copy(), g.H.cpu().numpy().copy()
        if not np.isfinite(e).all() or not np.isfinite(h).all():
            raise FloatingPointError('Non-finite fields detected.')
        kernel_steps = dict(graph.kernel_steps) if graph is not None else dict(two_step=0, one_pass=0, split=eager_steps)
        summary = run_summary(p, run, e, completed, termination, seconds, setup_seconds, kernel_steps=kernel_steps,
                              launch=run.launch_shape() if run.one_pass else None, launch_tuning=run.launch_record,
                              cuda_graph=graph is not None, cuda_graph_steps=width, cuda_graph_replays=replays,
                              diagnostics=monitor.history(0) if monitor is not None else [],
                              diagnostic_seconds=diagnostic_seconds,
                              diagnostic_backend=monitor.backend if monitor is not None else None,
                              slice_index=plane_index)
        signature = run_signature(p, completed)
        planes = [dict(m.result(), run_signature=signature) for m in run.planes]
        return Result(p, summary, frames, np.array(frame_steps), eps_plane,
                      run.traces[:completed].cpu().numpy(), np.arange(1, completed+1)*g.time_step, e, h, planes)


class DisplaySlice:
    """The displayed field plane of a run (solver.py:747-756): axis, index and a <=256x256 subsample."""

    def __init__(self, project):
        r = project.region
        self.axis = 'xyz'.index(r.slice_axis)
        self.index = index_at(tuple(r.slice_position if i == self.axis else 0 for i in range(3)), r, r.field)[self.axis]
        self.component = 'xyz'.index(r.field[1].lower())
        self.electric = r.field[0] == 'E'

    def plane(self, a):
        sl = [slice(None)]*3
        sl[self.axis] = self.index
        data = a[tuple(sl)]
        return data[::max(1, math.ceil(data.shape[0]/256)), ::max(1, math.ceil(data.shape[1]/256))]

    def frame(self, grid):
        return self.device_frame(grid).cpu().numpy()

    def device_frame(self, grid):
        """A device copy of the displayed plane (no host synchronisation)."""
        return self.plane((grid.E if self.electric else grid.H)[..., self.component]).clone()

    def epsilon(self, eps):
        return np.array(self.plane(eps[..., self.component] if eps.ndim == 4 else eps))


def run_signature(project, steps):
    """Physical identity of a run, for matching a sample plane with its reference (solver.py:493-508):
    SHA-256 of the canonical JSON of the region settings without placement and display fields, the
    resolved sources, the step count and the mesh nodes."""
    r = project.region
    placement = {'backend', 'cuda_kernel', 'cuda_monitor_kernel', 'execution_mode', 'tiling', 'resident_cell_limit', 'field',
                 'slice_axis', 'slice_position', 'complex_display', 'snapshot_interval'}
    identity = dict(region=r.model_dump(exclude=placement),
                    sources=[project.resolved_source(s).model_dump() for s in project.sources], steps=steps,
                    nodes=[np.asarray(a).tolist() for a in r.mesh_nodes])
    return hashlib.sha256(json.dumps(identity, sort_keys=True).encode()).hexdigest()


def check_frames(monitor, frames, frame_steps):
    """Host copies of the device frames, after every run-control verdict has arrived.  The first
    failure in step order raises: a divergence verdict, or a non-finite displayed frame (a verdict
    at the same barrier comes first, as the verdict is evaluated before the frame is taken)."""
    failure = None
    if monitor is not None:
        try:
            monitor.finish()
        except FloatingPointError as exc:
            failure = exc
    host = np.array([f.cpu().numpy() for f in frames])
    for frame, step in zip(host, frame_steps):
        if not np.isfinite(frame).all():
            if failure is None or step < monitor.failed_step:
                raise FloatingPointError('Fields diverged. Check mesh, sources and precision.')
            break
    if failure is not None:
        raise failure
    return host


def run_summary(project, run, e, completed, termination, seconds, setup_seconds, kernel_steps=None, **extra):
    """Result.summary of one run (the keys of the reference's solver.py summary)."""
    r = project.region
    props = torch.cuda.get_device_properties(run.grid.device)
    summary = dict(shape=list(r.shape), cells=math.prod(r.shape), dt_fs=r.time_step*1e15,
                   duration_fs=r.time_step*r.steps*1e15, backend='cuda', precision=r.precision,
                   gpu=props.name, gpu_arch=getattr(props, 'gcnArchName', None),
                   cuda_kernel=run.kernel, cuda_monitor_kernel='hip', engine='torchfdtd_amd HIP gfx950',
                   kernel_steps=kernel_steps,
                   steps_per_pass=(0 if kernel_steps is None else 2 if kernel_steps['two_step'] else
                                   1 if kernel_steps['one_pass'] else 0),
                   steps=completed, requested_steps=r.steps, cancelled=termination == 'cancelled',
                   termination_reason=termination, auto_shutoff=termination == 'decayed',
                   seconds=seconds, setup_seconds=setup_seconds,
                   mcells_per_second=math.prod(r.shape)*completed/max(seconds, 1e-9)/1e6,
                   field_peak=float(np.max(np.abs(e))),
                   material_sampling=r.material_sampling, boundaries=r.boundaries.model_dump(),
                   plan_hash=resolve_plan(project).plan_hash,
                   units='geometry: um; time: s; E/H: reduced fields; Bloch phase: rad', warnings=[])
    summary.update(extra)
    return summary
