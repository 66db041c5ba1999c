"""The resolved simulation plan and its hash (the reference's resolve_plan / plan_hash,
torchfdtd/plan.py:497-539).

``resolve_plan(project)`` evaluates the host planning once and freezes the physical inputs of a run:
the mesh nodes and Yee component coordinates, the time base and its run-control shutoff settings,
the boundary faces with their CPML coefficient rows (FP64), the ADE coefficients of the dispersive
materials in use, the rasterization inputs of the material (enabled structures, active materials,
background, sampling), every enabled source's realized boxes and FP64 samples, and every monitor's
index or quadrature, interpolation maps and frequencies.

``plan_hash`` is SHA-256 over a canonical JSON of six sections -- mesh, time, exterior, material,
sources, monitors -- in which an array enters as its dtype, shape and the SHA-256 of its bytes.
Placement (backend, kernels, memory and execution mode, tiling, precision) stays outside the hash,
and so do object ids and names and the raw settings whose effect is already hashed through what
they produce (a source's pulse settings through its samples, a monitor's spectrum settings through
its frequencies and window).  The voxelized permittivity is a deterministic function of the hashed
rasterization inputs and is produced on demand (``plan.material``).

The section layout and the invariants follow the reference (tests/test_plan.py): equal physics
gives equal hashes whatever the placement; ``diff`` names the dotted keys that differ.  The
hash values themselves are this engine's: the reference's serialization is not reproduced byte
for byte.
"""
from __future__ import annotations

import copy
import hashlib
import json
from dataclasses import dataclass, field, fields, is_dataclass
from functools import cached_property

import numpy as np

COMPONENTS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')
# E is sampled after its update at (n+1) dt, H half a step later.
SAMPLE_TIME_STEPS = {'E': 1., 'H': 1.5}
FOURIER_CONVENTION = 'exp(+2 pi i f t)'
PLACEMENT_FIELDS = ('backend', 'cuda_kernel', 'cuda_monitor_kernel', 'memory_mode', 'execution_mode', 'tiling',
                    'precision')
_NOT_HASHED = {'hashed': False}


class PlanInvalidated(ValueError):
    """A prepared grid or plane no longer matches the plan it was built from."""


def _readonly(value):
    out = np.array(value, copy=True)
    out.setflags(write=False)
    return out


# ------------------------------------------------------------------ canonical form

def _array_record(a):
    """An array as dtype, shape and digest, after widening to FP64 / complex128 / int64."""
    a = np.asarray(a)
    kind = 'complex128' if np.iscomplexobj(a) else 'float64' if a.dtype.kind == 'f' else 'int64'
    a = np.ascontiguousarray(a, dtype=kind)
    return {'dtype': kind, 'shape': list(a.shape), 'sha256': hashlib.sha256(a.tobytes()).hexdigest()}


def _canon(value, browser=False):
    """JSON-able form of a plan value.  browser=True keeps 1-D arrays inline (to_json)."""
    if isinstance(value, np.ndarray):
        if browser and value.ndim <= 1:
            return ({'real': value.real.tolist(), 'imag': value.imag.tolist()} if np.iscomplexobj(value)
                    else value.tolist())
        return _array_record(value)
    if is_dataclass(value) and not isinstance(value, type):
        return {f.name: _canon(getattr(value, f.name), browser) for f in fields(value)
                if not f.name.startswith('_') and (browser or f.metadata.get('hashed', True))}
    if isinstance(value, dict):
        return {str(k): _canon(v, browser) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_canon(v, browser) for v in value]
    if isinstance(value, complex):
        return {'real': value.real, 'imag': value.imag}
    if isinstance(value, np.generic):
        return value.item()
    if value is None or isinstance(value, (bool, int, float, str)):
        return value
    raise TypeError(f'a plan value of type {type(value).__name__} has no canonical form')


def _leaves(value, key, out):
    """Dotted-key leaves of a canonical form (an array record is one leaf)."""
    if isinstance(value, dict) and 'sha256' not in value:
        for k, v in value.items():
            _leaves(v, f'{key}.{k}' if key else k, out)
    elif isinstance(value, list):
        for i, v in enumerate(value):
            _leaves(v, f'{key}[{i}]', out)
    else:
        out[key] = json.dumps(value, sort_keys=True)
    return out


def _sha256(payload):
    return hashlib.sha256(json.dumps(payload, sort_keys=True, allow_nan=False).encode()).hexdigest()


# ------------------------------------------------------------------ plan records

@dataclass(frozen=True, eq=False)
class CPMLSegment:
    """One CPML slab: rows [start, stop) of `axis` for the curl term differentiating `component`."""
    forward: bool
    axis: int
    component: int
    side: int
    start: int
    stop: int
    kappa: np.ndarray
    sigma: np.ndarray
    alpha: np.ndarray
    b: np.ndarray
    c: np.ndarray


@dataclass(frozen=True, eq=False)
class BoundaryPlan:
    faces: tuple
    bloch_phase: tuple
    courant_number: float
    wrap: dict
    pec_upper: dict
    cpml: tuple


@dataclass(frozen=True, eq=False)
class ADEPlan:
    """Trapezoidal ADE coefficients of one dispersive material at the plan's time step."""
    material: str
    epsilon_inf: float
    oscillators: tuple
    a: np.ndarray
    d: np.ndarray
    k: np.ndarray


@dataclass(frozen=True, eq=False)
class SourceTerm:
    component: str
    location: tuple
    sample_times: np.ndarray
    samples: np.ndarray


@dataclass(frozen=True, eq=False)
class SourcePlan:
    id: str = field(metadata=_NOT_HASHED)
    name: str = field(metadata=_NOT_HASHED)
    kind: str = 'point'
    injection: str = 'soft'
    settings: dict = field(default_factory=dict, metadata=_NOT_HASHED)
    polarization: tuple = ()
    terms: tuple = ()


@dataclass(frozen=True, eq=False)
class MonitorPlan:
    id: str = field(metadata=_NOT_HASHED)
    name: str = field(metadata=_NOT_HASHED)
    kind: str = 'point'
    components: tuple = ()
    index: tuple | None = None
    normal: str | None = None
    points_um: np.ndarray | None = None
    weights: np.ndarray | None = None
    shape: tuple | None = None
    maps: tuple = ()
    frequency_hz: np.ndarray | None = None
    apodization: dict = field(default_factory=dict)
    time_downsample: int = 1
    dft_precision: str | None = None
    spectrum: dict = field(default_factory=dict, metadata=_NOT_HASHED)


@dataclass(frozen=True, eq=False)
class SampledMaterial:
    """The voxelized relative permittivity and the material index of every sample."""
    epsilon: np.ndarray
    ownership: np.ndarray


SECTIONS = {
    'mesh': ('dimension', 'size', 'shape', 'nodes', 'axes'),
    'time': ('time_step', 'steps', 'run_control', 'sample_time_steps', 'fourier_convention'),
    'exterior': ('background_index', 'boundaries', 'pml_dispersion'),
    'material': ('material_sampling', 'interface_method', 'subpixel_quadrature', 'background_index', 'structures',
                 'materials', 'ade'),
}
JSON_FIELDS = ('dimension', 'size', 'shape', 'nodes', 'axes', 'time_step', 'steps', 'run_control', 'sample_time_steps',
               'fourier_convention', 'background_index', 'material_sampling', 'interface_method', 'subpixel_quadrature',
               'boundaries', 'pml_dispersion', 'structures', 'materials', 'ade', 'sources', 'monitors', 'placement')


@dataclass(frozen=True, eq=False)
class SimulationPlan:
    dimension: str
    size: tuple
    shape: tuple
    nodes: tuple
    axes: dict
    time_step: float
    steps: int
    run_control: dict
    sample_time_steps: dict
    fourier_convention: str
    background_index: float
    material_sampling: str
    interface_method: str
    subpixel_quadrature: int
    boundaries: BoundaryPlan
    pml_dispersion: str
    structures: tuple
    materials: tuple
    ade: tuple
    sources: tuple
    monitors: tuple
    resources: dict
    placement: dict
    plan_hash: str
    _sections: dict = field(repr=False)
    _project: object = field(repr=False)

    @cached_property
    def material(self):
        """The sampled permittivity and ownership, voxelized on first use by the engine's own planner."""
        from .planning import voxelize
        epsilon, ownership, _ = voxelize(self._project, with_ownership=True)
        return SampledMaterial(_readonly(epsilon), _readonly(ownership))

    @property
    def source_terms(self):
        """(component, location, FP64 samples) of every enabled source term, in source order."""
        return tuple((t.component, t.location, t.samples) for s in self.sources for t in s.terms)

    @property
    def sections(self):
        """A copy of the hashed canonical form, by section."""
        return copy.deepcopy(self._sections)

    def to_json(self):
        out = {name: _canon(getattr(self, name), browser=True) for name in JSON_FIELDS}
        out['resources'] = json.loads(json.dumps(self.resources, default=str))
        out['plan_hash'] = self.plan_hash
        return out

    def diff(self, other):
        """Sorted dotted keys of the hashed sections and of the placement that differ from `other`."""
        if not isinstance(other, SimulationPlan):
            raise TypeError('diff compares two SimulationPlan records')
        mine = _leaves(dict(self._sections, placement=_canon(self.placement)), '', {})
        theirs = _leaves(dict(other._sections, placement=_canon(other.placement)), '', {})
        return sorted(k for k in mine.keys() | theirs.keys() if mine.get(k) != theirs.get(k))

    def verify_grid(self, grid):
        """Raise PlanInvalidated where a prepared grid's time base, Courant number or CPML rows differ
        from the plan (coefficients compared in the grid's precision)."""
        if grid.time_step != self.time_step:
            raise PlanInvalidated('the grid time step differs from the plan')
        if grid.courant_number != self.boundaries.courant_number:
            raise PlanInvalidated('the grid Courant number differs from the plan')
        planned = {}
        for seg in self.boundaries.cpml:
            planned.setdefault((seg.forward, seg.axis, seg.component), []).append(seg)
        actual = {k: v for k, v in grid.cpml.items() if v}
        if actual.keys() != planned.keys() or any(len(actual[k]) != len(planned[k]) for k in planned):
            raise PlanInvalidated('the grid CPML slab layout differs from the plan')
        for key, segments in planned.items():
            for seg, entry in zip(segments, actual[key]):
                if (entry['lo'], entry['hi']) != (seg.start, seg.stop):
                    raise PlanInvalidated(f'the grid CPML rows of axis {seg.axis} differ from the plan')
                for name, expected in (('b', seg.b), ('c', seg.c), ('inv_k', 1/seg.kappa)):
                    got = entry[name].detach().cpu().numpy().reshape(-1)
                    if not np.array_equal(got, expected.astype(got.dtype)):
                        raise PlanInvalidated(f'the grid CPML coefficient {name} of axis {seg.axis} differs from the plan')


# ------------------------------------------------------------------ resolvers

def _run_control(region):
    """The run-control settings the completed step count depends on: with auto_shutoff the decay test
    decides when the run stops; the divergence checks only abort a run, never change its result."""
    c = region.run_control
    if not c.auto_shutoff:
        return {'auto_shutoff': False}
    return {'auto_shutoff': True, 'decay_threshold': float(c.decay_threshold), 'check_interval': int(c.check_interval),
            'consecutive_checks': int(c.consecutive_checks), 'min_steps': int(c.min_steps),
            'source_tail_amplitude': float(c.source_tail_amplitude), 'after_source_s': float(c.after_source_s)}


def _boundaries(region):
    from .planning import cpml_slabs, pec_upper, periodic_axes
    faces = []
    for axis in range(3):
        for side, face in enumerate(region.boundaries.pair(axis)):
            settings = {k: v for k, v in face.model_dump().items() if k not in ('kind', 'layers')}
            faces.append({'axis': 'xyz'[axis], 'side': ('min', 'max')[side], 'kind': face.kind,
                          'layers': region.pml_layers(axis, side), **settings})
    segments = []
    for (forward, axis, comp), slabs in cpml_slabs(region, dtype=np.float64).items():
        for s in slabs:
            segments.append(CPMLSegment(bool(forward), int(axis), int(comp), int(s.side), int(s.lo), int(s.hi),
                                        *(_readonly(np.asarray(v, dtype=np.float64))
                                          for v in (s.kappa, s.sigma, s.alpha, s.b, s.c))))
    wrap = {a: complex(1.0) for a in periodic_axes(region)}
    return BoundaryPlan(tuple(faces), tuple(float(v) for v in region.bloch_phase), float(region.rectangular_courant),
                        wrap, {a: float(v) for a, v in pec_upper(region).items()}, tuple(segments))


def _ade(project):
    """a, d, k of every active dispersive material, formed in FP64 as the step forms them."""
    r = project.region
    active = {s.material for s in project.structures if s.enabled}
    dt = r.time_step
    out = []
    for m in project.materials:
        if not m.oscillators or m.name not in active:
            continue
        rates = np.asarray(m.oscillators, dtype=np.float64).reshape(-1, 3)
        w0, strength, gamma = rates[:, 0], rates[:, 1], rates[:, 2]
        denominator = 1+.5*gamma*dt+.25*(w0*dt)**2
        out.append(ADEPlan(m.name, float(m.epsilon_inf), tuple(tuple(float(v) for v in o) for o in m.oscillators),
                           _readonly(.5*(w0*dt)**2), _readonly(denominator), _readonly(strength*dt*dt/(4*denominator))))
    return tuple(out)


def _sources(project):
    from .planning import source_terms
    r = project.region
    times = np.arange(1, r.steps+1)*r.time_step
    out = []
    for raw in project.sources:
        if not raw.enabled:
            continue
        s = project.resolved_source(raw)
        terms = []
        for t in source_terms(project, dtype=np.float64, only=raw):
            component = t.family+'xyz'[t.comp]
            offset = (SAMPLE_TIME_STEPS[t.family]-1.)*r.time_step
            box = tuple((int(b.start), int(b.stop)) if isinstance(b, slice) else int(b) for b in t.box)
            terms.append(SourceTerm(component, box, _readonly(times+offset), _readonly(t.values)))
        out.append(SourcePlan(s.id, s.name, s.kind, s.injection,
                              s.model_dump(mode='json', exclude={'id', 'name', 'enabled', 'signal'}),
                              tuple((c, float(w)) for c, w in s.polarization_components), tuple(terms)))
    return tuple(out)


def _monitors(project):
    from .monitors import frequency_samples, interpolation_map, plane_plan
    from .planning import index_at
    r = project.region
    out = []
    for raw in project.monitors:
        if not raw.enabled:
            continue
        m = project.resolved_monitor(raw)
        apodization = {'mode': m.spectrum.apodization}
        if m.spectrum.apodization in ('start', 'end', 'full'):
            apodization.update(center_s=m.spectrum.apodization_center, time_width_s=m.spectrum.apodization_time_width)
        spectrum = m.spectrum.model_dump(mode='json')
        if m.kind == 'field':
            plan = plane_plan(r, m)
            maps = tuple((c, *(_readonly(v) for v in interpolation_map(r, c, plan['points_um']))) for c in m.required_fields)
            out.append(MonitorPlan(m.id, m.name, 'field', tuple(m.required_fields), None, m.normal,
                                   _readonly(plan['points_um']), _readonly(plan['weights']), tuple(plan['shape']), maps,
                                   _readonly(frequency_samples(m.spectrum)), apodization, m.time_downsample,
                                   m.dft_precision, spectrum))
        else:
            frequency = None if m.spectrum.sampling == 'fft' else _readonly(frequency_samples(m.spectrum))
            out.append(MonitorPlan(m.id, m.name, 'point', (m.component,),
                                   tuple(int(i) for i in index_at(m.center, r, m.component)), None, None, None, None, (),
                                   frequency, apodization, m.time_downsample, None, spectrum))
    return tuple(out)


def _resources(project):
    """Cells, steps and the resident device bytes of the fields, permittivity and CPML memory."""
    from .planning import psi_elements
    r = project.region
    cells = int(np.prod(r.shape))
    item = 8 if r.precision == 'float64' else 4
    eps = 3 if r.material_sampling == 'yee' else 1
    psi = psi_elements(r)
    return {'cells': cells, 'steps': int(r.steps), 'precision': r.precision, 'psi_elements': int(psi),
            'field_bytes': (6+eps)*cells*item, 'cpml_bytes': psi*item}


def resolve_plan(project):
    """Resolve the physical inputs of `project` once and return the frozen SimulationPlan."""
    from .models import Project
    from .planning import field_axes
    project = Project.model_validate(project.model_dump())
    r = project.region
    active = {s.material for s in project.structures if s.enabled}
    values = dict(
        dimension=r.dimension, size=tuple(float(v) for v in r.size), shape=tuple(int(n) for n in r.shape),
        nodes=tuple(_readonly(a) for a in r.mesh_nodes),
        axes={c: tuple(_readonly(a) for a in field_axes(r, c)) for c in COMPONENTS},
        time_step=float(r.time_step), steps=int(r.steps), run_control=_run_control(r),
        sample_time_steps=dict(SAMPLE_TIME_STEPS), fourier_convention=FOURIER_CONVENTION,
        background_index=float(r.background_index), material_sampling=r.material_sampling,
        interface_method=r.interface_method, subpixel_quadrature=int(r.subpixel_quadrature),
        boundaries=_boundaries(r), pml_dispersion=r.pml_dispersion,
        structures=tuple(dict(s.model_dump(mode='json', exclude={'id', 'name', 'enabled'}), order=i)
                         for i, s in enumerate(project.structures) if s.enabled),
        materials=tuple({'name': m.name, 'model': m.model,
                         'epsilon': list(m.epsilon_tensor) if m.model == 'tensor' else m.instantaneous_epsilon,
                         'oscillators': [list(o) for o in m.oscillators]}
                        for m in project.materials if m.name in active),
        ade=_ade(project), sources=_sources(project), monitors=_monitors(project))
    sections = {name: {k: _canon(values[k]) for k in keys} for name, keys in SECTIONS.items()}
    sections['sources'] = _canon(values['sources'])
    sections['monitors'] = _canon(values['monitors'])
    placement = {k: (r.tiling.model_dump(mode='json') if k == 'tiling' else getattr(r, k)) for k in PLACEMENT_FIELDS}
    return SimulationPlan(**values, resources=_resources(project), placement=placement, plan_hash=_sha256(sections),
                          _sections=sections, _project=project)
