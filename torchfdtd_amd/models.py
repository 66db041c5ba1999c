"""Project schema: the reference's JSON layout (torchfdtd/models.py), so that projects
written for the reference load unchanged.

Fields keep the reference's names, types and defaults.  Fields that select physics
this engine does not run (dispersion, Bloch phases, subpixel interfaces, graded
meshes, one-way and TFSF sources, streamed execution) are accepted here and refused
with a ValueError by ``Simulation.run`` before anything is allocated.

``Region.cuda_kernel`` gains the value ``'hip'``: the hand-written gfx950 kernels of
this package.  ``'auto'`` and ``'fused'`` resolve to ``'hip'``.
"""
from __future__ import annotations

import math
from typing import Any, Literal
from uuid import uuid4

from pydantic import BaseModel, ConfigDict, Field, model_serializer, model_validator

C0 = 299792458.0


class Model(BaseModel):
    model_config = ConfigDict(extra='forbid', allow_inf_nan=False)


class Material(Model):
    name: str = Field(min_length=1, max_length=100)
    index: float = Field(default=1.5, ge=1, le=20)
    color: str = '#6ca8dd'
    model: Literal['dielectric', 'tensor', 'drude', 'lorentz', 'multipole'] = 'dielectric'
    epsilon_tensor: tuple[float, float, float, float, float, float] = (2.25, 2.25, 2.25, 0., 0., 0.)
    epsilon_inf: float = Field(default=1, ge=1, le=400)
    plasma_rad_s: float = Field(default=2e15, gt=0, le=1e18)
    collision_rad_s: float = Field(default=1e14, ge=0, le=1e18)
    resonance_rad_s: float = Field(default=2e15, gt=0, le=1e18)
    linewidth_rad_s: float = Field(default=1e14, ge=0, le=1e18)
    delta_epsilon: float = Field(default=1, gt=0, le=10000)
    poles: list[dict[str, Any]] = Field(default_factory=list, max_length=16)
    samples: dict[str, Any] | None = None
    fit_band_um: tuple[float, float] | None = None
    fit_dt_s: float | None = Field(default=None, gt=0)
    provenance: dict[str, Any] | None = None

    @property
    def instantaneous_epsilon(self):
        """models.py:150-154: index**2 for a dielectric."""
        if self.model == 'tensor':
            raise ValueError('Tensor permittivity is not run by this engine.')
        return self.index**2 if self.model == 'dielectric' else self.epsilon_inf

    @property
    def oscillators(self):
        """(w0, strength, damping) in rad/s of each pole (models.py:161-174): Drude (0, wp^2, collision),
        Lorentz (w0, delta_eps w0^2, 2 linewidth), multipole poles as listed; () for a dielectric."""
        if self.model == 'drude':
            return ((0., self.plasma_rad_s**2, self.collision_rad_s),)
        if self.model == 'lorentz':
            return ((self.resonance_rad_s, self.delta_epsilon*self.resonance_rad_s**2, 2*self.linewidth_rad_s),)
        if self.model == 'multipole':
            return tuple((float(p['resonance_rad_s']), float(p['strength_rad_s_squared']), float(p['damping_rad_s']))
                         for p in self.poles)
        return ()


class BoundaryFace(Model):
    kind: Literal['pml', 'periodic', 'bloch', 'pec', 'antisymmetric', 'pmc', 'symmetric'] = 'pml'
    layers: int | None = Field(default=None, ge=3, le=100)
    sigma_scale: float = Field(default=1, gt=0, le=20)
    kappa: float = Field(default=1, ge=1, le=50)
    alpha: float = Field(default=1e-8, ge=0, le=5)
    polynomial: float = Field(default=3, ge=1, le=10)
    alpha_polynomial: float = Field(default=0, ge=0, le=10)


class Boundaries(Model):
    x_min: BoundaryFace = Field(default_factory=BoundaryFace)
    x_max: BoundaryFace = Field(default_factory=BoundaryFace)
    y_min: BoundaryFace = Field(default_factory=BoundaryFace)
    y_max: BoundaryFace = Field(default_factory=BoundaryFace)
    z_min: BoundaryFace = Field(default_factory=BoundaryFace)
    z_max: BoundaryFace = Field(default_factory=BoundaryFace)

    def pair(self, axis):
        return tuple(getattr(self, 'xyz'[axis]+'_'+side) for side in ('min', 'max'))


class MeshRefinement(Model):
    name: str = Field(default='Refinement', min_length=1, max_length=100)
    center: tuple[float, float, float] = (0, 0, 0)
    size: tuple[float, float, float] = (1, 1, 1)
    enabled: bool = True


class RunControl(Model):
    auto_shutoff: bool = False
    decay_threshold: float = Field(default=1e-6, gt=0, lt=1)
    check_interval: int = Field(default=50, ge=1, le=10000)
    consecutive_checks: int = Field(default=3, ge=2, le=100)
    min_steps: int = Field(default=100, ge=10, le=100000)
    source_tail_amplitude: float = Field(default=1e-8, gt=0, le=1e-3)
    after_source_s: float = Field(default=0, ge=0)
    divergence_check: bool = True
    growth_limit: float = Field(default=1e6, gt=1)
    field_limit: float | None = Field(default=None, gt=0)


class Tiling(Model):
    size_um: float = Field(default=20, gt=0)
    overlap_um: float = Field(default=2, gt=0)
    max_angle_deg: float = Field(default=45, ge=0, lt=90)
    propagation_um: float | None = Field(default=None, ge=0)
    allow_approximate: bool = False


class Region(Model):
    dimension: Literal['2d', '3d'] = '2d'
    size: tuple[float, float, float] = (8, 6, 2)
    mesh: float = Field(default=0.05, gt=0, le=10)
    mesh_steps: tuple[float, float, float] | None = None
    mesh_coordinates: tuple[tuple[float, ...], tuple[float, ...], tuple[float, ...]] | None = None
    material_sampling: Literal['cell', 'yee'] = 'cell'
    interface_method: Literal['staircase', 'subpixel'] = 'staircase'
    subpixel_quadrature: int = Field(default=8, ge=2, le=32)
    mesh_type: Literal['uniform', 'graded', 'explicit'] = 'uniform'
    mesh_max: float = Field(default=.15, gt=0, le=10)
    mesh_grading: float = Field(default=1.25, ge=1.05, le=1.5)
    mesh_ppw: float = Field(default=24, ge=6, le=80)
    mesh_auto_refine: bool = True
    mesh_refinements: list[MeshRefinement] = Field(default_factory=list)
    courant_factor: float = Field(default=0.99, gt=0, le=0.99)
    time_step_override: float | None = Field(default=None, gt=0)
    steps: int = Field(default=1000, ge=10)
    run_control: RunControl = Field(default_factory=RunControl)
    pml_cells: int = Field(default=10, ge=3, le=50)
    pml_dispersion: Literal['ade', 'frozen', 'absorber'] = 'ade'
    boundaries: Boundaries = Field(default_factory=Boundaries)
    bloch_phase: tuple[float, float, float] = (0, 0, 0)
    background_index: float = Field(default=1, ge=1, le=20)
    backend: Literal['auto', 'cuda', 'cpu'] = 'auto'
    memory_mode: Literal['resident', 'streamed', 'budgeted'] = 'resident'
    execution_mode: Literal['auto', 'resident', 'streamed_host', 'streamed_disk', 'tiled'] = 'auto'
    tiling: Tiling = Field(default_factory=Tiling)
    resident_cell_limit: int | None = Field(default=None, ge=1, strict=True)
    cuda_kernel: Literal['auto', 'torch', 'fused', 'hip'] = 'auto'
    cuda_monitor_kernel: Literal['auto', 'torch', 'fused', 'hip'] = 'auto'
    precision: Literal['float32', 'float64'] = 'float32'
    snapshot_interval: int = Field(default=20, ge=1, le=10000)
    field: Literal['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'] = 'Ez'
    slice_axis: Literal['x', 'y', 'z'] = 'z'
    slice_position: float = 0
    complex_display: Literal['real', 'imag', 'magnitude', 'phase'] = 'real'

    @model_serializer(mode='wrap')
    def _omit_unset_fields(self, handler):
        # models.py:306-317: 'auto' kernels and an unset cell cap are not written.
        data = handler(self)
        if self.resident_cell_limit is None:
            data.pop('resident_cell_limit', None)
        for name in ('cuda_kernel', 'cuda_monitor_kernel'):
            if data.get(name) == 'auto':
                del data[name]
        return data

    @property
    def active_axes(self):
        return 2 if self.dimension == '2d' else 3

    @property
    def complex_fields(self):
        return any(self.boundaries.pair(i)[0].kind == 'bloch' for i in range(self.active_axes))

    def pml_layers(self, axis, side):
        face = self.boundaries.pair(axis)[side]
        return (face.layers or self.pml_cells) if face.kind == 'pml' else 0

    @property
    def axis_steps(self):
        return self.mesh_steps or (self.mesh,)*3

    @property
    def reference_step(self):
        return min(self.axis_steps[:self.active_axes])

    @property
    def rectangular_courant(self):
        """S = c dt / reference_step (models.py:377-381, uniform branch)."""
        return self.courant_factor/math.sqrt(self.active_axes)

    @property
    def base_shape(self):
        """Cells per axis (models.py:383-394): an integer ratio a few ulps off is not a new cell."""
        def cells(span, step):
            ratio = span/step
            nearest = round(ratio)
            return max(1, nearest if abs(ratio-nearest) <= 8*math.ulp(ratio) else math.ceil(ratio))
        return tuple(cells(s, self.axis_steps[i]) if i < 2 or self.dimension == '3d' else 1 for i, s in enumerate(self.size))

    @property
    def shape(self):
        return self.base_shape

    @property
    def time_step(self):
        """models.py:403-412, uniform branch."""
        return self.time_step_override or self.courant_factor/math.sqrt(self.active_axes)*self.mesh*1e-6/C0

    @property
    def actual_size(self):
        return tuple(n*self.axis_steps[i] if n > 1 else self.size[i] for i, n in enumerate(self.base_shape))

    @property
    def mesh_nodes(self):
        from .planning import mesh_nodes
        return mesh_nodes(self)

    def interior_bounds(self, axis):
        if axis == 2 and self.dimension == '2d':
            return (-self.size[2]/2, self.size[2]/2)
        nodes = self.mesh_nodes[axis]
        lo, hi = self.pml_layers(axis, 0), self.pml_layers(axis, 1)
        return (float(nodes[lo]) if lo else -self.actual_size[axis]/2,
                float(nodes[-hi-1]) if hi else self.actual_size[axis]/2)

    def unsupported(self):
        """Why this engine cannot run the region, or None (checked before any allocation)."""
        if self.mesh_type != 'uniform' or self.mesh_steps is not None or self.mesh_coordinates is not None:
            return 'graded, explicit and axis-specific meshes are not run by this engine'
        if self.time_step_override is not None:
            return 'time_step_override is not run by this engine'
        if self.interface_method != 'staircase':
            return 'subpixel interfaces are not run by this engine'
        if self.memory_mode != 'resident':
            return 'streamed and budgeted memory modes are not run by this engine'
        for axis in range(3):
            for face in self.boundaries.pair(axis):
                if face.kind in ('bloch', 'antisymmetric', 'pmc', 'symmetric'):
                    return f'{face.kind} faces are not run by this engine (pml, periodic and pec are)'
        return None

    @model_validator(mode='after')
    def valid_grid(self):
        if any(s <= 0 for s in self.size):
            raise ValueError('Region spans must be positive.')
        active = self.shape[:2] if self.dimension == '2d' else self.shape
        for axis, n in enumerate(active):
            low, high = self.boundaries.pair(axis)
            if low.kind != high.kind and ({low.kind, high.kind} & {'periodic', 'bloch'}):
                raise ValueError('Periodic/Bloch boundaries must be paired on the same axis.')
            if low.kind != 'bloch' and self.bloch_phase[axis] != 0:
                raise ValueError('A nonzero Bloch phase requires Bloch boundaries on that axis.')
            if n <= self.pml_layers(axis, 0)+self.pml_layers(axis, 1)+4:
                raise ValueError('Mesh must leave at least 5 cells between the PML boundaries.')
        if self.dimension == '2d' and (any(f.kind != 'pml' for f in self.boundaries.pair(2)) or self.bloch_phase[2] != 0):
            raise ValueError('The invariant z axis has no boundary condition in 2D; keep its defaults.')
        if max(self.shape) > 1_000_000:
            raise ValueError('A grid axis may contain at most one million cells.')
        if self.dimension == '2d' and self.slice_axis != 'z':
            raise ValueError('2D simulations use the XY (z-normal) field plane.')
        if abs(self.slice_position) > self.size['xyz'.index(self.slice_axis)]/2:
            raise ValueError('Field slice must lie inside the simulation region.')
        return self


class Item(Model):
    id: str = Field(default_factory=lambda: uuid4().hex[:12], min_length=1, max_length=100)
    name: str = Field(default='object', min_length=1, max_length=100)
    center: tuple[float, float, float] = (0, 0, 0)
    enabled: bool = True


class Structure(Item):
    kind: Literal['rectangle', 'circle', 'sphere', 'ring', 'polygon'] = 'rectangle'
    size: tuple[float, float, float] = (1, 1, 0.5)
    radius: float = Field(default=0.5, gt=0)
    inner_radius: float = Field(default=0.3, ge=0)
    rotation: float = 0
    rotation_axes: tuple[Literal['x', 'y', 'z', 'none'], Literal['x', 'y', 'z', 'none'], Literal['x', 'y', 'z', 'none']] = ('z', 'x', 'y')
    rotation_angles: tuple[float, float, float] = (0, 0, 0)
    make_ellipsoid: bool = False
    radius_2: float = Field(default=.5, gt=0)
    radius_3: float = Field(default=.5, gt=0)
    inner_radius_2: float = Field(default=.3, ge=0)
    theta_start: float = 0
    theta_stop: float = 360
    vertices: tuple[tuple[float, float], ...] = ((-.5, -.5), (.5, -.5), (0, .5))
    holes: tuple[tuple[tuple[float, float], ...], ...] = ()
    material: str = 'SiN (constant n)'
    mesh_order: int = Field(default=2, ge=1, le=100)

    @model_validator(mode='after')
    def valid_shape(self):
        if any(s <= 0 for s in self.size):
            raise ValueError('Object spans must be positive.')
        if self.kind == 'ring':
            if self.inner_radius >= self.radius:
                raise ValueError('Ring inner radius must be less than its outer radius.')
            if self.make_ellipsoid and (self.inner_radius_2 >= self.radius_2
                                        or (self.inner_radius == 0) != (self.inner_radius_2 == 0)):
                raise ValueError('An elliptic ring needs both inner radii zero, or both positive and below the outer radii.')
            if not 0 < abs(self.theta_stop-self.theta_start) <= 360:
                raise ValueError('A ring arc spans more than 0 and at most 360 degrees.')
        if self.kind == 'polygon':
            check_contour(self.vertices)
            check_holes(self.vertices, self.holes)
        elif self.holes:
            raise ValueError('Holes belong to polygons only.')
        return self

    @property
    def angular_span(self):
        return (self.theta_stop-self.theta_start) % 360 or 360.


def _cross(a, b):
    return a[..., 0]*b[..., 1]-a[..., 1]*b[..., 0]


def _touching(p, q, eps, tol):
    """[len(p), len(q)] booleans: segment p_i of contour p meets (crosses or touches) segment q_j of q.
    Orientation tests carry `eps`, bounding boxes `tol`."""
    import numpy as np
    a, b = p[:, None], np.roll(p, -1, axis=0)[:, None]
    c, d = q[None], np.roll(q, -1, axis=0)[None]
    o1, o2 = _cross(b-a, c-a), _cross(b-a, d-a)
    o3, o4 = _cross(d-c, a-c), _cross(d-c, b-c)
    straddles = (((o1 <= eps) & (o2 >= -eps)) | ((o2 <= eps) & (o1 >= -eps))) & \
                (((o3 <= eps) & (o4 >= -eps)) | ((o4 <= eps) & (o3 >= -eps)))
    boxes = (np.maximum(np.minimum(a, b), np.minimum(c, d)) <= np.minimum(np.maximum(a, b), np.maximum(c, d))+tol).all(-1)
    return straddles & boxes


def _scale(*contours):
    import numpy as np
    extent = max(float(np.ptp(c, axis=0).max()) for c in contours)
    tol = 64*2.220446049250313e-16*max(extent, 1e-300)
    return tol, tol*extent


def check_contour(vertices):
    """One simple closed contour (either orientation): 3..2048 distinct consecutive vertices, nonzero
    area, no edge folding back on its neighbour, no edge meeting a non-adjacent edge."""
    import numpy as np
    v = np.asarray(vertices, dtype=float)
    if v.ndim != 2 or v.shape[1] != 2 or not 3 <= len(v) <= 2048 or not np.isfinite(v).all():
        raise ValueError('A polygon has 3 to 2048 finite (x, y) vertices.')
    tol, eps = _scale(v)
    edge = np.roll(v, -1, axis=0)-v
    if np.any(np.hypot(edge[:, 0], edge[:, 1]) <= tol):
        raise ValueError('A polygon edge has zero length (do not repeat the first vertex at the end).')
    w = v-v[0]
    if abs(float(np.sum(_cross(w, np.roll(w, -1, axis=0))))) <= eps:
        raise ValueError('A polygon needs a nonzero area.')
    before = np.roll(edge, 1, axis=0)
    if np.any((abs(_cross(before, edge)) <= eps) & ((before*edge).sum(-1) < 0)):
        raise ValueError('Adjacent polygon edges may not fold back onto each other.')
    n = len(v)
    meets = _touching(v, v, eps, tol)
    i, j = np.indices((n, n))
    apart = (j >= i+2) & ~((i == 0) & (j == n-1))
    if np.any(meets & apart):
        raise ValueError('Polygon edges may not cross or touch non-adjacent edges.')


def check_holes(outer, holes):
    """Holes: at most 1024 simple contours strictly inside the outer contour, apart from it and
    from each other (no crossing, touching or nesting)."""
    import numpy as np
    if len(holes) > 1024:
        raise ValueError('A polygon has at most 1024 holes.')
    outer = np.asarray(outer, dtype=float)
    loops = []
    for hole in holes:
        check_contour(hole)
        loop = np.asarray(hole, dtype=float)
        tol, eps = _scale(outer, loop)
        if np.any(_touching(loop, outer, eps, tol)) or not np.all(_strictly_inside(loop, outer)):
            raise ValueError('A polygon hole lies strictly inside the outer contour.')
        for other in loops:
            tol, eps = _scale(other, loop)
            if (np.any(_touching(loop, other, eps, tol)) or np.any(_inside_or_on(loop, other))
                    or np.any(_inside_or_on(other, loop))):
                raise ValueError('Polygon holes may not overlap, nest or touch.')
        loops.append(loop)


def _inside_or_on(points, contour):
    from .planning import _contour
    parity, edge = _contour(points[:, 0], points[:, 1], contour)
    return parity | edge


def _strictly_inside(points, contour):
    from .planning import _contour
    parity, edge = _contour(points[:, 0], points[:, 1], contour)
    return parity & ~edge


class TimeSignal(Model):
    time_s: list[float] = Field(min_length=2)
    amplitude: list[float] = Field(min_length=2)
    phase_rad: list[float] = Field(min_length=2)


class SourceTimeSettings(Model):
    wavelength: float = Field(default=1.55, gt=0)
    pulse: Literal['gaussian', 'continuous', 'sampled', 'broadband'] = 'gaussian'
    pulse_cycles: float = Field(default=3, ge=1, le=50)
    time_definition: Literal['cycles', 'standard', 'wavelength', 'frequency'] = 'cycles'
    pulse_length: float = Field(default=20e-15, gt=0)
    pulse_offset: float = Field(default=50e-15, ge=0)
    signal: TimeSignal | None = None
    wavelength_start: float = Field(default=1.3, gt=0)
    wavelength_stop: float = Field(default=1.8, gt=0)
    optimize_for_short_pulse: bool = True
    eliminate_discontinuities: bool = False
    chirp_bandwidth_hz: float = Field(default=100e12, gt=0)

    @model_validator(mode='after')
    def valid_waveform(self):
        if self.pulse == 'sampled' and self.signal is None:
            raise ValueError('A sampled source requires time, amplitude and phase data.')
        if self.wavelength_stop < self.wavelength_start:
            raise ValueError('Source stop wavelength must be at least its start wavelength.')
        if self.time_definition in ('wavelength', 'frequency') and self.pulse != 'broadband':
            raise ValueError('Automatic wavelength/frequency ranges require the broadband pulse generator.')
        if self.pulse == 'broadband' and self.time_definition == 'cycles':
            raise ValueError('Broadband pulses require a wavelength/frequency range or standard time-domain parameters.')
        return self


class Source(Item, SourceTimeSettings):
    name: str = 'source'
    kind: Literal['point', 'plane', 'tfsf'] = 'point'
    injection: Literal['soft', 'oneway'] = 'soft'
    normal: Literal['x', 'y', 'z'] = 'x'
    direction: Literal['+', '-'] = '+'
    incident_pml_cells: int = Field(default=96, ge=32, le=512)
    size: tuple[float, float, float] = (0, 2, 0)
    component: Literal['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'] = 'Ez'
    theta: float | None = Field(default=None, ge=0, le=180)
    phi: float = 0
    amplitude: float = Field(default=1, gt=0, le=100)
    phase: float = 0
    use_global_source: bool = False
    extend_through_pml: bool = False

    @property
    def polarization_components(self):
        """models.py:595-605: Cartesian component, or a unit vector in its family."""
        if self.theta is None:
            return ((self.component, 1.),)
        def sincos(degrees):
            cardinal = {0: (0., 1.), 90: (1., 0.), 180: (0., -1.), 270: (-1., 0.)}
            degrees %= 360
            return cardinal[degrees] if degrees in cardinal else (math.sin(math.radians(degrees)), math.cos(math.radians(degrees)))
        st, ct = sincos(self.theta)
        sp, cp = sincos(self.phi)
        return tuple((self.component[0]+axis, value) for axis, value in zip('xyz', (st*cp, st*sp, ct)) if value != 0)

    @property
    def time_offset_steps(self):
        return .5 if self.component.startswith('H') else 0.


class SpectrumSettings(Model):
    sampling: Literal['fft', 'frequency', 'wavelength', 'chebyshev', 'custom'] = 'fft'
    chebyshev_wavelength: bool = False
    chebyshev_nodes: Literal['roots', 'lobatto'] = 'roots'
    use_source_limits: bool = False
    custom_frequencies_hz: list[float] = Field(default_factory=list)
    wavelength_start: float = Field(default=1.3, gt=0)
    wavelength_stop: float = Field(default=1.8, gt=0)
    frequency_points: int = Field(default=101, ge=1)
    apodization: Literal['hann', 'none', 'start', 'end', 'full'] = 'hann'
    apodization_center: float = Field(default=20e-15, ge=0)
    apodization_time_width: float = Field(default=10e-15, gt=0)


class Monitor(Item):
    kind: Literal['point'] = 'point'
    name: str = 'monitor'
    component: Literal['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'] = 'Ez'
    spectrum: SpectrumSettings = Field(default_factory=SpectrumSettings)
    use_global_monitor: bool = False
    inherit_apodization: bool = True
    time_downsample: int = Field(default=1, ge=1, le=10000)


class FieldMonitor(Monitor):
    kind: Literal['field'] = 'field'
    name: str = 'frequency monitor'
    normal: Literal['x', 'y', 'z'] = 'x'
    size: tuple[float, float, float] = (0, 1, 1)
    downsample: int = Field(default=1, ge=1, le=32)
    downsample_xyz: tuple[int, int, int] | None = None
    record_fields: tuple[Literal['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'], ...] = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')
    record_poynting: tuple[Literal['x', 'y', 'z'], ...] = ('x', 'y', 'z')
    record_flux: bool = True
    dft_precision: Literal['field', 'float64'] = 'field'
    spatial_interpolation: Literal['specified', 'nearest'] = 'specified'
    spectrum: SpectrumSettings = Field(default_factory=lambda: SpectrumSettings(sampling='frequency', apodization='none'))

    @model_validator(mode='after')
    def valid_plane(self):
        axis = 'xyz'.index(self.normal)
        if self.size[axis] != 0 or any(v <= 0 for i, v in enumerate(self.size) if i != axis):
            raise ValueError('Frequency monitor normal span must be zero and transverse spans positive.')
        if not self.use_global_monitor and self.spectrum.sampling == 'fft':
            raise ValueError('Distributed frequency monitors require explicit frequency samples, not FFT bins.')
        if not self.use_global_monitor and self.spectrum.apodization == 'hann':
            raise ValueError('Distributed monitors support none/start/end/full apodization.')
        return self

    @property
    def required_fields(self):
        """models.py:693-699: recorded components plus the tangential pairs of Poynting and flux."""
        fields = set(self.record_fields)
        axes = set(self.record_poynting) | ({self.normal} if self.record_flux else set())
        for axis in axes:
            for component in 'xyz':
                if component != axis:
                    fields.update(('E'+component, 'H'+component))
        return tuple(c for c in ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz') if c in fields)


def default_materials():
    return [Material(name='Air', index=1, color='#b7c5d7'),
            Material(name='SiO2 (constant n)', index=1.444, color='#80c4d7'),
            Material(name='SiN (constant n)', index=2, color='#69a1e8'),
            Material(name='Si (constant n)', index=3.48, color='#cc8bdc')]


class Project(Model):
    schema_version: Literal[1] = 1
    name: str = Field(default='Untitled', min_length=1, max_length=120)
    region: Region = Field(default_factory=Region)
    limits: dict[str, Any] | None = None
    materials: list[Material] = Field(default_factory=default_materials, min_length=1)
    structures: list[Structure] = Field(default_factory=list)
    sources: list[Source] = Field(default_factory=list)
    global_source: SourceTimeSettings | None = Field(default_factory=SourceTimeSettings)
    monitors: list[FieldMonitor | Monitor] = Field(default_factory=list)
    global_monitor: SpectrumSettings = Field(default_factory=lambda: SpectrumSettings(sampling='frequency', apodization='none'))
    import_provenance: dict[str, Any] | None = None
    revision: int = Field(default=0, ge=0)
    content_sha256: str | None = Field(default=None, pattern=r'^[a-f0-9]{64}$')

    @model_serializer(mode='wrap')
    def _omit_default_limits(self, handler):
        data = handler(self)
        if not self.limits:
            data.pop('limits', None)
        return data

    @model_validator(mode='after')
    def valid_scene(self):
        names = [m.name for m in self.materials]
        if len(set(names)) != len(names):
            raise ValueError('Material names must be unique.')
        items = self.structures+self.sources+self.monitors
        ids = [x.id for x in items]
        if len(set(ids)) != len(ids):
            raise ValueError('Object IDs must be unique.')
        if any(s.material not in names for s in self.structures):
            raise ValueError('A structure references an unknown material.')
        r = self.region
        for obj in self.sources+self.monitors:
            if not obj.enabled:
                continue
            for axis in range(r.active_axes):
                spread = (isinstance(obj, Source) and obj.kind in ('plane', 'tfsf')) or isinstance(obj, FieldMonitor)
                half = obj.size[axis]/2 if spread else 0
                lower, upper = r.interior_bounds(axis)
                if isinstance(obj, Source) and obj.extend_through_pml and 'xyz'[axis] != obj.normal:
                    lower, upper = -r.actual_size[axis]/2, r.actual_size[axis]/2
                tolerance = 16*math.ulp(max(abs(lower), abs(upper), r.size[axis]))
                if obj.center[axis]-half < lower-tolerance or obj.center[axis]+half > upper+tolerance or (half == 0 and obj.center[axis] >= upper):
                    raise ValueError(f'{obj.name} must lie entirely inside the non-PML region.')
            if r.dimension == '2d' and obj.center[2] != 0:
                raise ValueError(f'{obj.name}: z must be 0 in a 2D simulation.')
        return self

    def resolved_source(self, source: Source):
        """models.py:777-784: pulse inheritance with local geometry, amplitude and phase."""
        if not source.use_global_source:
            return source
        if self.global_source is None:
            raise ValueError('Global source settings are unavailable.')
        return Source.model_validate({**source.model_dump(), **self.global_source.model_dump(), 'use_global_source': False})

    def resolved_monitor(self, monitor):
        """models.py:759-775 without source limits."""
        spec = (self.global_monitor if monitor.use_global_monitor else monitor.spectrum).model_dump()
        if monitor.use_global_monitor and not monitor.inherit_apodization:
            for key in ('apodization', 'apodization_center', 'apodization_time_width'):
                spec[key] = getattr(monitor.spectrum, key)
        if spec['use_source_limits']:
            bands = []
            for raw in self.sources:
                if not raw.enabled:
                    continue
                s = self.resolved_source(raw)
                if s.time_definition not in ('wavelength', 'frequency'):
                    raise ValueError(f'{monitor.name}: source limits require explicitly ranged wavelength/frequency sources.')
                bands.append((s.wavelength_start, s.wavelength_stop))
            if not bands:
                raise ValueError(f'{monitor.name}: source limits require an enabled ranged source.')
            spec.update(wavelength_start=min(a for a, _ in bands), wavelength_stop=max(b for _, b in bands))
        spec['use_source_limits'] = False
        return type(monitor).model_validate({**monitor.model_dump(), 'spectrum': spec, 'use_global_monitor': False})

    def save(self, path):
        from pathlib import Path
        Path(path).write_text(self.model_dump_json(indent=2), encoding='utf-8')

    @classmethod
    def load(cls, path):
        from pathlib import Path
        return cls.model_validate_json(Path(path).read_text(encoding='utf-8'))
