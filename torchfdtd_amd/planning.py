"""Host-side planning of a run, in FP64 NumPy before anything reaches the GPU.

Mesh nodes (torchfdtd/mesh.py:74-105, uniform branch), CPML coefficients
(torchfdtd/boundaries.py:307-355), staircase voxelization (torchfdtd/solver.py:49-111,
torchfdtd/geometry.py:106-148), point indices and source boxes (solver.py:114-149),
waveforms (torchfdtd/waveforms.py:30-74) and soft source terms
(torchfdtd/injection.py:117-141).  The expressions are the reference's, in the same
order, so every coefficient rounds to the same FP32/FP64 value.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

C0 = 299792458.0
TAIL_INNER = math.sqrt(2*math.log(1e4))
TAIL_OUTER = 1.1*TAIL_INNER

# (derivative axis, input component, output component, curl sign): boundaries.py:21-23
CURL_TERMS = ((0, 2, 1, -1), (0, 1, 2, 1),
              (1, 2, 0, 1), (1, 0, 2, -1),
              (2, 1, 0, -1), (2, 0, 1, 1))


def mesh_nodes(region):
    """Node coordinates per axis in micrometres (uniform mesh; 2D z keeps its two bounding nodes)."""
    nodes = []
    for axis, n in enumerate(region.base_shape):
        if region.dimension == '2d' and axis == 2:
            a = np.array([-region.size[2]/2, region.size[2]/2])
        else:
            h = region.axis_steps[axis]
            a = np.arange(n+1)*h-n*h/2
        a.setflags(write=False)
        nodes.append(a)
    return tuple(nodes)


def real_dtype(region):
    return np.float64 if region.precision == 'float64' else np.float32


@dataclass
class CpmlSlab:
    """One auxiliary slab: target rows [lo, hi) of `axis` for the curl term differentiating `comp`."""
    forward: bool
    axis: int
    comp: int
    side: int
    lo: int
    hi: int
    b: np.ndarray
    c: np.ndarray
    inv_k: np.ndarray
    sigma: np.ndarray
    kappa: np.ndarray
    alpha: np.ndarray

    def psi_shape(self, shape):
        s = list(shape)
        s[self.axis] = self.hi-self.lo
        return tuple(s)


def cpml_slabs(region, dtype=None):
    """Per (forward, axis, comp) the CPML slabs of every PML face, in face order; b, c and inv_k in
    `dtype` (default: the field precision)."""
    shape = region.shape
    courant = region.rectangular_courant
    real = real_dtype(region) if dtype is None else dtype
    table = {}
    for forward in (False, True):
        for axis, comp, _, _ in CURL_TERMS:
            n = shape[axis]
            if n == 1:
                continue
            slabs = []
            for side, face in enumerate(region.boundaries.pair(axis)):
                if face.kind != 'pml':
                    continue
                layers = region.pml_layers(axis, side)
                base = 0 if side == 0 else n-layers
                # E derivatives are backward (rows 1..n-1), H derivatives forward (rows 0..n-2).
                first, last = (0, n-1) if forward else (1, n)
                lo, hi = max(base, first), min(base+layers, last)
                if lo >= hi:
                    continue
                row = np.arange(lo, hi)-base
                if side == 0:
                    depth = layers-row-(1 if forward else .5)
                else:
                    depth = row+(1 if forward else .5)
                rho = depth/(layers+1)
                sigma = face.sigma_scale*40*rho**face.polynomial/(layers+1)
                kappa = 1+(face.kappa-1)*rho**face.polynomial
                alpha = face.alpha*(1-rho)**face.alpha_polynomial
                decay = np.exp(-(sigma/kappa+alpha)*courant)
                denominator = sigma*kappa+alpha*kappa*kappa
                coupling = np.divide((decay-1)*sigma, denominator, out=np.zeros_like(sigma), where=denominator != 0)
                slabs.append(CpmlSlab(forward, axis, comp, side, lo, hi, decay.astype(real), coupling.astype(real),
                                      (1/kappa).astype(real), sigma, kappa, alpha))
            table[forward, axis, comp] = slabs
    return table


def psi_elements(region):
    """Number of CPML auxiliary values of a grid (both families)."""
    return sum(math.prod(s.psi_shape(region.shape)) for slabs in cpml_slabs(region).values() for s in slabs)


def pec_upper(region):
    """{axis: reference_step / last cell width} for upper PEC walls (boundaries.py:262-266)."""
    out = {}
    for axis, nodes in enumerate(region.mesh_nodes):
        if region.shape[axis] > 1 and region.boundaries.pair(axis)[1].kind == 'pec':
            out[axis] = float(region.reference_step/(nodes[-1]-nodes[-2]))
    return out


def periodic_axes(region):
    return tuple(a for a in range(3) if region.shape[a] > 1 and region.boundaries.pair(a)[0].kind == 'periodic')


# ------------------------------------------------------------------ geometry

def _axis_turn(axis, degrees):
    """Rotation about one axis (geometry.py:9-16); multiples of 90 degrees are exact."""
    if axis == 'none':
        return np.eye(3)
    degrees %= 360
    exact = {0: (1., 0.), 90: (0., 1.), 180: (-1., 0.), 270: (0., -1.)}
    c, s = exact.get(degrees) or (math.cos(math.radians(degrees)), math.sin(math.radians(degrees)))
    a = 'xyz'.index(axis)
    b, d = (a+1) % 3, (a+2) % 3
    m = np.eye(3)
    m[b, b] = m[d, d] = c
    m[b, d], m[d, b] = -s, s
    return m


def rotation_matrix(obj):
    """Local -> world rotation (geometry.py:19-26): the legacy z angle, unreduced, then each listed
    (axis, angle) about the fixed world axes in turn."""
    a = math.radians(obj.rotation)
    m = np.array([[math.cos(a), -math.sin(a), 0.], [math.sin(a), math.cos(a), 0.], [0., 0., 1.]])
    for axis, angle in zip(obj.rotation_axes, obj.rotation_angles):
        m = _axis_turn(axis, angle)@m
    return m


def _contour(x, y, contour):
    """(crossing parity, on-boundary) of points against one closed contour (geometry.py:83-93).
    The boundary carries a tolerance of 64 ulps of the contour's extent."""
    v = np.asarray(contour, dtype=float)
    extent = np.ptp(v, axis=0).max()
    tol = 64*np.finfo(float).eps*max(extent, 1e-300)
    shape = np.broadcast_shapes(np.shape(x), np.shape(y))
    parity, edge = np.zeros(shape, bool), np.zeros(shape, bool)
    for (ax, ay), (bx, by) in zip(v, np.roll(v, -1, axis=0)):
        side = (bx-ax)*(y-ay)-(by-ay)*(x-ax)
        near = abs(side) <= tol*max(abs(bx-ax), abs(by-ay))
        edge |= near & (x >= min(ax, bx)-tol) & (x <= max(ax, bx)+tol) & (y >= min(ay, by)-tol) & (y <= max(ay, by)+tol)
        if by != ay:
            parity ^= ((ay > y) != (by > y)) & (x < (bx-ax)*(y-ay)/(by-ay)+ax)
    return parity, edge


class Solid:
    """Membership and support box of one structure (geometry.py:104-148): a rectangle, sphere or
    ellipsoid, circle or elliptic cylinder, ring (elliptic, with an inner radius and an arc) or
    polygon with holes, extruded along its local w axis, placed by its rotation and centre."""

    def __init__(self, obj):
        self.obj, self.kind = obj, obj.kind
        self.matrix = rotation_matrix(obj)
        self.ellipse = obj.make_ellipsoid
        self.radii = (obj.radius, obj.radius_2 if self.ellipse else obj.radius, obj.radius_3 if self.ellipse else obj.radius)
        self.arc = (obj.theta_stop-obj.theta_start) % 360 or 360.
        z_only = all(axis in ('none', 'z') or angle % 360 == 0 for axis, angle in zip(obj.rotation_axes, obj.rotation_angles))
        rotation_free = (self.kind == 'sphere' and not self.ellipse) or (
            self.kind in ('circle', 'ring') and not self.ellipse and z_only and (self.kind != 'ring' or self.arc == 360))
        if rotation_free or (obj.rotation == 0 and not any(obj.rotation_angles)):
            self.frame = 'translate'
        elif not any(obj.rotation_angles):
            self.frame = 'legacy'          # z angle only: the historical two-term form
        else:
            self.frame = 'matrix'

    def support(self):
        """(centre, span) of a box holding the solid: exact for boxes, ellipsoids and cylinders,
        conservative for arcs and polygons."""
        m, obj = self.matrix, self.obj
        if self.kind == 'polygon':
            corners = np.array([[x, y, z] for x, y in obj.vertices for z in (-obj.size[2]/2, obj.size[2]/2)])@m.T+obj.center
            lo, hi = corners.min(axis=0), corners.max(axis=0)
            return tuple((lo+hi)/2), tuple(hi-lo)
        if self.kind == 'rectangle':
            half = np.abs(m)@(np.array(obj.size)/2)
        elif self.kind == 'sphere':
            half = np.sqrt((m*m)@(np.array(self.radii)**2))
        else:
            rx, ry, _ = self.radii
            half = np.sqrt((m[:, 0]*rx)**2+(m[:, 1]*ry)**2)+abs(m[:, 2])*obj.size[2]/2
        return tuple(obj.center), tuple(2*half)

    def local(self, x, y, z):
        d = (x-self.obj.center[0], y-self.obj.center[1], z-self.obj.center[2])
        if self.frame == 'translate':
            return d
        if self.frame == 'legacy':
            a = math.radians(self.obj.rotation)
            c, s = math.cos(a), math.sin(a)
            return c*d[0]+s*d[1], -s*d[0]+c*d[1], d[2]
        m = self.matrix
        # R^T d, skipping exact zeros so that quarter turns move coordinates without rounding.
        return tuple(sum(m[a, b]*d[a] for a in range(3) if m[a, b] != 0) for b in range(3))

    def contains(self, x, y, z):
        u, v, w = self.local(x, y, z)
        obj = self.obj
        slab = abs(w) <= obj.size[2]/2
        if self.kind == 'rectangle':
            return (abs(u) <= obj.size[0]/2) & (abs(v) <= obj.size[1]/2) & slab
        if self.kind == 'polygon':
            inside, edge = _contour(u, v, obj.vertices)
            for hole in obj.holes:
                parity, rim = _contour(u, v, hole)
                inside, edge = inside ^ parity, edge | rim
            return (inside | edge) & slab
        rx, ry, rz = self.radii
        if self.kind == 'sphere':
            return (u/rx)**2+(v/ry)**2+(w/rz)**2 <= 1 if self.ellipse else u*u+v*v+w*w <= rx*rx
        disc = ((u/rx)**2+(v/ry)**2 <= 1 if self.ellipse else u*u+v*v <= rx*rx) & slab
        if self.kind == 'ring':
            if obj.inner_radius > 0:
                disc &= ((u/obj.inner_radius)**2+(v/obj.inner_radius_2)**2 >= 1 if self.ellipse
                         else u*u+v*v >= obj.inner_radius**2)
            if self.arc < 360:
                polar = (np.degrees(np.arctan2(v, u))-obj.theta_start) % 360
                disc &= (polar <= self.arc+1e-12) | ((u == 0) & (v == 0))
        return disc


def object_bounds(obj):
    """Support box (centre, span) of a structure (geometry.py:136-148)."""
    return Solid(obj).support()


def contains(obj, x, y, z):
    """Analytic membership of a structure at world points (geometry.py:104-130)."""
    return Solid(obj).contains(x, y, z)


def field_axes(region, component):
    """Yee sample coordinates of a component: E on its own edge, H on its dual face (solver.py:37-46)."""
    axis = 'xyz'.index(component[1].lower())
    out = []
    for i, nodes in enumerate(region.mesh_nodes):
        half = (i == axis) if component[0] == 'E' else (i != axis)
        if region.dimension == '2d' and i == 2:
            out.append(np.array([0.]))
        else:
            out.append((nodes[:-1]+nodes[1:])/2 if half else nodes[:-1])
    return out


def voxelize(project, with_ownership=False):
    """Relative permittivity: (Nx,Ny,Nz) at cell centres ('cell') or (Nx,Ny,Nz,3) per E component
    ('yee') (solver.py:49-111).  with_ownership: (eps, ownership, counts) -- int32 index of the
    owning material in project.materials (-1 background) per sample, and per structure id the
    number of samples it set (the largest over the three components for 'yee')."""
    r = project.region
    if r.material_sampling == 'yee':
        parts = [_voxelize_at(project, field_axes(r, c)) for c in ('Ex', 'Ey', 'Ez')]
        eps = np.stack([p[0] for p in parts], axis=-1)
        if not with_ownership:
            return eps
        owner = np.stack([p[1] for p in parts], axis=-1)
        return eps, owner, {k: max(p[2][k] for p in parts) for k in parts[0][2]}
    axes = [(np.arange(n)+.5)*r.mesh-s/2 for n, s in zip(r.shape, r.actual_size)]
    if r.dimension == '2d':
        axes[2] = np.array([0.])
    eps, owner, counts = _voxelize_at(project, axes)
    return (eps, owner, counts) if with_ownership else eps


def _voxelize_at(project, axes):
    r = project.region
    shape = tuple(len(a) for a in axes)
    eps = np.full(shape, r.background_index**2, dtype=real_dtype(r))
    owner = np.full(shape, -1, dtype=np.int32)
    counts = {}
    number = {m.name: i for i, m in enumerate(project.materials)}
    materials = {m.name: m for m in project.materials}
    # Lower mesh order wins; at equal order the later structure wins.
    for obj in sorted(project.structures, key=lambda s: -s.mesh_order):
        if not obj.enabled:
            continue
        solid = Solid(obj)
        center, size = solid.support()
        # Only the samples inside the support box are tested (with a rounding margin).
        window = []
        for axis, c, span in zip(axes, center, size):
            margin = 128*np.finfo(float).eps*max(abs(c), span, 1e-300)
            window.append(slice(int(np.searchsorted(axis, c-span/2-margin, 'left')),
                                int(np.searchsorted(axis, c+span/2+margin, 'right'))))
        window = tuple(window)
        if any(w.start == w.stop for w in window):
            counts[obj.id] = 0
            continue
        mask = solid.contains(*np.meshgrid(*(axis[w] for axis, w in zip(axes, window)), indexing='ij', sparse=True))
        material = materials[obj.material]
        if material.model == 'tensor':
            raise ValueError(f'{material.name}: tensor materials are not run by this engine')
        eps[window][mask] = material.instantaneous_epsilon      # epsilon_inf for a dispersive material
        owner[window][mask] = number[obj.material]
        counts[obj.id] = int(np.count_nonzero(mask))
    return eps, owner, counts


def index_at(center, region, component=None):
    """Cell index of a point (solver.py:114-118)."""
    if component and region.material_sampling == 'yee':
        return tuple(int(np.argmin(abs(axis-c))) for axis, c in zip(field_axes(region, component), center))
    return tuple(0 if n == 1 else min(n-1, max(0, int(math.floor((v+s/2)/region.mesh))))
                 for v, s, n in zip(center, region.actual_size, region.shape))


def source_box(src, region):
    """Slices of the cells a soft source drives (solver.py:121-149)."""
    if src.kind == 'point':
        return tuple(slice(k, k+1) for k in index_at(src.center, region, src.component))
    parts = []
    if region.material_sampling == 'yee':
        for axis, c, span, n in zip(field_axes(region, src.component), src.center, src.size, region.shape):
            if n == 1:
                parts.append(slice(0, 1))
            elif span == 0:
                k = int(np.argmin(abs(axis-c)))
                parts.append(slice(k, k+1))
            else:
                idx = np.flatnonzero((axis >= c-span/2-1e-12) & (axis <= c+span/2+1e-12))
                if len(idx) == 0:
                    raise ValueError(f'{src.name} is smaller than one mesh cell.')
                parts.append(slice(int(idx[0]), int(idx[-1])+1))
        return tuple(parts)
    for c, span, total, n in zip(src.center, src.size, region.actual_size, region.shape):
        if n == 1:
            parts.append(slice(0, 1))
        elif span == 0:
            k = min(n-1, max(0, int((c+total/2)/region.mesh)))
            parts.append(slice(k, k+1))
        else:
            lo = max(0, int(math.ceil((c-span/2+total/2)/region.mesh-.5)))
            hi = min(n, int(math.floor((c+span/2+total/2)/region.mesh-.5))+1)
            if hi <= lo:
                raise ValueError(f'{src.name} is smaller than one mesh cell.')
            parts.append(slice(lo, hi))
    return tuple(parts)


@dataclass(frozen=True)
class PulseParameters:
    frequency_hz: float
    sigma_s: float
    offset_s: float
    frequency_span_hz: float
    chirped: bool


def pulse_parameters(source):
    """Centre frequency, Gaussian sigma, offset and chirp span (waveforms.py:30-46)."""
    frequency = C0/(source.wavelength*1e-6)
    span = source.chirp_bandwidth_hz if source.pulse == 'broadband' else 0.
    if source.time_definition in ('wavelength', 'frequency'):
        low = C0/(source.wavelength_stop*1e-6)
        high = C0/(source.wavelength_start*1e-6)
        frequency = (low+high)/2
        span = high-low
        length = (2 if source.optimize_for_short_pulse else 8)/(high+.01*low)
        sigma = length/(2*math.sqrt(math.log(2)))
        offset = TAIL_OUTER*sigma
        chirped = span > math.sqrt(math.log(2))/(math.pi*sigma)
    elif source.time_definition == 'standard':
        sigma = source.pulse_length/(2*math.sqrt(math.log(2)))
        offset = source.pulse_offset
        chirped = source.pulse == 'broadband'
    else:
        sigma = source.pulse_cycles/frequency
        offset = 4*sigma
        chirped = False
    return PulseParameters(frequency, sigma, offset, span, chirped)


def source_time_signal(source, times):
    """Soft-source drive samples at `times` (waveforms.py:57-74)."""
    times = np.asarray(times, dtype=np.float64)
    if source.pulse == 'sampled':
        signal = source.signal
        amplitude = np.interp(times, signal.time_s, signal.amplitude, left=0, right=0)
        phase = np.interp(times, signal.time_s, signal.phase_rad)+math.radians(source.phase)
        return source.amplitude*amplitude*np.sin(phase)
    p = pulse_parameters(source)
    x = times-p.offset_s
    u = x/p.sigma_s
    omega = 2*math.pi*p.frequency_hz
    phase = math.radians(source.phase)
    if source.pulse == 'continuous':
        envelope = 1-np.exp(-(times/p.sigma_s)**2)
        carrier = omega*times
    else:
        envelope = np.exp(-.5*u*u)
        if source.eliminate_discontinuities:
            h = TAIL_OUTER-TAIL_INNER
            distance = np.maximum(TAIL_OUTER-np.abs(u), 0)
            a = (3e-4-TAIL_INNER*1e-4*h)/(h*h)
            b = (TAIL_INNER*1e-4*h-2e-4)/(h*h*h)
            envelope = np.where(np.abs(u) <= TAIL_INNER, envelope, a*distance**2+b*distance**3)
        carrier = omega*times if source.time_definition == 'cycles' else -omega*x
        if p.chirped:
            inside = p.frequency_span_hz*p.sigma_s*(4*np.cos(np.pi*u/4)-2*np.pi)
            outside = -np.pi*p.frequency_span_hz*np.abs(x)
            carrier = carrier+np.where(np.abs(u) <= 2, inside, outside)
    return source.amplitude*envelope*np.sin(carrier+phase)


@dataclass
class SourceTerm:
    """A soft additive drive: field `family` component `comp` on `box` += values[step]."""
    family: str
    comp: int
    box: tuple
    values: np.ndarray


def source_terms(project, steps=None, dtype=None, only=None):
    """Soft point and plane sources (injection.py:117-141, Cartesian or polarized, real fields);
    values in `dtype` (default: the field precision); `only`: the terms of that one source."""
    r = project.region
    steps = r.steps if steps is None else steps
    out = []
    for raw in project.sources if only is None else [only]:
        s = project.resolved_source(raw)
        if not s.enabled:
            continue
        if s.kind == 'tfsf' or s.injection == 'oneway':
            raise ValueError(f'{s.name}: one-way and TFSF sources are not run by this engine')
        waveform = source_time_signal(s, np.arange(1, steps+1)*r.time_step+s.time_offset_steps*r.time_step)
        for component, weight in s.polarization_components:
            scalar = s.model_copy(update={'component': component, 'theta': None})
            box = source_box(scalar, r)
            out.append(SourceTerm(component[0], 'xyz'.index(component[1].lower()), box,
                                  (waveform*weight).astype(real_dtype(r) if dtype is None else dtype)))
    return out
