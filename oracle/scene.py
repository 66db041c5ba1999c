"""Oracle: scene planning restated from the reference (uniform meshes only).

Input is a project dictionary in the reference's JSON layout (torchfdtd/models.py);
missing keys take the reference defaults.  Test infrastructure only.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

C0 = 299792458.0
TAIL_INNER = math.sqrt(2*math.log(1e4))
TAIL_OUTER = 1.1*TAIL_INNER

# (derivative axis, input component, output component, sign): torchfdtd/boundaries.py:21-23
CURL_TERMS = ((0, 2, 1, -1), (0, 1, 2, 1),
              (1, 2, 0, 1), (1, 0, 2, -1),
              (2, 1, 0, -1), (2, 0, 1, 1))

FACE_DEFAULT = dict(kind='pml', layers=None, sigma_scale=1.0, kappa=1.0, alpha=1e-8, polynomial=3.0, alpha_polynomial=0.0)
MATERIALS_DEFAULT = [dict(name='Air', index=1.0), dict(name='SiO2 (constant n)', index=1.444),
                     dict(name='SiN (constant n)', index=2.0), dict(name='Si (constant n)', index=3.48)]


def _face(d):
    out = dict(FACE_DEFAULT)
    out.update(d or {})
    return SimpleNamespace(**out)


class Scene:
    """Region geometry of torchfdtd/models.py:236-465 for uniform meshes."""

    def __init__(self, project):
        p = project
        r = p.get('region', {})
        self.project = p
        self.dimension = r.get('dimension', '2d')
        self.size = tuple(float(v) for v in r.get('size', (8, 6, 2)))
        self.mesh = float(r.get('mesh', 0.05))
        self.courant_factor = float(r.get('courant_factor', 0.99))
        self.steps = int(r.get('steps', 1000))
        self.pml_cells = int(r.get('pml_cells', 10))
        self.background_index = float(r.get('background_index', 1))
        self.material_sampling = r.get('material_sampling', 'cell')
        self.precision = r.get('precision', 'float32')
        self.snapshot_interval = int(r.get('snapshot_interval', 20))
        self.field = r.get('field', 'Ez')
        self.slice_axis = r.get('slice_axis', 'z')
        self.slice_position = float(r.get('slice_position', 0))
        if r.get('mesh_type', 'uniform') != 'uniform' or r.get('mesh_steps') is not None:
            raise NotImplementedError('oracle: uniform meshes only')
        b = r.get('boundaries', {})
        self.faces = {(a, s): _face(b.get('xyz'[a]+'_'+('min', 'max')[s])) for a in range(3) for s in range(2)}
        self.real = np.float64 if self.precision == 'float64' else np.float32
        active = 2 if self.dimension == '2d' else 3
        # models.py:383-394 base_shape
        def cells(span, step):
            ratio = span/step
            nearest = round(ratio)
            return max(1, nearest if abs(ratio-nearest) <= 8*math.ulp(ratio) else math.ceil(ratio))
        self.shape = tuple(cells(s, self.mesh) if (i < 2 or self.dimension == '3d') else 1 for i, s in enumerate(self.size))
        self.actual_size = tuple(n*self.mesh if n > 1 else self.size[i] for i, n in enumerate(self.shape))
        # mesh.py:86 uniform nodes; 2D z has the two bounding nodes (mesh.py:84-85)
        self.nodes = []
        for axis, n in enumerate(self.shape):
            if self.dimension == '2d' and axis == 2:
                self.nodes.append(np.array([-self.size[2]/2, self.size[2]/2]))
            else:
                self.nodes.append(np.arange(n+1)*self.mesh-n*self.mesh/2)
        # models.py:377-381 rectangular_courant, 403-412 time_step
        self.courant = self.courant_factor/math.sqrt(active)
        self.dt = self.courant_factor/math.sqrt(active)*self.mesh*1e-6/C0
        self.active = active

    def pml_layers(self, axis, side):
        face = self.faces[axis, side]
        return (face.layers or self.pml_cells) if face.kind == 'pml' else 0

    # ---------------------------------------------------------------- CPML
    def cpml_segments(self):
        """{(forward, axis, comp): [segment]} as torchfdtd/boundaries.py:307-355 (uniform branch)."""
        out = {}
        for forward in (False, True):
            for axis, comp, _, _ in CURL_TERMS:
                n = self.shape[axis]
                if n == 1:
                    continue
                segs = []
                for side in (0, 1):
                    face = self.faces[axis, side]
                    if face.kind != 'pml':
                        continue
                    L = self.pml_layers(axis, side)
                    base = 0 if side == 0 else n-L
                    start, stop = (0, n-1) if forward else (1, n)
                    lo, hi = max(base, start), min(base+L, stop)
                    if lo >= hi:
                        continue
                    t = np.arange(lo, hi)-base
                    if side == 0:
                        depth = L-t-(1 if forward else .5)
                    else:
                        depth = t+(1 if forward else .5)
                    rho = depth/(L+1)
                    sigma = face.sigma_scale*40*rho**face.polynomial/(L+1)
                    kappa = 1+(face.kappa-1)*rho**face.polynomial
                    alpha = face.alpha*(1-rho)**face.alpha_polynomial
                    b = np.exp(-(sigma/kappa+alpha)*self.courant)
                    den = sigma*kappa+alpha*kappa*kappa
                    c = np.divide((b-1)*sigma, den, out=np.zeros_like(sigma), where=den != 0)
                    shape = list(self.shape)
                    shape[axis] = hi-lo
                    cshape = [1, 1, 1]
                    cshape[axis] = hi-lo
                    segs.append(dict(lo=lo, hi=hi, start=start, psi=np.zeros(shape, self.real),
                                     b=b.reshape(cshape).astype(self.real), c=c.reshape(cshape).astype(self.real),
                                     inv_k=(1/kappa).reshape(cshape).astype(self.real)))
                out[forward, axis, comp] = segs
        return out

    # ---------------------------------------------------------------- materials
    def materials(self):
        return {m['name']: m for m in self.project.get('materials', MATERIALS_DEFAULT)}

    def field_axes(self, component):
        """Physical Yee coordinates of a component (torchfdtd/solver.py:37-46, no PMC)."""
        axis = 'xyz'.index(component[1].lower())
        out = []
        for i, nodes in enumerate(self.nodes):
            half = (i == axis) if component[0] == 'E' else (i != axis)
            if self.dimension == '2d' and i == 2:
                out.append(np.array([0.]))
            else:
                out.append((nodes[:-1]+nodes[1:])/2 if half else nodes[:-1])
        return out

    def voxelize(self, with_ownership=False):
        """Relative permittivity per cell ('cell') or per E component ('yee'): solver.py:49-111.
        with_ownership: (eps, ownership, counts) -- the int32 index of the owning material in the
        project's material list (-1 background) and, per structure id, the samples it set."""
        if self.material_sampling == 'yee':
            parts = [self._voxelize_at(self.field_axes(c)) for c in ('Ex', 'Ey', 'Ez')]
            eps = np.stack([p[0] for p in parts], axis=-1)
            if not with_ownership:
                return eps
            counts = {k: max(p[2].get(k, 0) for p in parts) for k in parts[0][2]}
            return eps, np.stack([p[1] for p in parts], axis=-1), counts
        axes = [(np.arange(n)+.5)*self.mesh-s/2 for n, s in zip(self.shape, self.actual_size)]
        if self.dimension == '2d':
            axes[2] = np.array([0.])
        eps, owner, counts = self._voxelize_at(axes)
        return (eps, owner, counts) if with_ownership else eps

    def _voxelize_at(self, axes):
        shape = tuple(len(a) for a in axes)
        eps = np.full(shape, self.background_index**2, dtype=self.real)
        owner = np.full(shape, -1, dtype=np.int32)
        counts = {}
        listed = self.project.get('materials', MATERIALS_DEFAULT)
        index = {m['name']: i for i, m in enumerate(listed)}
        mats = self.materials()
        structures = [_structure(s) for s in self.project.get('structures', [])]
        for obj in sorted(structures, key=lambda s: -s.mesh_order):
            if not obj.enabled:
                continue
            center, span = _bounds(obj)
            slices = []
            for axis, c, w in zip(axes, center, span):
                tol = 128*np.finfo(float).eps*max(abs(c), w, 1e-300)
                slices.append(slice(int(np.searchsorted(axis, c-w/2-tol, 'left')),
                                    int(np.searchsorted(axis, c+w/2+tol, 'right'))))
            slices = tuple(slices)
            if any(s.start == s.stop for s in slices):
                counts[obj.id] = 0
                continue
            xyz = np.meshgrid(*(a[s] for a, s in zip(axes, slices)), indexing='ij', sparse=True)
            mask = _contains(obj, *xyz)
            m = mats[obj.material]
            # models.py:156-159: a dispersive material's instantaneous permittivity is epsilon_inf.
            dispersive = m.get('model', 'dielectric') in ('drude', 'lorentz', 'multipole')
            eps[slices][mask] = m.get('epsilon_inf', 1.) if dispersive else m.get('index', 1.5)**2
            owner[slices][mask] = index[obj.material]
            counts[obj.id] = int(np.count_nonzero(mask))
        return eps, owner, counts

    # ---------------------------------------------------------------- sources
    def index_at(self, center, component=None):
        """solver.py:114-118."""
        if component and self.material_sampling == 'yee':
            return tuple(int(np.argmin(abs(a-c))) for a, c in zip(self.field_axes(component), center))
        return tuple(0 if n == 1 else min(n-1, max(0, int(math.floor((v+s/2)/self.mesh))))
                     for v, s, n in zip(center, self.actual_size, self.shape))

    def source_slice(self, src):
        """solver.py:121-149 (point and soft plane sources)."""
        if src['kind'] == 'point':
            return self.index_at(src['center'], src['component'])
        parts = []
        if self.material_sampling == 'yee':
            for axis, c, span, n in zip(self.field_axes(src['component']), src['center'], src['size'], self.shape):
                if n == 1:
                    parts.append(slice(0, 1))
                elif span == 0:
                    k = int(np.argmin(abs(axis-c)))
                    parts.append(slice(k, k+1))
                else:
                    idx = np.flatnonzero((axis >= c-span/2-1e-12) & (axis <= c+span/2+1e-12))
                    parts.append(slice(int(idx[0]), int(idx[-1])+1))
            return tuple(parts)
        for c, span, total, n in zip(src['center'], src['size'], self.actual_size, self.shape):
            if n == 1:
                parts.append(slice(0, 1))
            elif span == 0:
                k = min(n-1, max(0, int((c+total/2)/self.mesh)))
                parts.append(slice(k, k+1))
            else:
                lo = max(0, int(math.ceil((c-span/2+total/2)/self.mesh-.5)))
                hi = min(n, int(math.floor((c+span/2+total/2)/self.mesh-.5))+1)
                parts.append(slice(lo, hi))
        return tuple(parts)

    def source_terms(self):
        """[(component, loc, values in field dtype)] as injection.py:117-141 (soft sources, Cartesian)."""
        out = []
        for raw in self.project.get('sources', []):
            s = _source(raw)
            if not s['enabled']:
                continue
            if s['kind'] != 'point' and s['kind'] != 'plane':
                raise NotImplementedError('oracle: point and soft plane sources only')
            offset = .5 if s['component'].startswith('H') else 0.
            times = np.arange(1, self.steps+1)*self.dt+offset*self.dt
            waveform = source_time_signal(s, times)
            out.append((s['component'], self.source_slice(s), (waveform*1.).astype(self.real)))
        return out


SOURCE_DEFAULT = dict(wavelength=1.55, pulse='gaussian', pulse_cycles=3.0, time_definition='cycles', pulse_length=20e-15,
                      pulse_offset=50e-15, wavelength_start=1.3, wavelength_stop=1.8, optimize_for_short_pulse=True,
                      eliminate_discontinuities=False, chirp_bandwidth_hz=100e12, center=(0, 0, 0), enabled=True,
                      kind='point', size=(0, 2, 0), component='Ez', amplitude=1.0, phase=0.0)


def _source(raw):
    s = dict(SOURCE_DEFAULT)
    s.update({k: v for k, v in raw.items() if v is not None or k not in s})
    if s.get('theta') is not None or s.get('injection', 'soft') != 'soft' or s.get('use_global_source'):
        raise NotImplementedError('oracle: Cartesian soft sources without inheritance only')
    return s


def pulse_parameters(s):
    """waveforms.py:30-46."""
    frequency = C0/(s['wavelength']*1e-6)
    span = s['chirp_bandwidth_hz'] if s['pulse'] == 'broadband' else 0.
    if s['time_definition'] in ('wavelength', 'frequency'):
        low = C0/(s['wavelength_stop']*1e-6)
        high = C0/(s['wavelength_start']*1e-6)
        frequency = (low+high)/2
        span = high-low
        length = (2 if s['optimize_for_short_pulse'] else 8)/(high+.01*low)
        sigma = length/(2*math.sqrt(math.log(2)))
        offset = TAIL_OUTER*sigma
        chirped = span > math.sqrt(math.log(2))/(math.pi*sigma)
    elif s['time_definition'] == 'standard':
        sigma = s['pulse_length']/(2*math.sqrt(math.log(2)))
        offset = s['pulse_offset']
        chirped = s['pulse'] == 'broadband'
    else:
        sigma = s['pulse_cycles']/frequency
        offset = 4*sigma
        chirped = False
    return frequency, sigma, offset, span, chirped


def gaussian_envelope(u, taper):
    """waveforms.py:45-54 gaussian_envelope: exp(-u^2/2), with taper the tail beyond TAIL_INNER is the cubic
    a d^2 + b d^3 in d = TAIL_OUTER - |u| that meets the Gaussian's 1e-4 value and slope at TAIL_INNER and
    reaches 0 with zero slope at TAIL_OUTER."""
    u = np.asarray(u, dtype=np.float64)
    envelope = np.exp(-.5*u*u)
    if taper:
        h = TAIL_OUTER-TAIL_INNER
        distance = np.maximum(TAIL_OUTER-np.abs(u), 0)
        a = (3e-4-TAIL_INNER*1e-4*h)/(h*h)
        b = (TAIL_INNER*1e-4*h-2e-4)/(h*h*h)
        envelope = np.where(np.abs(u) <= TAIL_INNER, envelope, a*distance**2+b*distance**3)
    return envelope


def source_time_signal(s, times):
    """waveforms.py:57-74 for gaussian, broadband and continuous pulses."""
    times = np.asarray(times, dtype=np.float64)
    if s['pulse'] == 'sampled':
        raise NotImplementedError('oracle: sampled pulses are not restated')
    f, sigma, offset, span, chirped = pulse_parameters(s)
    x = times-offset
    u = x/sigma
    omega = 2*math.pi*f
    phase = math.radians(s['phase'])
    if s['pulse'] == 'continuous':
        envelope = 1-np.exp(-(times/sigma)**2)
        carrier = omega*times
    else:
        envelope = gaussian_envelope(u, s['eliminate_discontinuities'])
        carrier = omega*times if s['time_definition'] == 'cycles' else -omega*x
        if chirped:
            inside = span*sigma*(4*np.cos(np.pi*u/4)-2*np.pi)
            outside = -np.pi*span*np.abs(x)
            carrier = carrier+np.where(np.abs(u) <= 2, inside, outside)
    return s['amplitude']*envelope*np.sin(carrier+phase)


STRUCTURE_DEFAULT = dict(kind='rectangle', size=(1, 1, 0.5), radius=0.5, inner_radius=0.3, center=(0, 0, 0), enabled=True,
                         material='SiN (constant n)', mesh_order=2, rotation=0, rotation_axes=('z', 'x', 'y'),
                         rotation_angles=(0, 0, 0), make_ellipsoid=False, radius_2=.5, radius_3=.5, inner_radius_2=.3,
                         theta_start=0, theta_stop=360, vertices=((-.5, -.5), (.5, -.5), (0, .5)), holes=(), id=None)


def _structure(raw):
    s = dict(STRUCTURE_DEFAULT)
    s.update(raw)
    s['angular_span'] = (s['theta_stop']-s['theta_start']) % 360 or 360.
    return SimpleNamespace(**s)


def _quarter_turn(axis, angle):
    """geometry.py:9-16: rotation about one axis, exact at multiples of 90 degrees."""
    if axis == 'none':
        return np.eye(3)
    angle %= 360
    c, s = {0: (1., 0.), 90: (0., 1.), 180: (-1., 0.), 270: (0., -1.)}.get(
        angle, (math.cos(math.radians(angle)), math.sin(math.radians(angle))))
    a = 'xyz'.index(axis)
    b, d = (a+1) % 3, (a+2) % 3
    r = np.eye(3)
    r[b, b] = r[d, d] = c
    r[b, d] = -s
    r[d, b] = s
    return r


def rotation_matrix(obj):
    """geometry.py:19-26: local -> world; the legacy z angle (unreduced) first, then the listed axes."""
    a = math.radians(obj.rotation)
    c, s = math.cos(a), math.sin(a)
    r = np.array([[c, -s, 0.], [s, c, 0.], [0., 0., 1.]])
    for axis, angle in zip(obj.rotation_axes, obj.rotation_angles):
        r = _quarter_turn(axis, angle)@r
    return r


def _radii(obj):
    return (obj.radius, obj.radius_2 if obj.make_ellipsoid else obj.radius,
            obj.radius_3 if obj.make_ellipsoid else obj.radius)


def _bounds(obj):
    """geometry.py:136-148: support box (centre, span)."""
    r = rotation_matrix(obj)
    center = np.asarray(obj.center, dtype=float)
    if obj.kind == 'polygon':
        points = np.array([[x, y, z] for x, y in obj.vertices for z in (-obj.size[2]/2, obj.size[2]/2)])@r.T+center
        low, high = points.min(axis=0), points.max(axis=0)
        return tuple((low+high)/2), tuple(high-low)
    if obj.kind == 'rectangle':
        half = np.abs(r)@(np.array(obj.size)/2)
    elif obj.kind == 'sphere':
        half = np.sqrt((r*r)@(np.array(_radii(obj))**2))
    else:
        rx, ry, _ = _radii(obj)
        half = np.sqrt((r[:, 0]*rx)**2+(r[:, 1]*ry)**2)+abs(r[:, 2])*obj.size[2]/2
    return tuple(center), tuple(2*half)


def _contour_parity(x, y, vertices):
    """geometry.py:83-93: even-odd crossing parity and closed-boundary membership of one contour."""
    shape = np.broadcast_shapes(np.shape(x), np.shape(y))
    inside, boundary = np.zeros(shape, bool), np.zeros(shape, bool)
    v = np.asarray(vertices, dtype=float)
    span = np.ptp(v, axis=0).max()
    tol = 64*np.finfo(float).eps*max(span, 1e-300)
    for (ax, ay), (bx, by) in zip(v, np.roll(v, -1, axis=0)):
        cross = (bx-ax)*(y-ay)-(by-ay)*(x-ax)
        boundary |= ((abs(cross) <= tol*max(abs(bx-ax), abs(by-ay))) & (x >= min(ax, bx)-tol) & (x <= max(ax, bx)+tol)
                     & (y >= min(ay, by)-tol) & (y <= max(ay, by)+tol))
        if by != ay:
            inside ^= ((ay > y) != (by > y)) & (x < (bx-ax)*(y-ay)/(by-ay)+ax)
    return inside, boundary


def _polygon_contains(x, y, vertices, holes):
    """geometry.py:96-101: even-odd over the outer contour and its holes; every boundary is material."""
    inside, boundary = _contour_parity(x, y, vertices)
    for hole in holes:
        parity, edge = _contour_parity(x, y, hole)
        inside ^= parity
        boundary |= edge
    return inside | boundary


def _contains(obj, x, y, z):
    """geometry.py:104-130: membership of the local coordinates (u, v, w) = R^T (p - centre)."""
    delta = [x-obj.center[0], y-obj.center[1], z-obj.center[2]]
    isotropic = not obj.make_ellipsoid
    only_z = all(axis in ('none', 'z') or angle % 360 == 0 for axis, angle in zip(obj.rotation_axes, obj.rotation_angles))
    invariant = ((obj.kind == 'sphere' and isotropic)
                 or (obj.kind in ('circle', 'ring') and isotropic and only_z and (obj.kind != 'ring' or obj.angular_span == 360)))
    if invariant or (obj.rotation == 0 and not any(obj.rotation_angles)):
        u, v, w = delta
    elif not any(obj.rotation_angles):
        a = math.radians(obj.rotation)
        c, s = math.cos(a), math.sin(a)
        u, v, w = c*delta[0]+s*delta[1], -s*delta[0]+c*delta[1], delta[2]
    else:
        m = rotation_matrix(obj)
        u, v, w = [sum(m[a, b]*delta[a] for a in range(3) if m[a, b] != 0) for b in range(3)]
    if obj.kind == 'rectangle':
        return (abs(u) <= obj.size[0]/2) & (abs(v) <= obj.size[1]/2) & (abs(w) <= obj.size[2]/2)
    if obj.kind == 'polygon':
        return _polygon_contains(u, v, obj.vertices, obj.holes) & (abs(w) <= obj.size[2]/2)
    rx, ry, rz = _radii(obj)
    if obj.kind == 'sphere':
        return (u/rx)**2+(v/ry)**2+(w/rz)**2 <= 1 if obj.make_ellipsoid else u*u+v*v+w*w <= rx*rx
    mask = ((u/rx)**2+(v/ry)**2 <= 1 if obj.make_ellipsoid else u*u+v*v <= rx*rx) & (abs(w) <= obj.size[2]/2)
    if obj.kind == 'ring':
        if obj.inner_radius > 0:
            if obj.make_ellipsoid:
                mask &= (u/obj.inner_radius)**2+(v/obj.inner_radius_2)**2 >= 1
            else:
                mask &= u*u+v*v >= obj.inner_radius**2
        if obj.angular_span < 360:
            angle = (np.degrees(np.arctan2(v, u))-obj.theta_start) % 360
            mask &= (angle <= obj.angular_span+1e-12) | ((u == 0) & (v == 0))
    return mask
