"""Oracle: point spectra and frequency-plane DFT monitors.

Restates torchfdtd/spectra.py:14-116 and torchfdtd/field_monitors.py:14-79,
121-183 (uniform meshes, real fields).  Test infrastructure only.
"""
from __future__ import annotations

from itertools import product

import numpy as np

C0 = 299792458.0

SPECTRUM_DEFAULT = dict(sampling='fft', chebyshev_wavelength=False, chebyshev_nodes='roots', use_source_limits=False,
                        custom_frequencies_hz=[], wavelength_start=1.3, wavelength_stop=1.8, frequency_points=101,
                        apodization='hann', apodization_center=20e-15, apodization_time_width=10e-15)
PLANE_SPECTRUM_DEFAULT = dict(SPECTRUM_DEFAULT, sampling='frequency', apodization='none')


def spectrum_settings(raw, plane=False):
    s = dict(PLANE_SPECTRUM_DEFAULT if plane else SPECTRUM_DEFAULT)
    s.update(raw or {})
    return s


def apodization_window(times, s):
    """spectra.py:14-27."""
    t = np.asarray(times, dtype=np.float64)
    mode = s['apodization']
    if mode == 'none':
        return np.ones_like(t)
    if mode == 'hann':
        return np.hanning(len(t))
    d = t-s['apodization_center']
    if mode == 'start':
        d = np.minimum(d, 0)
    elif mode == 'end':
        d = np.maximum(d, 0)
    return np.exp(-2*np.log(2)*(d/s['apodization_time_width'])**2)


def frequency_samples(s):
    """spectra.py:30-44 (frequency, wavelength and custom sampling)."""
    if s['sampling'] == 'custom':
        return np.asarray(s['custom_frequencies_hz'], dtype=np.float64)
    a, b, n = s['wavelength_start'], s['wavelength_stop'], s['frequency_points']
    if s['sampling'] not in ('frequency', 'wavelength'):
        raise NotImplementedError('oracle: frequency/wavelength/custom sampling only')
    if n == 1:
        return np.array([C0/(1e-6*(a+b)/2) if s['sampling'] == 'wavelength' else .5*C0*1e6*(1/a+1/b)])
    if s['sampling'] == 'wavelength':
        return C0/(np.linspace(a, b, n)*1e-6)
    return np.linspace(C0/(b*1e-6), C0/(a*1e-6), n)


def direct_transform(times, signal, frequencies, window=None):
    """dt * sum(x(t) w(t) exp(+2 pi i f t)): spectra.py:76-92."""
    t, e, f = np.asarray(times), np.asarray(signal), np.asarray(frequencies)
    dt = t[1]-t[0]
    weighted = e if window is None else e*np.asarray(window)
    return (np.exp(2j*np.pi*f[:, None]*t)@weighted)*dt


def point_spectrum(times, signal, s):
    """spectra.py:95-116."""
    window = apodization_window(times, s)
    dt = times[1]-times[0]
    if s['sampling'] != 'fft':
        f = frequency_samples(s)
        return f, direct_transform(times, signal, f, window)
    f = np.fft.rfftfreq(len(signal), dt)
    values = np.fft.rfft(signal*window)/(len(signal)/2)
    keep = f > 0
    values = values*np.exp(-2j*np.pi*f*times[0])
    return f[keep], values[keep]


PLANE_FIELDS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


def required_fields(m):
    """models.py:693-699."""
    fields = set(m.get('record_fields', PLANE_FIELDS))
    axes = set(m.get('record_poynting', ('x', 'y', 'z'))) | ({m.get('normal', 'x')} if m.get('record_flux', True) else set())
    for axis in axes:
        for c in 'xyz':
            if c != axis:
                fields.update(('E'+c, 'H'+c))
    return tuple(c for c in PLANE_FIELDS if c in fields)


def plane_plan(scene, m):
    """field_monitors.py:14-41 for 'specified' interpolation."""
    axes, widths = [], []
    normal = 'xyz'.index(m.get('normal', 'x'))
    center, size = m.get('center', (0, 0, 0)), m.get('size', (0, 1, 1))
    for a, nodes in enumerate(scene.nodes):
        if scene.dimension == '2d' and a == 2:
            axes.append(np.array([0.]))
            widths.append(np.array([1.]))
            continue
        if a == normal:
            axes.append(np.array([center[a]]))
            widths.append(np.array([1.]))
            continue
        lo, hi = center[a]-size[a]/2, center[a]+size[a]/2
        edges = np.r_[lo, nodes[(nodes > lo+1e-12) & (nodes < hi-1e-12)], hi]
        ds = m.get('downsample_xyz')
        stride = ds[a] if ds is not None else m.get('downsample', 1)
        edges = edges[np.r_[np.arange(0, len(edges)-1, stride), len(edges)-1]]
        axes.append((edges[:-1]+edges[1:])/2)
        widths.append(np.diff(edges)*1e-6)
    points = np.stack([v.ravel() for v in np.meshgrid(*axes, indexing='ij')], axis=1)
    area = np.prod(np.stack(np.broadcast_arrays(*[v.reshape(tuple(-1 if a == i else 1 for a in range(3)))
                                                  for i, v in enumerate(widths)])), axis=0).ravel()
    return points, area


def interpolation_map(scene, component, points):
    """Eight-corner trilinear maps with flat indices ((x*Ny+y)*Nz+z)*3+c: field_monitors.py:56-79 (non-cyclic)."""
    pairs = []
    for a, coords in enumerate(scene.field_axes(component)):
        x = points[:, a]
        n = len(coords)
        if n == 1:
            pairs.append([(np.zeros(len(x), dtype=int), np.ones(len(x)))])
            continue
        if scene.faces[a, 0].kind == 'periodic':
            length = scene.actual_size[a]
            low = np.searchsorted(coords, x, side='right')-1
            high = low+1
            lower = coords[low % n]-np.where(low < 0, length, 0)
            upper = coords[high % n]+np.where(high >= n, length, 0)
            q = (x-lower)/(upper-lower)
            pairs.append([(low % n, 1-q), (high % n, q)])
            continue
        low = np.searchsorted(coords, x, side='right')-1
        low = np.clip(low, 0, n-2)
        high = low+1
        q = (x-coords[low])/(coords[high]-coords[low])
        pairs.append([(low, 1-q), (high, q)])
    ny, nz = scene.shape[1:]
    c = 'xyz'.index(component[1].lower())
    indices, weights = [], []
    for terms in product(*pairs):
        indices.append(((terms[0][0]*ny+terms[1][0])*nz+terms[2][0])*3+c)
        weights.append(terms[0][1]*terms[1][1]*terms[2][1])
    return np.array(indices), np.array(weights)


class FrequencyPlane:
    """Online plane DFT, field_monitors.py:121-164, in NumPy (the reference's CPU branch)."""

    def __init__(self, scene, m):
        self.scene = scene
        self.m = m
        self.spectrum = spectrum_settings(m.get('spectrum'), plane=True)
        self.frequency = frequency_samples(self.spectrum)
        self.points, self.weights = plane_plan(scene, m)
        high = scene.precision == 'float64' or m.get('dft_precision', 'field') == 'float64'
        self.complex_dtype = np.complex128 if high else np.complex64
        self.components = required_fields(m)
        self.maps = [interpolation_map(scene, c, self.points) for c in self.components]
        self.maps = [(i, w.astype(scene.real)) for i, w in self.maps]
        self.value = np.zeros((len(self.frequency), len(self.weights), len(self.components)), self.complex_dtype)
        self.omega = 2j*np.pi*self.frequency*scene.dt
        self.half_phase = np.exp(1j*np.pi*self.frequency*scene.dt).astype(self.complex_dtype)
        times = np.arange(1, scene.steps+1)*scene.dt
        ds = m.get('time_downsample', 1)
        mask = (np.arange(scene.steps) % ds == 0)*ds
        self.windows = [(apodization_window(times+shift, self.spectrum)*scene.dt*mask).astype(scene.real)
                        for shift in (0, scene.dt/2)]

    def sample(self, E, H, n):
        fields = []
        for comp, (indices, weights) in zip(self.components, self.maps):
            family = int(comp.startswith('H'))
            flat = (H if family else E).reshape(-1)
            fields.append((flat[indices]*weights).sum(axis=0)*self.windows[family][n])
        return np.stack(fields, axis=-1)

    def update(self, E, H, n):
        values = self.sample(E, H, n)
        phase = np.exp(self.omega*(n+1)).astype(self.complex_dtype)
        ne = sum(c.startswith('E') for c in self.components)
        self.value[..., :ne] += phase[:, None, None]*values[None, :, :ne]
        self.value[..., ne:] += (phase*self.half_phase)[:, None, None]*values[None, :, ne:]
