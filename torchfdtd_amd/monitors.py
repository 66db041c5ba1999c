"""Point traces, point spectra and frequency-plane DFT monitors.

Planning (torchfdtd/field_monitors.py:14-79, torchfdtd/spectra.py:14-116) runs on the
host in FP64; the per-step plane update is the HIP kernel ``tfdtd_dft_update``
(the three launches of torchfdtd/cuda_monitors.py:65-155 fused into one): FP64
absolute-time phases exp(2 pi i f dt (n+1)) evaluated per block, eight-corner
trilinear samples in the field precision, accumulation in complex64 or complex128.
"""
from __future__ import annotations

import ctypes
from itertools import product

import numpy as np
import torch

from . import _native as N
from .planning import field_axes

C0 = 299792458.0


PLANE_FIELDS = ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')


def apodization_window(times, settings):
    t = np.asarray(times, dtype=np.float64)
    mode = settings.apodization
    if mode == 'none':
        return np.ones_like(t)
    if mode == 'hann':
        return np.hanning(len(t))
    d = t-settings.apodization_center
    if mode == 'start':
        d = np.minimum(d, 0)
    elif mode == 'end':
        d = np.maximum(d, 0)
    return np.exp(-2*np.log(2)*(d/settings.apodization_time_width)**2)


def frequency_samples(settings):
    if settings.sampling == 'custom':
        return np.asarray(settings.custom_frequencies_hz, dtype=np.float64)
    a, b, n = settings.wavelength_start, settings.wavelength_stop, settings.frequency_points
    if settings.sampling == 'fft':
        raise ValueError('FFT samples depend on the recorded time grid.')
    if settings.sampling == 'chebyshev':
        u = ((np.array([.5]) if n == 1 else (1-np.cos(np.pi*np.arange(n)/(n-1)))/2) if settings.chebyshev_nodes == 'lobatto'
             else (1-np.cos(np.pi*(np.arange(n)+.5)/n))/2)
        if settings.chebyshev_wavelength:
            return C0/((a+(b-a)*u)*1e-6)
        return C0/(b*1e-6)+(C0/(a*1e-6)-C0/(b*1e-6))*u
    if n == 1:
        return np.array([C0/(1e-6*(a+b)/2) if settings.sampling == 'wavelength' else .5*C0*1e6*(1/a+1/b)])
    if settings.sampling == 'wavelength':
        return C0/(np.linspace(a, b, n)*1e-6)
    return np.linspace(C0/(b*1e-6), C0/(a*1e-6), n)


def direct_transform(times, signal, frequencies, window=None):
    t, e, f = np.asarray(times), np.asarray(signal), np.asarray(frequencies)
    if t.ndim != 1 or e.shape != t.shape or f.ndim != 1 or len(t) < 2:
        raise ValueError('DFT requires a one-dimensional trace with at least two time samples.')
    dt = t[1]-t[0]
    if dt <= 0 or not np.allclose(np.diff(t), dt, rtol=1e-7, atol=0):
        raise ValueError('DFT requires a uniform increasing time grid.')
    if not np.all(np.isfinite(f)) or np.any(f <= 0) or np.any(f >= .5/dt):
        raise ValueError('DFT frequencies must be positive and below the temporal Nyquist limit.')
    weighted = e if window is None else e*np.asarray(window)
    values = np.empty(len(f), dtype=np.complex128)
    chunk = max(1, 1_000_000//len(t))
    for start in range(0, len(f), chunk):
        block = f[start:start+chunk]
        values[start:start+chunk] = (np.exp(2j*np.pi*block[:, None]*t)@weighted)*dt
    return values


def point_spectrum(times, signal, settings):
    """Spectrum of a point trace: direct DFT for explicit samples, scaled rFFT for 'fft' (spectra.py:95-116)."""
    window = apodization_window(times, settings)
    if len(times) < 2:
        return {'frequency_hz': np.array([]), 'value': np.array([], dtype=complex),
                'units': 'unavailable', 'transform': 'insufficient time samples'}
    dt = times[1]-times[0]
    if settings.sampling != 'fft':
        f = frequency_samples(settings)
        return {'frequency_hz': f, 'value': direct_transform(times, signal, f, window),
                'units': 'reduced field * s', 'transform': 'dt * sum(field * window * exp(+2pi i f t))'}
    if np.iscomplexobj(signal):
        f = np.fft.fftfreq(len(signal), dt)
        values = np.fft.fft(signal*window)/len(signal)
    else:
        f = np.fft.rfftfreq(len(signal), dt)
        values = np.fft.rfft(signal*window)/(len(signal)/2)
    keep = f > 0
    values = values*np.exp(-2j*np.pi*f*times[0])
    return {'frequency_hz': f[keep], 'value': values[keep], 'units': 'reduced field',
            'transform': 'FFT exp(-2pi i f t); scale 2/N real or 1/N complex; no window-gain correction'}


def plane_plan(region, monitor):
    """Quadrature points and areas of a plane (field_monitors.py:14-41)."""
    axes, widths = [], []
    normal = 'xyz'.index(monitor.normal)
    for a, nodes in enumerate(region.mesh_nodes):
        if region.dimension == '2d' and a == 2:
            axes.append(np.array([0.]))
            widths.append(np.array([1.]))
            continue
        if a == normal:
            position = monitor.center[a]
            if monitor.spatial_interpolation == 'nearest':
                position = nodes[np.argmin(abs(nodes-position))]
            axes.append(np.array([position]))
            widths.append(np.array([1.]))
            continue
        lo, hi = monitor.center[a]-monitor.size[a]/2, monitor.center[a]+monitor.size[a]/2
        edges = np.r_[lo, nodes[(nodes > lo+1e-12) & (nodes < hi-1e-12)], hi]
        stride = monitor.downsample_xyz[a] if monitor.downsample_xyz is not None else monitor.downsample
        edges = edges[np.r_[np.arange(0, len(edges)-1, stride), len(edges)-1]]
        axes.append((edges[:-1]+edges[1:])/2)
        widths.append(np.diff(edges)*1e-6)
    points = np.stack([v.ravel() for v in np.meshgrid(*axes, indexing='ij')], axis=1)
    area = np.prod(np.stack(np.broadcast_arrays(*[v.reshape(tuple(-1 if a == i else 1 for a in range(3)))
                                                  for i, v in enumerate(widths)])), axis=0).ravel()
    return dict(points_um=points, weights=area, shape=tuple(len(a) for a in axes), normal=normal)


def interpolation_map(region, component, points):
    """Trilinear corners: int64 flat indices ((x*Ny+y)*Nz+z)*3+c and weights, shape (corners, points)."""
    pairs = []
    for a, coords in enumerate(field_axes(region, component)):
        x = points[:, a]
        n = len(coords)
        if n == 1:
            pairs.append([(np.zeros(len(x), dtype=np.int64), np.ones(len(x)))])
            continue
        low = np.searchsorted(coords, x, side='right')-1
        high = low+1
        if region.boundaries.pair(a)[0].kind == 'periodic':
            length = region.actual_size[a]
            lower = coords[low % n]-np.where(low < 0, length, 0)
            upper = coords[high % n]+np.where(high >= n, length, 0)
            q = (x-lower)/(upper-lower)
            pairs.append([(low % n, 1-q), (high % n, q)])
        else:
            if np.any(x < coords[0]-1e-10) or np.any(x > coords[-1]+1e-10):
                raise ValueError('Monitor interpolation extends beyond the Yee grid.')
            low = np.clip(low, 0, n-2)
            high = low+1
            q = (x-coords[low])/(coords[high]-coords[low])
            pairs.append([(low, 1-q), (high, q)])
    ny, nz = region.shape[1:]
    c = 'xyz'.index(component[1].lower())
    indices, weights = [], []
    for terms in product(*pairs):
        indices.append(((terms[0][0]*ny+terms[1][0])*nz+terms[2][0])*3+c)
        weights.append(terms[0][1]*terms[1][1]*terms[2][1])
    return np.array(indices, dtype=np.int64), np.array(weights)


class FrequencyPlane:
    """One plane monitor with HIP accumulation (the state of field_monitors.py:121-164)."""

    def __init__(self, grid, monitor):
        r = grid.region
        self.grid, self.monitor = grid, monitor
        if monitor.spatial_interpolation != 'specified' and monitor.spatial_interpolation != 'nearest':
            raise ValueError('unknown spatial interpolation')
        self.plan = plane_plan(r, monitor)
        self.frequency = frequency_samples(monitor.spectrum)
        high = r.precision == 'float64' or monitor.dft_precision == 'float64'
        self.complex_dtype = torch.complex128 if high else torch.complex64
        self.components = monitor.required_fields
        device, real = grid.E.device, grid.dtype
        self.maps = []
        for comp in self.components:
            indices, weights = interpolation_map(r, comp, self.plan['points_um'])
            self.maps.append((torch.as_tensor(indices, device=device).contiguous(),
                              torch.as_tensor(weights, device=device, dtype=real).contiguous()))
        shape = (len(self.frequency), len(self.plan['weights']), len(self.components))
        self.value = torch.zeros(shape, device=device, dtype=self.complex_dtype)
        grid.memory_states.append(self.value)
        self.omega = torch.as_tensor((2j*np.pi*self.frequency*r.time_step).imag.copy(), device=device, dtype=torch.float64)
        self.half_phase = torch.as_tensor(np.exp(1j*np.pi*self.frequency*r.time_step), device=device, dtype=self.complex_dtype)
        times = np.arange(1, r.steps+1)*r.time_step
        mask = (np.arange(r.steps) % monitor.time_downsample == 0)*monitor.time_downsample
        self.windows = [torch.as_tensor(apodization_window(times+shift, monitor.spectrum)*r.time_step*mask, device=device, dtype=real)
                        for shift in (0, r.time_step/2)]
        # Blocked accumulation: the samples of `block` steps are staged and their phase products summed
        # in FP64 once per block, so the accumulators are read and written once per block, not per step.
        spec_bytes = 8 if self.complex_dtype == torch.complex128 else 4
        self.block = max(1, min(32, 65536//(4*spec_bytes*len(self.frequency))))
        self.stage = torch.zeros((self.block, shape[1]*shape[2]), device=device, dtype=real)
        self.flushed = torch.zeros(1, device=device, dtype=torch.long)
        grid.memory_states.append(self.flushed)
        self.counter = None

    def descriptor(self, counter, E=None, H=None, indices=None):
        """tfdtd_plane sampling `E`/`H` (the grid's fields by default) at `indices` (one index tensor per
        component, the maps' own by default)."""
        d = N.Plane()
        g = self.grid
        d.dtype = N.F64 if g.dtype == torch.float64 else N.F32
        d.spectral = 1 if self.complex_dtype == torch.complex128 else 0
        d.points = self.value.shape[1]
        d.frequencies = self.value.shape[0]
        d.components = len(self.components)
        d.electric = sum(c.startswith('E') for c in self.components)
        d.corners = self.maps[0][0].shape[0]
        d.time_downsample = self.monitor.time_downsample
        d.E, d.H = (g.E if E is None else E).data_ptr(), (g.H if H is None else H).data_ptr()
        for k, (own, weights) in enumerate(self.maps):
            d.indices[k] = (own if indices is None else indices[k]).data_ptr()
            d.weights[k] = weights.data_ptr()
        d.window_e, d.window_h = self.windows[0].data_ptr(), self.windows[1].data_ptr()
        d.omega, d.half, d.value = self.omega.data_ptr(), self.half_phase.data_ptr(), self.value.data_ptr()
        d.samples = self.windows[0].numel()
        d.counter = counter.data_ptr()
        d.block = self.block
        d.stage, d.flushed = self.stage.data_ptr(), self.flushed.data_ptr()
        self.counter = counter
        return d

    def flush(self):
        """Accumulate the staged steps of an unfinished block (before reading the accumulators)."""
        if self.counter is None or self.block <= 1:
            return
        d = self.descriptor(self.counter)
        N.check(N.library().tfdtd_dft_flush(ctypes.byref(d), 1, N.stream_handle(self.grid.device)), 'tfdtd_dft_flush')

    def result(self):
        self.flush()
        return plane_result(self.monitor, self.plan, self.frequency, self.components,
                            self.value.detach().cpu().numpy(), self.grid.region.dimension)


def plane_result(monitor, plan, frequency, components, values, dimension):
    """Fields, Poynting density and signed flux of a plane (field_monitors.py:167-183)."""
    values = np.asarray(values).astype(np.complex128)
    by_component = {c: values[..., i] for i, c in enumerate(components)}
    density = {}
    for axis in set(monitor.record_poynting) | ({monitor.normal} if monitor.record_flux else set()):
        i = 'xyz'.index(axis)
        b, c = 'xyz'[(i+1) % 3], 'xyz'[(i+2) % 3]
        density[axis] = .5*np.real(by_component['E'+b]*by_component['H'+c].conj()-by_component['E'+c]*by_component['H'+b].conj())
    flux = density[monitor.normal]@plan['weights'] if monitor.record_flux else None
    fields = values[..., [components.index(c) for c in monitor.record_fields]]
    poynting = np.stack([density[c] for c in monitor.record_poynting], axis=-1) if monitor.record_poynting else np.empty((*values.shape[:2], 0))
    return dict(id=monitor.id, name=monitor.name, frequency_hz=frequency, fields=fields, poynting=poynting, flux=flux, **plan,
                components=list(monitor.record_fields), poynting_components=list(monitor.record_poynting),
                accumulated_components=list(components), time_downsample=monitor.time_downsample,
                settings=monitor.model_dump(), flux_units='reduced E*H * s^2 * '+('m per invariant length' if dimension == '2d' else 'm^2'),
                field_units='reduced field * s', normal_axis=monitor.normal)


def normalize_flux(sample, reference, *, subtract_incident=False, min_reference_fraction=.01):
    """Plane power over a reference plane's positive power, per frequency (field_monitors.py:209-248).

    Both planes must come from matched runs (same run_signature: mesh, sources, steps), share
    frequencies, points, weights, normal and time downsampling, and hold unapodized fields and a
    recorded flux.  With `subtract_incident` the reference's tangential fields are subtracted from
    the sample's first (a reflection monitor), and the flux of the difference is used.  Ratios keep
    their sign; a frequency whose reference power is zero or below `min_reference_fraction` of the
    band's largest gives NaN, with the reason in `reasons`."""
    if not 0 < min_reference_fraction < 1:
        raise ValueError('min_reference_fraction lies strictly between 0 and 1.')
    for key in ('frequency_hz', 'points_um', 'weights'):
        a, b = np.asarray(sample[key]), np.asarray(reference[key])
        if a.shape != b.shape or not np.allclose(a, b, rtol=1e-10, atol=0):
            raise ValueError(f'The reference plane differs from the sample in {key}.')
    if sample['normal_axis'] != reference['normal_axis']:
        raise ValueError('The reference plane has another normal.')
    if sample['settings'].get('time_downsample', 1) != reference['settings'].get('time_downsample', 1):
        raise ValueError('The reference plane has another time downsampling.')
    if any(p['settings']['spectrum']['apodization'] != 'none' for p in (sample, reference)):
        raise ValueError('Flux normalization needs unapodized planes.')
    signature = sample.get('run_signature')
    if not signature or signature != reference.get('run_signature'):
        raise ValueError('The reference run differs from the sample run (mesh, sources, steps or region settings).')
    if sample['flux'] is None or reference['flux'] is None:
        raise ValueError('Flux normalization needs the recorded flux of both planes.')
    axis = 'xyz'.index(sample['normal_axis'])
    if subtract_incident:
        b, c = 'xyz'[(axis+1) % 3], 'xyz'[(axis+2) % 3]
        difference = {}
        for name in ('E'+b, 'E'+c, 'H'+b, 'H'+c):
            parts = []
            for plane in (sample, reference):
                names = plane.get('components', list(PLANE_FIELDS))
                if name not in names:
                    raise ValueError(f'Subtracting the incident field needs {name} stored in both planes.')
                parts.append(plane['fields'][..., names.index(name)])
            difference[name] = parts[0]-parts[1]
        density = .5*np.real(difference['E'+b]*difference['H'+c].conj()-difference['E'+c]*difference['H'+b].conj())
        flux = density@np.asarray(sample['weights'])
    else:
        flux = np.asarray(sample['flux'])
    power = np.abs(np.asarray(reference['flux']))
    valid = (power > 0) & (power > min_reference_fraction*power.max())
    ratio = np.full(len(flux), np.nan)
    ratio[valid] = flux[valid]/power[valid]
    reasons = np.array([None if ok else 'reference is zero' if p <= 0 else
                        f'reference below {min_reference_fraction:g} of its band peak: no supported ratio'
                        for ok, p in zip(valid, power)], dtype=object)
    return dict(frequency_hz=np.asarray(sample['frequency_hz']), ratio=ratio, valid=valid,
                subtract_incident=subtract_incident, reasons=reasons)
