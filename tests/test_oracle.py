"""Pin the NumPy oracle before trusting it (CPU only).

* the reference-written golden runs (tests/golden/tiny_2d_*.npz, CPU FP32, copied
  from the reference's tests/fixtures/npz) are reproduced bit for bit;
* the closed-form discrete plane-wave eigenmode of the Yee scheme holds in FP64
  (pattern of tests/test_boundaries.py:32-61) -- a wrong sign, index or term fails it;
* the CPML absorbs a plane pulse (tests/test_boundaries.py:77-86 pattern);
* the plane DFT equals its definition (direct sum of sampled fields) and the
  trilinear maps reproduce affine fields exactly.
"""
import json
import math

import numpy as np
import pytest

from oracle import simulate
from oracle.monitors import FrequencyPlane, interpolation_map, plane_plan
from oracle.scene import Scene
from oracle.yee import Grid
from tests.conftest import GOLDEN


@pytest.mark.parametrize('name,steps', [('tiny_2d_complete', 40), ('tiny_2d_cancelled', 20)])
def test_oracle_reproduces_reference_golden_runs_bitwise(name, steps):
    data = np.load(GOLDEN/f'{name}.npz', allow_pickle=False)
    project = json.loads(str(data['project']))
    assert json.loads(str(data['summary']))['steps'] == steps
    out = simulate.run(project, steps=steps)
    np.testing.assert_array_equal(out['E'], data['E'])
    np.testing.assert_array_equal(out['H'], data['H'])
    np.testing.assert_array_equal(out['signals'], data['signals'])
    np.testing.assert_array_equal(out['epsilon'][:, :, 0], data['epsilon'])
    np.testing.assert_array_equal(out['times'], data['times'])
    for axis in range(3):
        np.testing.assert_array_equal(out['scene'].nodes[axis], data['mesh_'+'xyz'[axis]+'_um'])
    if name == 'tiny_2d_complete':
        np.testing.assert_array_equal(out['frames'], data['frames'])
        f, v = out['spectra'][0]
        np.testing.assert_array_equal(f, data['monitor_0_frequency_hz'])
        np.testing.assert_allclose(v, data['monitor_0_spectrum'], rtol=1e-12, atol=0)


def test_golden_project_file_matches_npz_project():
    data = np.load(GOLDEN/'tiny_2d_complete.npz', allow_pickle=False)
    stored = json.loads(str(data['project']))
    on_disk = json.loads((GOLDEN/'tiny_result.json').read_text())
    assert stored['region'] == on_disk['region'] and stored['sources'] == on_disk['sources']


def periodic_project(dimension):
    faces = {a+'_'+s: {'kind': 'periodic'} for a in 'xyz'[:2 if dimension == '2d' else 3] for s in ('min', 'max')}
    return {'region': {'dimension': dimension, 'size': (2.4, 2, 1.6), 'mesh': .1, 'precision': 'float64',
                       'boundaries': faces}}


@pytest.mark.parametrize('dimension', ['2d', '3d'])
def test_discrete_plane_wave_eigenmode(dimension):
    """omega = 2 asin(S |sin(q/2)|) for the exact N-cell period, after 67 steps, FP64 2e-13."""
    scene = Scene(periodic_project(dimension))
    g = Grid(scene)
    dims = 2 if dimension == '2d' else 3
    q = np.array([2*np.pi*m/n if n > 1 else 0 for m, n in zip((1, -2, 1), scene.shape)])
    k = np.sin(q/2)
    pol = np.array([0., 0., 1.]) if dims == 2 else np.cross(k, [0, 0, 1.])
    pol /= np.linalg.norm(pol)
    omega = 2*np.arcsin(g.courant*np.linalg.norm(k))
    coords = np.indices(scene.shape)
    base = np.exp(1j*np.einsum('i,i...->...', q, coords))
    electric = base[..., None]*pol*np.exp(.5j*q)
    magnetic = base[..., None]*(np.cross(k, pol)/np.linalg.norm(k))*np.exp(.5j*(q.sum()-q-omega))
    g.E[:] = electric.real
    g.H[:] = magnetic.real
    steps = 67
    for _ in range(steps):
        g.update_E()
        g.update_H()
    np.testing.assert_allclose(g.E, (electric*np.exp(-1j*steps*omega)).real, atol=2e-13, rtol=2e-13)
    np.testing.assert_allclose(g.H, (magnetic*np.exp(-1j*steps*omega)).real, atol=2e-13, rtol=2e-13)


def test_eigenmode_fails_for_a_planted_sign_error(monkeypatch):
    """The closed form detects a flipped curl sign (the check is not vacuous)."""
    import oracle.yee as yee
    flipped = list(yee.CURL_TERMS)
    flipped[3] = (1, 0, 2, 1)
    monkeypatch.setattr(yee, 'CURL_TERMS', tuple(flipped))
    with pytest.raises(AssertionError):
        test_discrete_plane_wave_eigenmode('3d')


def test_cpml_absorbs_a_periodic_sheet_pulse():
    """A plane pulse along x in a periodic-y strip leaves < 0.5 % residual after the PML round trip."""
    project = {'region': {'dimension': '2d', 'size': (8, 1, 1), 'mesh': .05, 'steps': 900, 'pml_cells': 14, 'precision': 'float64',
                          'boundaries': {'x_min': {'layers': 16, 'kappa': 3, 'alpha': .02, 'alpha_polynomial': 1},
                                         'x_max': {'layers': 20, 'kappa': 3, 'alpha': .02, 'alpha_polynomial': 1},
                                         'y_min': {'kind': 'periodic'}, 'y_max': {'kind': 'periodic'}}},
               'sources': [{'kind': 'plane', 'center': (-1, 0, 0), 'size': (0, 1, 0), 'wavelength': 1, 'pulse_cycles': 2}],
               'monitors': [{'center': (0, 0, 0)}]}
    out = simulate.run(project, record_frames=False)
    signal, times = out['signals'][:, 0], out['times']
    incident = np.max(abs(signal[times < 40e-15]))
    residual = np.max(abs(signal[times > 55e-15]))
    assert residual/incident < .005, residual/incident
    assert np.max(np.ptp(out['E'][..., 2], axis=1)) < 1e-12


def test_cpml_without_layers_reflects():
    """Control: the same pulse between periodic x walls is not absorbed."""
    project = {'region': {'dimension': '2d', 'size': (8, 1, 1), 'mesh': .05, 'steps': 900, 'precision': 'float64',
                          'boundaries': {'x_min': {'kind': 'periodic'}, 'x_max': {'kind': 'periodic'},
                                         'y_min': {'kind': 'periodic'}, 'y_max': {'kind': 'periodic'}}},
               'sources': [{'kind': 'plane', 'center': (-1, 0, 0), 'size': (0, 1, 0), 'wavelength': 1, 'pulse_cycles': 2}],
               'monitors': [{'center': (0, 0, 0)}]}
    out = simulate.run(project, record_frames=False)
    signal, times = out['signals'][:, 0], out['times']
    assert np.max(abs(signal[times > 55e-15])) > .2*np.max(abs(signal[times < 40e-15]))


def small_3d():
    return {'region': {'dimension': '3d', 'size': (2.4, 2.2, 2), 'mesh': .1, 'pml_cells': 3, 'steps': 30,
                       'precision': 'float64', 'material_sampling': 'yee'}}


def test_trilinear_maps_are_exact_for_affine_fields():
    scene = Scene(small_3d())
    m = {'kind': 'field', 'normal': 'x', 'center': (.13, .05, -.07), 'size': (0, 1.03, .9)}
    points, _ = plane_plan(scene, m)
    rng = np.random.default_rng(3)
    for comp in ('Ex', 'Hy', 'Ez'):
        a = rng.normal(size=4)
        axes = scene.field_axes(comp)
        X, Y, Z = np.meshgrid(*axes, indexing='ij')
        field = np.zeros((*scene.shape, 3))
        field[..., 'xyz'.index(comp[1].lower())] = a[0]+a[1]*X+a[2]*Y+a[3]*Z
        idx, w = interpolation_map(scene, comp, points)
        got = (field.reshape(-1)[idx]*w).sum(axis=0)
        np.testing.assert_allclose(got, a[0]+points@a[1:], rtol=1e-12, atol=1e-12)


def test_plane_dft_equals_direct_sum_of_samples():
    scene = Scene(small_3d())
    m = {'kind': 'field', 'normal': 'y', 'center': (0, .1, 0), 'size': (1, 0, .8), 'dft_precision': 'float64',
         'spectrum': {'sampling': 'frequency', 'frequency_points': 5, 'wavelength_start': 1, 'wavelength_stop': 1.4,
                      'apodization': 'end'}}
    plane = FrequencyPlane(scene, m)
    rng = np.random.default_rng(9)
    samples = []
    for n in range(25):
        E = rng.normal(size=(*scene.shape, 3))
        H = rng.normal(size=(*scene.shape, 3))
        samples.append(plane.sample(E, H, n))
        plane.update(E, H, n)
    samples = np.array(samples)       # (steps, points, comps)
    t = (np.arange(25)+1)*scene.dt
    ne = sum(c.startswith('E') for c in plane.components)
    shift = np.array([0.]*ne+[scene.dt/2]*(len(plane.components)-ne))
    kernel = np.exp(2j*np.pi*plane.frequency[:, None, None]*(t[None, :, None]+shift[None, None, :]))
    expected = np.einsum('fsc,spc->fpc', kernel, samples)
    np.testing.assert_allclose(plane.value, expected, rtol=1e-11, atol=1e-11*np.abs(expected).max())
