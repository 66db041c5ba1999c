"""Result files and plane post-processing (CPU).

* Result.load restores the reference-written NPZ runs and Result.save writes a file that loads back
  equal (the pattern of the reference's tests/test_project_compatibility.py:100-125);
* run_signature changes with the physics of a run and not with its placement or display settings
  (solver.py:493-508);
* normalize_flux (field_monitors.py:209-248) on planes whose ratios are worked out by hand.
"""
import numpy as np
import pytest

from tests.conftest import GOLDEN
from torchfdtd_amd.models import Project
from torchfdtd_amd.monitors import normalize_flux
from torchfdtd_amd.solver import Result, run_signature


@pytest.mark.parametrize('name', ['tiny_2d_complete', 'tiny_2d_cancelled'])
def test_reference_npz_loads_and_round_trips(name, tmp_path):
    path = GOLDEN/f'{name}.npz'
    result = Result.load(path)
    expected = Project.model_validate_json((GOLDEN/'tiny_result.json').read_text())
    assert result.project.model_dump() == expected.model_dump()
    summary = result.summary
    assert summary['backend'] == 'cpu' and summary['precision'] == 'float32'
    assert summary['requested_steps'] == 40 and isinstance(summary['warnings'], list)
    if name == 'tiny_2d_cancelled':
        assert summary['cancelled'] is True and summary['termination_reason'] == 'cancelled' and summary['steps'] == 20
    else:
        assert summary['cancelled'] is False and summary['steps'] == 40
    completed = summary['steps']
    assert result.times.shape == (completed,) and result.signals.shape == (completed, 1)
    assert result.frames.shape[0] == len(result.frame_steps) and result.frame_steps[-1] == completed
    assert np.isclose(result.times[1]-result.times[0], summary['dt_fs']*1e-15)
    assert result.electric.shape[:2] == result.epsilon.shape == tuple(result.project.region.shape[:2])
    with np.load(path, allow_pickle=False) as raw:
        for axis, nodes in zip('xyz', result.project.region.mesh_nodes):
            np.testing.assert_array_equal(raw[f'mesh_{axis}_um'], nodes)
        np.testing.assert_allclose(result.spectra[0]['value'], raw['monitor_0_spectrum'], rtol=1e-12, atol=0)
        np.testing.assert_array_equal(result.spectra[0]['frequency_hz'], raw['monitor_0_frequency_hz'])
        keys = set(raw.files)
    copy = tmp_path/'copy.npz'
    result.save(copy)
    again = Result.load(copy)
    assert again.summary == summary and again.project.model_dump() == result.project.model_dump()
    for attribute in ('frames', 'frame_steps', 'epsilon', 'signals', 'times', 'electric', 'magnetic'):
        np.testing.assert_array_equal(getattr(again, attribute), getattr(result, attribute))
    with np.load(copy, allow_pickle=False) as written:
        assert set(written.files) == keys


def test_run_signature_follows_physics_not_placement():
    base = Project.model_validate_json((GOLDEN/'tiny_result.json').read_text())
    signature = run_signature(base, 40)
    assert len(signature) == 64 and signature == run_signature(base.model_copy(deep=True), 40)
    placed = base.model_copy(deep=True)
    placed.region.backend, placed.region.snapshot_interval, placed.region.slice_position = 'cuda', 7, .1
    assert run_signature(placed, 40) == signature
    assert run_signature(base, 39) != signature
    moved = base.model_copy(deep=True)
    moved.sources[0].center = (moved.sources[0].center[0]+.05, *moved.sources[0].center[1:])
    assert run_signature(moved, 40) != signature
    finer = base.model_copy(deep=True)
    finer.region.courant_factor = .9
    assert run_signature(finer, 40) != signature


def plane(flux, fields=None, signature='s', apodization='none'):
    nf = len(flux)
    return dict(frequency_hz=np.linspace(1e14, 2e14, nf), points_um=np.zeros((2, 3)), weights=np.array([.5, .5]),
                normal_axis='x', settings=dict(time_downsample=1, spectrum=dict(apodization=apodization)),
                run_signature=signature, flux=np.asarray(flux, dtype=float), components=['Ey', 'Ez', 'Hy', 'Hz'],
                fields=fields if fields is not None else np.zeros((nf, 2, 4), complex))


def test_normalize_flux_ratios_guards_and_incident_subtraction():
    out = normalize_flux(plane([2., -1., 1e-4, 3.]), plane([4., 2., 1e-3, 0.]))
    np.testing.assert_array_equal(out['ratio'][:2], [.5, -.5])
    assert np.isnan(out['ratio'][2]) and np.isnan(out['ratio'][3])
    assert out['valid'].tolist() == [True, True, False, False]
    assert out['reasons'][3] == 'reference is zero' and 'below 0.01' in out['reasons'][2] and out['reasons'][0] is None
    # Incident subtraction: the sample holds reference + a reflected wave (Ey, Hz) = (a, -a) on both
    # points, so S_x = .5 Re(Ey Hz* - Ez Hy*) = -.5 |a|^2 and the weighted flux is -.5 |a|^2.
    a = np.array([1+1j, 2.])
    incident = np.zeros((2, 2, 4), complex)
    incident[..., 0], incident[..., 3] = 3., 3.
    sample_fields = incident.copy()
    sample_fields[..., 0] += a[:, None]
    sample_fields[..., 3] -= a[:, None]
    out = normalize_flux(plane([1., 1.], sample_fields), plane([4.5, 9.], incident), subtract_incident=True)
    np.testing.assert_allclose(out['ratio'], -.5*np.abs(a)**2/np.array([4.5, 9.]), rtol=1e-15)
    for bad in (plane([1., 1.], signature='t'), plane([1., 1.], apodization='hann'), dict(plane([1., 1.]), normal_axis='y')):
        with pytest.raises(ValueError):
            normalize_flux(plane([1., 1.]), bad)
    with pytest.raises(ValueError):
        normalize_flux(plane([1.]), plane([1.]), min_reference_fraction=1.)
