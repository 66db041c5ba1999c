"""The reference's recorded inverse-design run as a golden vector.

tests/golden/splitter.json is the thresholded 2D wavelength splitter of the reference's
adjoint example (examples/adjoint_inverse_design): 200 x 160 cells at 40 nm, 12-cell CPML,
FP64, 'yee' material sampling, a soft Ez sheet source with a 'standard' Gaussian pulse,
1200 steps.  tests/golden/splitter_fields.npz holds what the reference wrote for it
(examples/adjoint_inverse_design/reference.npz): the permittivity on the Ez nodes and the
spectral Ez of 173 x-normal lines (x = nodes 14 .. 186, y in +-2.6 um) at 1.31 and 1.55 um,
produced by DifferentiablePlaneSimulation (optimize.py:154-167).

The oracle reproduces both (CPU); the HIP engine reproduces the fields in FP64 through the
forward frequency planes and through DifferentiablePlaneSimulation (GPU).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import simulate
from oracle.scene import Scene

GOLDEN = Path(__file__).resolve().parent/'golden'
FREQUENCIES = [299792458.0/(w*1e-6) for w in (1.31, 1.55)]


def splitter():
    return json.loads((GOLDEN/'splitter.json').read_text())


def fixture():
    with np.load(GOLDEN/'splitter_fields.npz', allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def view_monitors(project, fields=('Ez',)):
    """The 173 observation lines of optimize.py:159-160, each recording only `fields`."""
    nodes = Scene(project).nodes[0]
    return [dict(id=f'view-{i}', name=f'view-{i}', kind='field', normal='x', center=(float(x), 0, 0), size=(0, 5.2, .04),
                 inherit_apodization=False, record_fields=list(fields), record_poynting=[], record_flux=False,
                 spectrum=dict(sampling='custom', custom_frequencies_hz=FREQUENCIES, apodization='none'))
            for i, x in enumerate(nodes[14:-14])]


def relative_l2(a, b):
    return float(np.linalg.norm(a-b)/np.linalg.norm(b))


def test_oracle_yee_voxelization_equals_the_reference_permittivity():
    """'yee' sampling of the 283 rectangles puts exactly the reference's permittivity on the Ez nodes."""
    eps = Scene(splitter()).voxelize()
    expected = fixture()['binary_epsilon']
    assert eps.shape == (200, 160, 1, 3) and eps.dtype == np.float64
    np.testing.assert_array_equal(eps[..., 2], expected)
    assert set(np.unique(expected)) == {1.0, 4.0}


def test_oracle_reproduces_the_reference_splitter_fields():
    project = splitter()
    project['monitors'] = view_monitors(project)
    out = simulate.run(project, record_frames=False)
    ref = fixture()
    got = np.stack([np.stack([plane.value[k, :, 0] for plane in out['planes']]) for k in range(2)])
    assert got.shape == ref['field_ez'].shape == (2, 173, 130)
    np.testing.assert_allclose(out['planes'][0].points[:, 1], ref['field_y_um'], rtol=0, atol=1e-12)
    np.testing.assert_allclose([p.points[0, 0] for p in out['planes']], ref['field_x_um'], rtol=0, atol=1e-12)
    assert relative_l2(got, ref['field_ez']) < 1e-12


@pytest.mark.gpu
def test_hip_frequency_planes_match_the_reference_splitter_fields(gpu):
    from torchfdtd_amd import Project, Simulation
    raw = splitter()
    raw['monitors'] = view_monitors(raw)
    raw['region'].update(backend='cuda', cuda_kernel='hip')
    project = Project.model_validate(raw)
    got = Simulation(project).run(cuda_graph_steps=50)
    assert got.summary['cuda_kernel'] == 'hip'
    planes = {m['id']: m for m in got.frequency_fields}
    fields = np.stack([np.stack([planes[f'view-{i}']['fields'][k, :, 0] for i in range(173)]) for k in range(2)])
    ref = fixture()['field_ez']
    assert relative_l2(fields, ref) < 1e-10


@pytest.mark.gpu
def test_hip_differentiable_planes_match_the_reference_splitter_fields(gpu):
    """The call of optimize.py:154-167: DifferentiablePlaneSimulation on the recorded binary permittivity."""
    import torch

    from torchfdtd_amd import Project
    from torchfdtd_amd.adjoint import DifferentiablePlaneSimulation
    raw = splitter()
    raw['monitors'] = [dict(m, record_fields=['Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'], record_flux=True,
                            record_poynting=['x', 'y', 'z']) for m in view_monitors(raw)]
    raw['region'].update(backend='cuda', cuda_kernel='hip')
    ref = fixture()
    epsilon = torch.as_tensor(ref['binary_epsilon'], device='cuda', dtype=torch.float64)
    model = DifferentiablePlaneSimulation(Project.model_validate(raw))
    with torch.no_grad():
        planes = model(epsilon, FREQUENCIES)
    fields = np.stack([np.stack([planes[f'view-{i}'].fields[k, :, 2].cpu().numpy() for i in range(173)]) for k in range(2)])
    assert relative_l2(fields, ref['field_ez']) < 1e-10
    np.testing.assert_allclose(planes['view-0'].points_um[:, 1].cpu().numpy(), ref['field_y_um'], rtol=0, atol=1e-12)
