"""The HIP discrete adjoint against the oracle's restated reference adjoint and finite differences.

* gradients of a weighted point-signal objective equal the NumPy oracle's reverse sweep
  (a restatement of torchfdtd/differentiable.py:496-519, itself pinned by the dot-product
  identity and central differences in tests/test_oracle_adjoint.py): FP64 to 1e-9 and FP32
  to 2e-4 of the largest gradient, for CPML, periodic and PEC faces and both forward kernels;
* checkpointed replays give the gradient of the record-everything sweep bit for bit
  (schedule independence, tests/test_differentiable.py:172-183 pattern);
* plane spectra of DifferentiablePlaneSimulation equal the forward run's frequency plane,
  and their objective's gradient matches central differences of FP64 forward runs.
"""
import numpy as np
import pytest
import torch

import oracle.adjoint as OA
from oracle.scene import Scene
from torchfdtd_amd.adjoint import AdjointOptions, DifferentiablePlaneSimulation, DifferentiableSimulation
from torchfdtd_amd.models import BoundaryFace, FieldMonitor, Monitor, Project, Region, Source, SpectrumSettings

pytestmark = pytest.mark.gpu


def point_project(precision='float64', layout='pml', steps=36):
    r = Region(dimension='3d', size=(1.4, 1.2, 1.2), mesh=.1, pml_cells=3, precision=precision, steps=steps, backend='cuda')
    if layout == 'mixed':
        r.boundaries.y_min = BoundaryFace(kind='periodic')
        r.boundaries.y_max = BoundaryFace(kind='periodic')
        r.boundaries.z_max = BoundaryFace(kind='pec')
        r.boundaries.x_min = BoundaryFace(layers=4, kappa=2, alpha=.05, alpha_polynomial=1)
    return Project(region=r, sources=[Source(id='s', center=(-.2, 0, 0), component='Ez', pulse_length=3e-15,
                                             pulse_offset=6e-15, time_definition='standard')],
                   monitors=[Monitor(id='a', center=(.2, .1, 0), component='Ez'),
                             Monitor(id='b', center=(0, -.1, .1), component='Hy'),
                             Monitor(id='c', center=(.1, 0, -.1), component='Ex')])


def oracle_gradient(project, eps, weights):
    scene = Scene(project.model_dump())
    model = DifferentiableSimulation(project)
    observers = [('H' if m else 'E', flat) for m, flat in model.observers]
    J, gradient, signals = OA.run_with_gradient(scene, eps, project.region.steps, observers, weights, scene.source_terms())
    return J, gradient, signals


def engine_gradient(project, eps, weights, options=None):
    model = DifferentiableSimulation(project, options)
    epsilon = torch.tensor(eps, device='cuda', requires_grad=True)
    result = model(epsilon)
    J = (result.signals*torch.as_tensor(weights, device='cuda', dtype=epsilon.dtype)).sum()
    J.backward()
    return float(J), epsilon.grad.cpu().numpy(), result


def interior(project):
    r = project.region
    lo, hi = [], []
    for a in range(3):
        low, high = r.boundaries.pair(a)
        lo.append(r.pml_layers(a, 0) if low.kind == 'pml' else 0)
        hi.append(r.shape[a]-(r.pml_layers(a, 1) if high.kind == 'pml' else 0))
    return tuple(slice(a, b) for a, b in zip(lo, hi))


@pytest.mark.parametrize('precision,layout,per_component', [('float64', 'pml', False), ('float64', 'mixed', True),
                                                             ('float32', 'pml', False), ('float32', 'mixed', False)])
def test_gradient_matches_oracle_adjoint(gpu, precision, layout, per_component):
    project = point_project(precision, layout)
    rng = np.random.default_rng(3)
    shape = tuple(project.region.shape)+((3,) if per_component else ())
    eps = rng.uniform(1, 3, shape).astype(np.float64 if precision == 'float64' else np.float32)
    weights = rng.normal(size=(project.region.steps, 3))
    J_ref, g_ref, s_ref = oracle_gradient(project, eps, weights)
    J, g, result = engine_gradient(project, eps, weights)
    tol = 1e-9 if precision == 'float64' else 2e-4
    np.testing.assert_allclose(result.signals.detach().cpu().numpy(), s_ref, rtol=0, atol=tol*np.abs(s_ref).max())
    # dJ/deps over epsilon's whole shape, the CPML cells included (differentiable.py:503-505).
    scale = np.abs(g_ref).max()
    assert scale > 0
    np.testing.assert_allclose(g, g_ref, rtol=0, atol=tol*scale)
    box = interior(project)
    slabs = np.ones(g.shape, bool)
    slabs[box] = False
    assert np.abs(g_ref[slabs]).max() > 1e-3*scale      # the CPML cells carry gradient too


def test_checkpointed_replay_is_bitwise_identical(gpu):
    project = point_project('float32', steps=40)
    rng = np.random.default_rng(4)
    eps = rng.uniform(1, 3, tuple(project.region.shape)).astype(np.float32)
    weights = rng.normal(size=(40, 3))
    _, every, r0 = engine_gradient(project, eps, weights, AdjointOptions(checkpoint_interval=0))
    _, replayed, r1 = engine_gradient(project, eps, weights, AdjointOptions(checkpoint_interval=7, cuda_graph_steps=5))
    assert r0.report['replayed_steps'] == 0 and r1.report['replayed_steps'] == 40
    assert r1.report['checkpoint_interval'] == 7      # odd intervals run as given
    assert np.array_equal(every, replayed)
    box = ((4, 10), (5, 8), (4, 6))
    _, small, _ = engine_gradient(project, eps, weights, AdjointOptions(design_box=box))
    sl = tuple(slice(a, b) for a, b in box)
    outside = np.ones(small.shape, bool)
    outside[sl] = False
    assert np.array_equal(small[sl], every[sl]) and not small[outside].any()


def plane_project(precision='float64', steps=60):
    r = Region(dimension='3d', size=(1.6, 1.2, 1.2), mesh=.1, pml_cells=3, precision=precision, steps=steps, backend='cuda')
    spectrum = SpectrumSettings(sampling='wavelength', frequency_points=3, wavelength_start=1.5, wavelength_stop=1.6,
                                apodization='none')
    return Project(region=r, sources=[Source(id='s', center=(-.3, 0, 0), component='Ez', pulse_length=3e-15,
                                             pulse_offset=6e-15, time_definition='standard')],
                   monitors=[FieldMonitor(id='out', center=(.35, 0, 0), normal='x', size=(0, .5, .3),
                                          dft_precision='float64', spectrum=spectrum)])


def test_plane_spectra_match_the_forward_frequency_plane(gpu):
    from torchfdtd_amd.monitors import frequency_samples
    from torchfdtd_amd.solver import Simulation
    project = plane_project()
    rng = np.random.default_rng(8)
    eps = rng.uniform(1, 2.5, tuple(project.region.shape))
    forward = Simulation(project, epsilon=eps).run()
    model = DifferentiablePlaneSimulation(project)
    frequency = frequency_samples(project.monitors[0].spectrum)
    planes = model(torch.tensor(eps, device='cuda'), frequency)
    fields = planes['out'].fields.detach().cpu().numpy()
    values = forward.frequency_fields[0]['fields']
    assert values.shape == fields.shape
    np.testing.assert_allclose(fields, values, rtol=0, atol=1e-10*np.abs(values).max())


def test_plane_objective_gradient_matches_central_differences(gpu):
    from torchfdtd_amd.monitors import frequency_samples
    project = plane_project()
    rng = np.random.default_rng(9)
    eps = rng.uniform(1, 2.5, tuple(project.region.shape))
    frequency = frequency_samples(project.monitors[0].spectrum)
    model = DifferentiablePlaneSimulation(project)

    def objective(e):
        return (model(e, frequency)['out'].fields[..., 2].abs()**2).sum()
    epsilon = torch.tensor(eps, device='cuda', requires_grad=True)
    objective(epsilon).backward()
    gradient = epsilon.grad.cpu().numpy()
    assert np.abs(gradient).max() > 0
    h = 1e-5
    cells = [(8, 6, 5), (10, 5, 4), tuple(int(v) for v in np.unravel_index(np.argmax(np.abs(gradient)), gradient.shape))]
    for cell in cells:
        plus, minus = eps.copy(), eps.copy()
        plus[cell] += h
        minus[cell] -= h
        with torch.no_grad():
            fd = (objective(torch.tensor(plus, device='cuda'))-objective(torch.tensor(minus, device='cuda'))).item()/(2*h)
        assert gradient[cell] == pytest.approx(fd, rel=1e-5, abs=1e-7*np.abs(gradient).max())
