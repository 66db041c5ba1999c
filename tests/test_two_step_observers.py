"""Plane monitors and traces on the two-step pass: the pass also writes step n+1 on the x planes the
observers read to a snapshot, and after each pass the DFT and trace launches sample step n+1 from it
and step n+2 from the fields (the launches of two one-pass steps, in their order).  A grid with an
x-normal plane monitor (E and H, float64 accumulation) and E/H traces runs the same steps on the two-step
pass (eager and in graphs), the one-pass step and the split step: fields, CPML memory, traces and
spectra are equal bit for bit (the split and one-pass steps equal the oracle bit for bit in
test_hip_parity.py and test_configs.py), and the two-step run equals the oracle directly."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def observed_project(steps=80, trace_x=1.0, reflection=False, wide=False, scattered=False, tall=False):
    """x-normal plane near high x and two traces; `reflection` adds an x-normal plane near low x (two
    snapshot boxes), `wide` a y-normal plane along x beside it (two boxes), `tall` makes the grid 60 x 150
    cells in y and z, so that the one box also passes through interior tiles (at 40 x 30 every tile
    holds CPML cells), `scattered` a z-normal and a y-normal plane across each other (one box of more cells than a
    snapshot takes)."""
    from torchfdtd_amd import FieldMonitor, Monitor, Project, Region, Source, Structure
    from torchfdtd_amd.models import SpectrumSettings
    region = Region(dimension='3d', size=(3.2, 2.4, 6.0) if tall else (3.2, 1.6, 1.2), mesh=.04, pml_cells=8, steps=steps, backend='cuda')
    spectrum = SpectrumSettings(sampling='wavelength', frequency_points=8, wavelength_start=1.45, wavelength_stop=1.65,
                                apodization='none')
    monitors = [FieldMonitor(id='plane', center=(1.0, 0, 0), normal='x', size=(0, .4, .5), dft_precision='float64',
                             spectrum=spectrum),
                Monitor(id='ez', component='Ez', center=(trace_x, .1, 0)),
                Monitor(id='hy', component='Hy', center=(trace_x + .04, 0, .05))]
    if reflection:
        monitors.append(FieldMonitor(id='reflection', center=(-.9, 0, 0), normal='x', size=(0, .4, .5), spectrum=spectrum))
    if wide:
        monitors.append(FieldMonitor(id='side', center=(0, -.3, 0), normal='y', size=(2.0, 0, .5), spectrum=spectrum))
    if scattered:
        monitors += [FieldMonitor(id='floor', center=(0, 0, 0), normal='z', size=(2.0, .8, 0), spectrum=spectrum),
                     FieldMonitor(id='wall', center=(0, 0, 0), normal='y', size=(2.0, 0, .5), spectrum=spectrum)]
    return Project(region=region,
                   structures=[Structure(center=(0, 0, 0), size=(2.4, .4, .3), material='Si (constant n)')],
                   sources=[Source(center=(.8, 0, 0), component='Ez', wavelength=1, pulse_cycles=1)],
                   monitors=monitors)


def state(run):
    for plane in run.planes:
        plane.flush()
    torch.cuda.synchronize()
    g = run.grid
    return dict(E=g.E.cpu(), H=g.H.cpu(), traces=run.traces.cpu(), planes=[p.value.cpu() for p in run.planes],
                memory=[m.cpu() for m in g.psi_buffers.values()])


def assert_same(got, expected):
    for key in ('E', 'H', 'traces'):
        assert torch.equal(got[key], expected[key]), key
    for a, b in zip(got['planes'], expected['planes']):
        assert torch.equal(a, b)
    for a, b in zip(got['memory'], expected['memory']):
        assert torch.equal(a, b)


@pytest.mark.parametrize('layout', ['one box', 'reflection', 'wide', 'tall'])
def test_two_step_observers_equal_the_one_step_paths(gpu, layout):
    """One snapshot box (a plane and two traces near high x), two (a reflection plane near low x as
    well, or a y-normal plane along x beside the x-normal one); one box on a grid with interior
    tiles."""
    reflection = layout == 'reflection'
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.solver import Stepper
    project = observed_project(reflection=reflection, wide=layout == 'wide', tall=layout == 'tall')
    results = {}
    for path in ('split', 'one_pass', 'two_step'):
        run = Stepper(project, one_pass=path != 'split', two_step=path == 'two_step')
        assert run.path == path and 'two_step' in run.paths()
        run.advance(80)
        results[path] = state(run)
        if path == 'two_step':
            boxes = run.snapshot_boxes
            total = sum(math.prod(b - a for a, b in zip(lo, hi)) for lo, hi in boxes)
            assert len(boxes) == (1 if layout in ('one box', 'tall') else 2), boxes
            assert 0 < total <= run.grid.E.numel() // 3 // 8, boxes
            interior = run.march2_tables[0][2][0]   # workgroups of the interior variant
            assert (interior > 0) == (layout == 'tall'), interior
            assert run.snap_e.shape[0] == total
    reference = results['split']
    assert np.abs(reference['traces'].numpy()).max() > 0 and reference['planes'][0].abs().max() > 0
    assert_same(results['one_pass'], reference)
    assert_same(results['two_step'], reference)
    # The same in graphs of eight steps (two passes of two steps each per graph).
    run = Stepper(project, one_pass=True, two_step=True)
    graphs = StepGraphs(run.step, run.states(), 80, 8, pair=run.step_pair, prepare=run.refresh_materials, quad=run.quad)
    graphs.advance(80)
    assert graphs.kernel_steps['two_step'] == 80, graphs.kernel_steps
    assert_same(state(run), reference)


def test_observers_spread_over_many_cells_keep_the_one_step_paths(gpu):
    from torchfdtd_amd.solver import Stepper
    project = observed_project(scattered=True)
    run = Stepper(project, one_pass=True)
    boxes = run.observed_boxes()
    assert sum(math.prod(b - a for a, b in zip(lo, hi)) for lo, hi in boxes) > run.grid.E.numel() // 3 // 8
    assert run.paths() == ['split', 'one_pass'] and not run.two_step
    with pytest.raises(ValueError, match='two-step'):
        Stepper(project, one_pass=True, two_step=True)


def test_two_step_observers_match_the_oracle(gpu):
    """The same grid for 80 steps on the two-step pass against the oracle of the reference's CPU path:
    fields and traces bit for bit, the float64 plane to 1e-12 of its largest entry."""
    from oracle import simulate
    from torchfdtd_amd.solver import Stepper
    project = observed_project()
    run = Stepper(project, one_pass=True, two_step=True)
    run.advance(80)
    got = state(run)
    ref = simulate.run(project.model_dump(), native=True, record_frames=False)
    np.testing.assert_array_equal(got['E'].numpy(), ref['E'])
    np.testing.assert_array_equal(got['H'].numpy(), ref['H'])
    np.testing.assert_array_equal(got['traces'].numpy(), ref['signals'])
    (expected,), plane = ref['planes'], run.planes[0]
    names = list(expected.components)
    values = expected.value[..., [names.index(c) for c in plane.components]]
    scale = np.abs(values).max()
    assert scale > 0 and np.abs(ref['signals']).max() > 0
    np.testing.assert_allclose(got['planes'][0].numpy()/scale, values/scale, rtol=1e-12, atol=1e-12)


def test_launch_tuner_accepts_the_two_step_pass_with_observers(gpu):
    """tuning.tune_launch times the two-step pass on an observed grid and finds its fields, CPML
    memory, traces and spectra identical to the split step's."""
    from torchfdtd_amd.solver import Stepper
    from torchfdtd_amd.tuning import tune_launch
    run = Stepper(observed_project())
    assert run.path == 'split' and 'two_step' in run.paths()
    record = tune_launch(run, candidates=[('split', {}), ('one_pass', dict(tile_y=8, chunk_x=6)),
                                          ('two_step', dict(waves2=12, chunk_x2=12))])
    assert record['rejected'] == [] and len(record['median_ms']) == 3, record
