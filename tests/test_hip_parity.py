"""HIP half steps and the fused step against the NumPy oracle of the reference path.

The HIP library keeps the reference's operation order and rounds every product and
sum separately (no FMA contraction, denormals kept), so results are compared bit for
bit.  Patterns follow the reference's own tests: random fields and every CPML state
over 23 steps (tests/test_cuda_kernels.py:72-111, seed 615), subnormal preservation
(:149-170), graph vs eager and complete solves (tests/test_solver.py:60-75).
"""
import numpy as np
import pytest
import torch

from oracle import simulate
from oracle.scene import Scene
from oracle.yee import Grid as OracleGrid
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def region_for(layout, precision):
    from torchfdtd_amd import BoundaryFace, Region
    r = Region(dimension='2d' if layout == 'pml2d' else '3d', size=(3.2, 2.8, 2.4), mesh=.1, pml_cells=4,
               precision=precision, material_sampling='yee', backend='cuda')
    r.boundaries.x_min = BoundaryFace(layers=5, kappa=3, alpha=.03, alpha_polynomial=1)
    if layout == 'mixed':
        r.boundaries.y_min = BoundaryFace(kind='periodic')
        r.boundaries.y_max = BoundaryFace(kind='periodic')
    if layout == 'pec':
        r.boundaries.z_min = BoundaryFace(kind='pec')
        r.boundaries.z_max = BoundaryFace(kind='pec')
        r.boundaries.y_max = BoundaryFace(kind='pec')
        r.boundaries.y_min = BoundaryFace(kind='pec')
    if layout in ('xpec', 'xpec2'):
        # CPML on y and z only at the high x end (both x ends for xpec2): the slab-free x middle runs
        # to the grid's last plane, so z-face and y-boundary tiles of its last chunk end at nx.
        r.boundaries.x_max = BoundaryFace(kind='pec')
        if layout == 'xpec2':
            r.boundaries.x_min = BoundaryFace(kind='pec')
    return r


@pytest.mark.parametrize('precision', ['float32', 'float64'])
@pytest.mark.parametrize('layout', ['pml2d', 'pml3d', 'mixed', 'pec'])
def test_random_fields_and_every_cpml_state_bitwise(gpu, precision, layout):
    from torchfdtd_amd import Project
    from torchfdtd_amd.grid import HipYee, YeeGrid
    region = region_for(layout, precision)
    project = Project(region=region)
    grid = YeeGrid(region)
    scene = Scene(project.model_dump())
    real = scene.real
    rng = np.random.default_rng(615)
    E = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    H = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    inv = rng.uniform(.25, 1, (*scene.shape, 3)).astype(real)
    ref = OracleGrid(scene, inv.copy())
    ref.E[:], ref.H[:] = E, H
    grid.E.copy_(torch.as_tensor(E))
    grid.H.copy_(torch.as_tensor(H))
    grid.inverse_permittivity.copy_(torch.as_tensor(inv))
    kernel = HipYee(grid)
    for _ in range(23):
        ref.update_E()
        ref.update_H()
        kernel.update_E()
        kernel.update_H()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(grid.E.cpu().numpy(), ref.E)
    np.testing.assert_array_equal(grid.H.cpu().numpy(), ref.H)
    assert len(grid.memory_states) == len(ref.memory_states) > 0 or layout == 'pec'
    for got, expected in zip(grid.memory_states, ref.memory_states):
        np.testing.assert_array_equal(got.cpu().numpy(), expected)


@pytest.mark.parametrize('precision', ['float32', 'float64'])
@pytest.mark.parametrize('layout,tile_y,chunk_x,nz,components', [
    ('pml3d', 8, 7, 131, 3), ('pml3d', 16, 64, 200, 3), ('pec', 8, 5, 131, 3), ('pml3d', 16, 0, 131, 3),
    ('pml3d', 8, 9, 200, 3), ('pec', 16, 6, 200, 3), ('pml3d', 8, 6, 200, 'index'), ('pec', 16, 4, 200, 1),
    ('xpec', 8, 7, 131, 3), ('xpec2', 16, 6, 131, 1), ('pml3d', 8, 2, 131, 3), ('pml3d', 16, 3, 200, 1)])
def test_one_pass_step_matches_oracle_bitwise(gpu, precision, layout, tile_y, chunk_x, nz, components):
    """The one-pass E+H kernel (ping-pong buffers, LDS tiles with halo recompute, x chunks, CPML-free
    interior and CPML boundary launches) equals the reference update bit for bit, with 1/eps per
    Yee component, per cell, or per cell through the one-byte material index (seven values)."""
    from torchfdtd_amd import Project
    from torchfdtd_amd.solver import Stepper
    region = region_for(layout, precision)
    region.size = (3.2, 7.4, nz/10)     # 32 x 74 x nz cells: partial y and z tiles
    indexed = components == 'index'
    if indexed:
        components = 1
    if components == 1:
        region.material_sampling = 'cell'

    project = Project(region=region)
    run = Stepper(project, steps=40, one_pass=True, tile_y=tile_y, chunk_x=chunk_x)
    g = run.grid
    scene = Scene(project.model_dump())
    real = scene.real
    rng = np.random.default_rng(7)
    E = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    H = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    inv = rng.uniform(.25, 1, (*scene.shape, components)).astype(real)
    if indexed:
        inv = rng.choice(np.unique(inv)[:7], size=inv.shape)
    assert tuple(run.grid.inverse_permittivity.shape) == inv.shape
    ref = OracleGrid(scene, inv.copy())
    ref.E[:], ref.H[:] = E, H
    g.E.copy_(torch.as_tensor(E))
    g.H.copy_(torch.as_tensor(H))
    g.inverse_permittivity.copy_(torch.as_tensor(inv))
    for _ in range(11):
        run.step_pair()
        for _ in range(2):
            ref.update_E()
            ref.update_H()
    run.step()          # an odd step through the half-step kernels
    ref.update_E()
    ref.update_H()
    torch.cuda.synchronize()
    assert run.materials.count == (7 if indexed else 0)
    np.testing.assert_array_equal(g.E.cpu().numpy(), ref.E)
    np.testing.assert_array_equal(g.H.cpu().numpy(), ref.H)
    for got, expected in zip(g.memory_states, ref.memory_states):
        np.testing.assert_array_equal(got.cpu().numpy(), expected)


@pytest.mark.parametrize('layout,chunk_x2,nz,components,waves2', [
    ('pml3d', 7, 131, 3, 12), ('pml3d', 20, 200, 'index', 12), ('pec', 5, 131, 1, 12), ('pml3d', 0, 131, 1, 12),
    ('pec', 20, 200, 3, 12), ('pml3d', 3, 64, 'index', 12), ('pml3d', 7, 131, 3, 16), ('pml3d', 20, 200, 'index', 16),
    ('pec', 5, 131, 1, 16), ('pml3d', 4, 64, 'index', 4),
    ('xpec', 7, 131, 3, 12), ('xpec', 20, 131, 1, 12), ('xpec', 9, 200, 'index', 16), ('xpec2', 6, 131, 3, 12),
    ('xpec2', 20, 64, 1, 4)])
def test_two_step_pass_matches_oracle_bitwise(gpu, layout, chunk_x2, nz, components, waves2):
    """The two-step pass (step n+1 recomputed on widened tiles and kept on chip, E, H and both CPML
    memories double-buffered) equals the reference update bit for bit, starting from random fields
    and random CPML memory, over five passes of four steps, one pair and one single step; interior
    tiles of 12 wavefronts (9 rows), 16 (13 rows, the last one cut) and 4 (3 rows each)."""
    from torchfdtd_amd import Project
    from torchfdtd_amd.solver import Stepper
    region = region_for(layout, 'float32')
    region.size = (3.2, 7.4, nz/10)     # 32 x 74 x nz cells: partial y and z tiles
    indexed = components == 'index'
    if indexed:
        components = 1
    if components == 1:
        region.material_sampling = 'cell'
    project = Project(region=region)
    run = Stepper(project, steps=40, one_pass=True, two_step=True, chunk_x2=chunk_x2, waves2=waves2)
    g = run.grid
    scene = Scene(project.model_dump())
    real = scene.real
    rng = np.random.default_rng(11)
    E = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    H = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    inv = rng.uniform(.25, 1, (*scene.shape, components)).astype(real)
    if indexed:
        inv = rng.choice(np.unique(inv)[:5], size=inv.shape)
    ref = OracleGrid(scene, inv.copy())
    ref.E[:], ref.H[:] = E, H
    g.E.copy_(torch.as_tensor(E))
    g.H.copy_(torch.as_tensor(H))
    g.inverse_permittivity.copy_(torch.as_tensor(inv))
    assert len(g.memory_states) == len(ref.memory_states) > 0
    for got, expected in zip(g.memory_states, ref.memory_states):
        expected[...] = (rng.normal(size=expected.shape)*.01).astype(real)
        got.copy_(torch.as_tensor(expected))
    for _ in range(5):
        run.quad()
        for _ in range(4):
            ref.update_E()
            ref.update_H()
    run.step_pair()     # the one-step kernel continues from the two-step pass's state
    run.step()
    for _ in range(3):
        ref.update_E()
        ref.update_H()
    torch.cuda.synchronize()
    assert run.materials.count == (5 if indexed else 0)
    assert int(run.counter.item()) == 23
    np.testing.assert_array_equal(g.E.cpu().numpy(), ref.E)
    np.testing.assert_array_equal(g.H.cpu().numpy(), ref.H)
    for got, expected in zip(g.memory_states, ref.memory_states):
        np.testing.assert_array_equal(got.cpu().numpy(), expected)


def test_two_step_pass_with_sources_equals_one_step_pairs_under_graphs(gpu):
    """The cube workload at 96^3 (16-cell CPML, two-material 1/eps, Ez dipole): graph replays of the
    two-step pass (graphs of fours then a pair and a single step, a graph of one four, single steps)
    equal the one-step pairs bit for bit, source included."""
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.solver import Stepper
    from torchfdtd_amd.workloads import cube, cube_inverse_permittivity, cube_noise_mask
    mask = cube_noise_mask(96)
    out = []
    for two_step in (False, True):
        run = Stepper(cube(n=96, steps=60), one_pass=True, two_step=two_step)
        assert run.two_step == two_step
        run.grid.inverse_permittivity.copy_(cube_inverse_permittivity(mask)[..., None])
        graphs = StepGraphs(run.step, run.states(), 60, 23, warmup=1, pair=run.step_pair,
                            prepare=run.refresh_materials, quad=run.quad if two_step else None)
        for _ in range(2):
            graphs.replay(23)
        if two_step:
            graphs.replay(4)            # the graph of one four
        else:
            for _ in range(4):
                graphs.replay(1)
        graphs.replay(1)
        torch.cuda.synchronize()
        assert int(run.counter.item()) == 51
        out.append([t.clone() for t in (run.grid.E, run.grid.H, *run.grid.memory_states)])
    assert float(out[0][0].abs().max()) > 0
    for a, b in zip(*out):
        assert torch.equal(a, b)


def test_two_step_entry_points_agree_bitwise(gpu):
    """tfdtd_yee_step_march2 (descriptor in kernel arguments) and tfdtd_batch_march2 (device table)
    produce the same pass from the same state: fields and both CPML memories."""
    import ctypes
    from torchfdtd_amd import Project
    from torchfdtd_amd import _native as N
    from torchfdtd_amd.solver import Stepper
    region = region_for('pml3d', 'float32')
    region.size = (4.4, 7.4, 13.1)
    run = Stepper(Project(region=region), steps=40, one_pass=True, two_step=True)
    g = run.grid
    gen = torch.Generator(device=g.device).manual_seed(5)
    E0 = torch.randn(g.E.shape, device=g.device, generator=gen)*.02
    H0 = torch.randn(g.H.shape, device=g.device, generator=gen)*.02
    stream = N.stream_handle(g.device)
    outs = []
    for launch in (lambda: run.march2_step(0, stream),
                   lambda: N.check(run.lib.tfdtd_yee_step_march2(ctypes.byref(run.march2[0]), stream), 'march2')):
        g.E.copy_(E0)
        g.H.copy_(H0)
        for t in g.memory_states:
            t.fill_(.01)
        run.E_b.zero_(), run.H_b.zero_(), run.psi_e_b.zero_(), run.psi_h_b.zero_()
        launch()
        torch.cuda.synchronize()
        outs.append([t.clone() for t in (run.E_b, run.H_b, run.psi_e_b, run.psi_h_b)])
    assert float(outs[0][3].abs().max()) > 0
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_subnormal_fields_are_preserved(gpu):
    from torchfdtd_amd import Project, Region
    from torchfdtd_amd.grid import HipYee, YeeGrid
    region = Region(size=(2, 2, 1), mesh=.1, pml_cells=4, backend='cuda')
    grid = YeeGrid(region)
    ref = OracleGrid(Scene(Project(region=region).model_dump()))
    grid.E[10, 10, 0, 2] = 1e-38
    ref.E[10, 10, 0, 2] = np.float32(1e-38)
    kernel = HipYee(grid)
    for _ in range(20):
        ref.update_E()
        ref.update_H()
        kernel.update_E()
        kernel.update_H()
    assert np.count_nonzero(ref.H) > 0
    assert np.count_nonzero(np.abs(ref.H) < np.finfo(np.float32).tiny) > 0   # subnormals occur
    np.testing.assert_array_equal(grid.E.cpu().numpy(), ref.E)
    np.testing.assert_array_equal(grid.H.cpu().numpy(), ref.H)


@pytest.mark.parametrize('name', ['tiny_2d_complete', 'tiny_2d_cancelled'])
def test_reference_golden_run_bitwise(gpu, name):
    """The reference-written CPU FP32 run (tests/fixtures/npz) reproduced by the HIP step, bit for bit."""
    import json
    from torchfdtd_amd import Project, Simulation
    data = np.load(GOLDEN/f'{name}.npz', allow_pickle=False)
    project = Project.model_validate_json(str(data['project']))
    project.region.backend = 'cuda'
    project.region.cuda_kernel = 'hip'
    steps = json.loads(str(data['summary']))['steps']
    project.region.steps = steps if name == 'tiny_2d_complete' else project.region.steps
    if name == 'tiny_2d_cancelled':
        from threading import Event
        event = Event()
        result = Simulation(project).run(progress=lambda s: event.set() if s['step'] >= 20 else None, cancel=event)
        assert result.summary['steps'] == 20
    else:
        result = Simulation(project).run()
    np.testing.assert_array_equal(result.electric, data['E'])
    np.testing.assert_array_equal(result.magnetic, data['H'])
    np.testing.assert_array_equal(result.signals, data['signals'])
    np.testing.assert_array_equal(result.frames, data['frames'])
    np.testing.assert_array_equal(result.times, data['times'])
    np.testing.assert_allclose(result.spectra[0]['value'], data['monitor_0_spectrum'], rtol=1e-12, atol=0)


def waveguide_2d(steps=1000):
    """The 2D TM waveguide of the reference's CPU configuration: 160x120 cells, 10-cell CPML, Ez point source."""
    import math
    from torchfdtd_amd import Material, Monitor, Project, Region, Source, SpectrumSettings, Structure
    return Project(region=Region(dimension='2d', size=(8, 6, .05), mesh=.05, steps=steps, pml_cells=10,
                                 background_index=math.sqrt(2.085), backend='cuda', snapshot_interval=50),
                   materials=[Material(name='core', index=math.sqrt(12.11))],
                   structures=[Structure(size=(8, .65, 1), material='core')],
                   sources=[Source(center=(-2.5, 0, 0), component='Ez', pulse='broadband', time_definition='wavelength',
                                   wavelength_start=1.5, wavelength_stop=1.6)],
                   monitors=[Monitor(center=(2, 0, 0), spectrum=SpectrumSettings(
                       sampling='wavelength', frequency_points=32, wavelength_start=1.5, wavelength_stop=1.6,
                       apodization='none'))])


def test_2d_waveguide_1000_steps_matches_oracle(gpu):
    project = waveguide_2d()
    from torchfdtd_amd import Simulation
    got = Simulation(project).run()
    ref = simulate.run(project.model_dump())
    np.testing.assert_array_equal(got.electric, ref['E'])
    np.testing.assert_array_equal(got.magnetic, ref['H'])
    np.testing.assert_array_equal(got.signals, ref['signals'])
    f, v = ref['spectra'][0]
    np.testing.assert_array_equal(got.spectra[0]['frequency_hz'], f)
    np.testing.assert_allclose(got.spectra[0]['value'], v, rtol=1e-12, atol=1e-12*np.abs(v).max())
    assert np.abs(got.signals).max() > 1e-3


def test_graph_unrolled_and_eager_steps_are_bitwise_equal(gpu):
    from torchfdtd_amd import Simulation
    project = waveguide_2d(steps=120)
    eager = Simulation(project).run(cuda_graph=False)
    graph = Simulation(project).run(cuda_graph=True)
    unrolled = Simulation(project).run(cuda_graph=True, cuda_graph_steps=16)
    assert graph.summary['cuda_graph'] and unrolled.summary['cuda_graph_steps'] == 16
    for other in (graph, unrolled):
        np.testing.assert_array_equal(other.electric, eager.electric)
        np.testing.assert_array_equal(other.signals, eager.signals)
        np.testing.assert_array_equal(other.frames, eager.frames)


def test_3d_waveguide_matches_oracle(gpu):
    """The 3D waveguide configuration (160x120x40, 12-cell CPML, Ey dipole) over its first 60 steps."""
    import math
    from torchfdtd_amd import Material, Monitor, Project, Region, Simulation, Source, Structure
    project = Project(region=Region(dimension='3d', size=(8, 6, 2), mesh=.05, steps=60, pml_cells=12,
                                    background_index=math.sqrt(2.085), backend='cuda', snapshot_interval=30),
                      materials=[Material(name='core', index=math.sqrt(12.11))],
                      structures=[Structure(size=(8, .65, .4), material='core')],
                      sources=[Source(center=(-2.5, 0, 0), component='Ey', pulse='broadband', time_definition='wavelength',
                                      wavelength_start=1.5, wavelength_stop=1.6)],
                      monitors=[Monitor(center=(-2.4, 0, 0), component='Ey')])
    got = Simulation(project).run()
    ref = simulate.run(project.model_dump())
    np.testing.assert_array_equal(got.electric, ref['E'])
    np.testing.assert_array_equal(got.magnetic, ref['H'])
    np.testing.assert_array_equal(got.signals, ref['signals'])
    assert np.abs(got.signals).max() > 0


@pytest.mark.parametrize('dft_precision', ['field', 'float64'])
def test_plane_dft_matches_oracle(gpu, dft_precision):
    from torchfdtd_amd import FieldMonitor, Monitor, Project, Region, Simulation, Source, SpectrumSettings, Structure
    project = Project(region=Region(dimension='3d', size=(3.2, 2.8, 2.4), mesh=.1, pml_cells=4, steps=80, backend='cuda',
                                    material_sampling='yee'),
                      structures=[Structure(kind='sphere', radius=.4, material='SiO2 (constant n)')],
                      sources=[Source(center=(-.8, 0, 0))],
                      monitors=[Monitor(center=(.5, 0, 0)),
                                FieldMonitor(id='x-plane', normal='x', center=(.7, 0, 0), size=(0, 1.03, 1), dft_precision=dft_precision,
                                             spectrum=SpectrumSettings(sampling='frequency', frequency_points=7, apodization='none')),
                                FieldMonitor(id='y-plane', normal='y', center=(0, .2, 0), size=(1.17, 0, .8), downsample=2,
                                             time_downsample=3, dft_precision=dft_precision,
                                             spectrum=SpectrumSettings(sampling='wavelength', frequency_points=3, apodization='end'))])
    got = Simulation(project).run()
    ref = simulate.run(project.model_dump())
    np.testing.assert_array_equal(got.electric, ref['E'])
    for plane, expected in zip(got.frequency_fields, ref['planes']):
        values = expected.value
        names = list(expected.components)
        fields = values[..., [names.index(c) for c in plane['components']]]
        scale = np.abs(fields).max()
        assert scale > 0
        tol = 3e-6 if dft_precision == 'field' else 1e-12
        np.testing.assert_allclose(plane['fields']/scale, fields/scale, rtol=tol, atol=tol)


def test_blocked_plane_flush_mid_window(gpu):
    """A run stopped inside an accumulation block: the flush adds exactly the staged steps, once."""
    from torchfdtd_amd import FieldMonitor, Project, Region, Source, SpectrumSettings
    from torchfdtd_amd.solver import Stepper
    project = Project(region=Region(dimension='3d', size=(2.4, 2.0, 2.0), mesh=.1, pml_cells=4, steps=80, backend='cuda'),
                      sources=[Source(center=(-.5, 0, 0))],
                      monitors=[FieldMonitor(id='p', normal='x', center=(.45, 0, 0), size=(0, .9, .8), dft_precision='float64',
                                             spectrum=SpectrumSettings(sampling='frequency', frequency_points=5, apodization='none'))])
    run = Stepper(project)
    plane = run.planes[0]
    assert plane.block > 1
    run.advance(45)
    first = plane.result()['fields']
    again = plane.result()['fields']      # a second flush at the same step adds nothing
    ref = simulate.run(project.model_dump(), steps=45)['planes'][0]
    names = list(ref.components)
    expected = ref.value[..., [names.index(c) for c in plane.monitor.record_fields]]
    scale = np.abs(expected).max()
    np.testing.assert_allclose(first/scale, expected/scale, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(first, again)
    run.advance(35)                       # the rest of the window: the block ending at step 64 starts at the flush
    ref = simulate.run(project.model_dump())['planes'][0]
    expected = ref.value[..., [names.index(c) for c in plane.monitor.record_fields]]
    np.testing.assert_allclose(plane.result()['fields']/scale, expected/scale, rtol=1e-12, atol=1e-12)


def test_material_index_follows_permittivity_changes_between_graph_replays(gpu):
    """The one-pass step reads 1/eps through a one-byte material index; changing the permittivity in
    place between graph replays (and to more values than the index holds, and back) must give the
    reference update on the new values, with the same captured graphs."""
    from torchfdtd_amd import Project
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.solver import Stepper
    region = region_for('pml3d', 'float32')
    region.size = (2.4, 2.6, 13.0)
    region.material_sampling = 'cell'
    project = Project(region=region)
    run = Stepper(project, steps=40, one_pass=True)
    g = run.grid
    scene = Scene(project.model_dump())
    rng = np.random.default_rng(11)
    E = (rng.normal(size=(*scene.shape, 3))*.02).astype(np.float32)
    H = (rng.normal(size=(*scene.shape, 3))*.02).astype(np.float32)
    few = rng.choice(np.float32([1, .5, .25, 1/12]), size=(*scene.shape, 1)).astype(np.float32)
    many = rng.uniform(.25, 1, (*scene.shape, 1)).astype(np.float32)
    g.inverse_permittivity.copy_(torch.as_tensor(few))
    graphs = StepGraphs(run.step, run.states(), 40, 4, pair=run.step_pair, prepare=run.refresh_materials)
    ref = OracleGrid(scene, few.copy())
    g.E.copy_(torch.as_tensor(E))
    g.H.copy_(torch.as_tensor(H))
    ref.E[:], ref.H[:] = E, H
    for inv, count in ((few, 4), (many, 0), (few[::-1].copy(), 4)):
        g.inverse_permittivity.copy_(torch.as_tensor(inv))
        ref.inverse_permittivity = inv.copy()
        graphs.replay(4)
        for _ in range(4):
            ref.update_E()
            ref.update_H()
        torch.cuda.synchronize()
        assert run.materials.count == count
        np.testing.assert_array_equal(g.E.cpu().numpy(), ref.E)
        np.testing.assert_array_equal(g.H.cpu().numpy(), ref.H)

