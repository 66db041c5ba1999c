"""The reference's unfused torch update (cuda_kernel='torch', torchfdtd_amd.torch_path; the reference's
torchfdtd/boundaries.py:363-438 restated with torch operators).  On the CPU it equals the oracle bit for
bit from random fields and CPML memory (PML, periodic, PEC and 2D layouts, both precisions); on the GPU
a whole Simulation.run with cuda_kernel='torch' equals the HIP kernels' run bit for bit."""
import numpy as np
import pytest
import torch

from oracle.scene import Scene
from oracle.yee import Grid as OracleGrid
from test_hip_parity import region_for


@pytest.mark.parametrize('precision', ['float32', 'float64'])
@pytest.mark.parametrize('layout', ['pml2d', 'pml3d', 'mixed', 'pec', 'xpec'])
def test_torch_update_equals_the_oracle_bitwise(precision, layout):
    from torchfdtd_amd import Project
    from torchfdtd_amd.grid import YeeGrid
    from torchfdtd_amd.torch_path import configure_torch_kernel
    region = region_for(layout, precision)
    project = Project(region=region)
    grid = YeeGrid(region, device='cpu')
    configure_torch_kernel(grid)
    scene = Scene(project.model_dump())
    rng = np.random.default_rng(17)
    E = (rng.normal(size=grid.E.shape)*.02).astype(scene.real)
    H = (rng.normal(size=grid.H.shape)*.02).astype(scene.real)
    inv = rng.uniform(.25, 1, grid.inverse_permittivity.shape).astype(scene.real)
    ref = OracleGrid(scene, inv.copy())
    ref.E[:], ref.H[:] = E, H
    grid.E.copy_(torch.as_tensor(E))
    grid.H.copy_(torch.as_tensor(H))
    grid.inverse_permittivity.copy_(torch.as_tensor(inv))
    assert len(grid.memory_states) == len(ref.memory_states)
    for got, expected in zip(grid.memory_states, ref.memory_states):
        expected[...] = (rng.normal(size=expected.shape)*.01).astype(scene.real)
        got.copy_(torch.as_tensor(expected))
    for _ in range(8):
        grid.update_E()
        grid.update_H()
        ref.update_E()
        ref.update_H()
    np.testing.assert_array_equal(grid.E.numpy(), ref.E)
    np.testing.assert_array_equal(grid.H.numpy(), ref.H)
    for got, expected in zip(grid.memory_states, ref.memory_states):
        np.testing.assert_array_equal(got.numpy(), expected)
    assert np.abs(ref.E).max() > 0


@pytest.mark.gpu
def test_torch_kernel_run_equals_the_hip_run(gpu):
    from torchfdtd_amd import Monitor, Project, Region, Simulation, Source, Structure
    results = []
    for kernel in ('torch', 'hip'):
        region = Region(dimension='3d', size=(1.6, 1.4, 1.2), mesh=.1, pml_cells=3, steps=30, backend='cuda',
                        cuda_kernel=kernel)
        project = Project(region=region, structures=[Structure(kind='sphere', radius=.3, material='Si (constant n)')],
                          sources=[Source(center=(-.3, 0, 0), wavelength=1, pulse_cycles=1)],
                          monitors=[Monitor(center=(.3, 0, 0))])
        results.append(Simulation(project).run(cuda_graph_steps=5))
    torch_run, hip_run = results
    assert torch_run.summary['cuda_kernel'] == 'torch' and hip_run.summary['cuda_kernel'] == 'hip'
    np.testing.assert_array_equal(torch_run.electric, hip_run.electric)
    np.testing.assert_array_equal(torch_run.magnetic, hip_run.magnetic)
    np.testing.assert_array_equal(torch_run.signals, hip_run.signals)
    assert np.abs(hip_run.signals).max() > 0
