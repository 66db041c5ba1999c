"""The density gradient of the full-size plane objective against central differences.

The workload of the adjoint benchmark: 200x160x60 cells at 0.04 um, 2000 steps, 12-cell CPML,
a 50x50x6 density design (eps = 2.085 + 10.025 rho, rho seeded uniform noise) and the objective
sum |DFT Ez|^2 over an output plane at 8 wavelengths from 1.50 to 1.60 um.  The FP32 and FP64
adjoint gradients with respect to rho are compared with FP64 central differences (h = 1e-3) at
16 seeded voxels of the design, the check that torchfdtd's own gradient tests make at small
sizes (tests/test_differentiable.py, tests/test_adjoint_planes.py), here at the benchmark's size
and with a checkpointed (replaying) reverse sweep.
"""
import numpy as np
import pytest
import torch

from torchfdtd_amd.adjoint import AdjointOptions, DifferentiablePlaneSimulation
from torchfdtd_amd.density import cell_density_map
from torchfdtd_amd.workloads import adjoint_case

pytestmark = pytest.mark.gpu

H = 1e-3
VOXELS = 16
TOLERANCE = 1e-3     # relative L2 over the 16 voxels


def objective_and_model(precision, interval):
    project, box, rho, frequency = adjoint_case(precision=precision)
    model = DifferentiablePlaneSimulation(project, AdjointOptions(design_box=box, checkpoint_interval=interval))
    dtype = torch.float32 if precision == 'float32' else torch.float64
    base = torch.full(tuple(project.region.shape), 2.085, device='cuda', dtype=dtype)

    def objective(density):
        eps = cell_density_map(density, base, box, background_epsilon=2.085, design_epsilon=12.11)
        return (model(eps, frequency)['out'].fields[..., 2].abs()**2).sum()
    return objective, rho


def gradient(precision, interval):
    objective, rho = objective_and_model(precision, interval)
    density = torch.tensor(rho, device='cuda', requires_grad=True)
    value = objective(density)
    value.backward()
    return float(value), density.grad.detach().cpu().numpy().astype(np.float64), rho


def test_full_size_density_gradient_matches_central_differences(gpu):
    j32, g32, rho = gradient('float32', 46)
    j64, g64, _ = gradient('float64', 46)
    assert np.abs(g64).max() > 0
    objective, _ = objective_and_model('float64', 46)
    rng = np.random.default_rng(2024)
    flat = rng.choice(rho.size, VOXELS, replace=False)
    voxels = [tuple(int(v) for v in np.unravel_index(k, rho.shape)) for k in flat]
    base = rho.astype(np.float64)
    fd = []
    with torch.no_grad():
        for voxel in voxels:
            plus, minus = base.copy(), base.copy()
            plus[voxel] += H
            minus[voxel] -= H
            up = float(objective(torch.tensor(plus, device='cuda')))
            down = float(objective(torch.tensor(minus, device='cuda')))
            fd.append((up-down)/(2*H))
    fd = np.array(fd)
    a32 = np.array([g32[v] for v in voxels])
    a64 = np.array([g64[v] for v in voxels])
    e32 = np.linalg.norm(a32-fd)/np.linalg.norm(fd)
    e64 = np.linalg.norm(a64-fd)/np.linalg.norm(fd)
    whole = np.linalg.norm(g32-g64)/np.linalg.norm(g64)
    print(f'objective fp32 {j32:.9e} fp64 {j64:.9e}; rel L2 vs central differences: fp32 {e32:.3e}, fp64 {e64:.3e}; '
          f'fp32 vs fp64 gradient over the design {whole:.3e}', flush=True)
    assert e64 < TOLERANCE
    assert e32 < TOLERANCE
    assert whole < TOLERANCE
