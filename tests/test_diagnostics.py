"""HIP state norms against a torch FP64 reduction of the same formula (run_control.py:127-159).

The reduction order differs from torch's, so energies agree to FP64 rounding (rtol 1e-12);
peaks and non-finite flags are exact.  Cohorts of grids must reproduce each grid's own
measurement bit for bit (same task split, same order).
"""
import ctypes

import numpy as np
import pytest
import torch

from torchfdtd_amd import _native as N
from torchfdtd_amd.models import Region

pytestmark = pytest.mark.gpu


def grid(shape, precision='float32', sampling='cell', seed=0):
    from torchfdtd_amd.grid import YeeGrid
    r = Region(dimension='3d', size=tuple(n*.05 for n in shape), mesh=.05, pml_cells=3, precision=precision,
               material_sampling=sampling, backend='cuda')
    g = YeeGrid(r)
    gen = torch.Generator(device='cpu').manual_seed(seed)
    g.E.copy_(torch.randn(g.E.shape, generator=gen, dtype=torch.float64).to(g.dtype))
    g.H.copy_(torch.randn(g.H.shape, generator=gen, dtype=torch.float64).to(g.dtype))
    inv = torch.rand(g.inverse_permittivity.shape, generator=gen, dtype=torch.float64)*.9+.1
    g.inverse_permittivity.copy_(inv.to(g.dtype))
    for psi in g.psi_buffers.values():
        psi.copy_(torch.randn(psi.shape, generator=gen, dtype=torch.float64).to(g.dtype))
    return g


def torch_norm(g, widths=None):
    if widths is None:
        volume = torch.ones(g.E.shape[:3], dtype=torch.float64, device=g.E.device)
    else:
        w = [torch.as_tensor(a, device=g.E.device) for a in widths]
        volume = w[0][:, None, None]*w[1][None, :, None]*w[2][None, None, :]
    inverse = g.inverse_permittivity.to(torch.float64)
    energy = torch.zeros((), dtype=torch.float64, device=g.E.device)
    for j in range(3):
        e = g.E[..., j].to(torch.float64)
        energy = energy+(e*e*(volume/inverse[..., j if inverse.shape[-1] == 3 else 0])).sum()
        h = g.H[..., j].to(torch.float64)
        energy = energy+(h*h*volume).sum()
    peak = torch.maximum(g.E.abs().amax(), g.H.abs().amax()).to(torch.float64)
    return energy.item(), peak.item()


@pytest.mark.parametrize('precision', ['float32', 'float64'])
@pytest.mark.parametrize('sampling', ['cell', 'yee'])
def test_state_norm_matches_torch_fp64(gpu, precision, sampling):
    from torchfdtd_amd.diagnostics import StateNorms
    g = grid((21, 17, 13), precision, sampling)
    energy, peak, bad = StateNorms([g]).values()[0]
    ref_energy, ref_peak = torch_norm(g)
    assert bad == 0
    assert energy == pytest.approx(ref_energy, rel=1e-12)
    assert peak == ref_peak


def test_cohort_measures_each_grid_exactly_like_alone(gpu):
    from torchfdtd_amd.diagnostics import StateNorms
    grids = [grid(s, seed=i) for i, s in enumerate([(21, 17, 13), (12, 30, 11), (40, 11, 16)])]
    together = StateNorms(grids).values()
    for g, row in zip(grids, together):
        assert np.array_equal(StateNorms([g]).values()[0], row)


def test_non_finite_state_is_flagged(gpu):
    from torchfdtd_amd.diagnostics import StateNorms, state_norm
    g = grid((16, 14, 12))
    norms = StateNorms([g])
    assert norms.values()[0][2] == 0
    g.psi_buffers[True].view(-1)[7] = float('nan')
    assert norms.values()[0][2] == 1
    with pytest.raises(FloatingPointError):
        state_norm(g)
    g.psi_buffers[True].view(-1)[7] = 0
    g.H.view(-1)[-1] = float('inf')
    assert norms.values()[0][2] == 1


def test_graded_cell_widths_through_the_c_abi(gpu):
    lib = N.library()
    g = grid((19, 15, 11))
    rng = np.random.default_rng(3)
    widths = [rng.uniform(.5, 1.5, n) for n in (19, 15, 11)]
    tensors = [torch.as_tensor(w, device=g.E.device) for w in widths]
    d = (N.NormGrid*1)()
    d[0].dtype = N.F32
    d[0].nx, d[0].ny, d[0].nz = 19, 15, 11
    d[0].E, d[0].H = g.E.data_ptr(), g.H.data_ptr()
    d[0].inverse_permittivity, d[0].eps_stride = g.inverse_permittivity.data_ptr(), 1
    for a in range(3):
        d[0].widths[a] = tensors[a].data_ptr()
    tasks = lib.tfdtd_norm_tasks(d, 1)
    nbytes = lib.tfdtd_norm_table_bytes(d, 1)
    host = (ctypes.c_ubyte*nbytes)()
    N.check(lib.tfdtd_norm_pack(d, 1, host, nbytes), 'pack')
    table = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(g.E.device)
    partial = torch.empty((tasks, 3), device=g.E.device, dtype=torch.float64)
    out = torch.empty((1, 3), device=g.E.device, dtype=torch.float64)
    N.check(lib.tfdtd_state_norm(table.data_ptr(), 1, tasks, N.F32, partial.data_ptr(), out.data_ptr(),
                                 N.stream_handle()), 'norm')
    energy, peak, bad = out.cpu().numpy()[0]
    ref_energy, ref_peak = torch_norm(g, widths)
    assert energy == pytest.approx(ref_energy, rel=1e-12) and peak == ref_peak and bad == 0
    # Only the product with the widths changes the energy: unit widths give the unweighted sum.
    assert abs(energy-torch_norm(g)[0]) > 1e-6*abs(energy)


def test_run_diagnostics_use_the_hip_reduction(gpu):
    from torchfdtd_amd import workloads as W
    from torchfdtd_amd.solver import Simulation
    p = W.waveguide_3d(steps=100)
    p.region.run_control.check_interval = 25
    result = Simulation(p).run(cuda_graph_steps=25)
    history = result.summary['diagnostics']
    assert [h['step'] for h in history] == [25, 50, 75, 100]
    assert result.summary['diagnostic_backend'] == 'hip'
    assert all(np.isfinite(h['state_norm']) and h['state_norm'] > 0 for h in history[1:])
