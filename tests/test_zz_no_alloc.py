"""tfdtd_state_norm_noalloc (partial sums in the library's own device buffer) against
tfdtd_state_norm (partial sums in a caller workspace): bit-for-bit equal results through the C ABI
on two calls in a row, the second on a smaller input, no device memory taken by the call, and a
clear error for inputs beyond the buffer."""
import ctypes

import numpy as np
import pytest
import torch

from torchfdtd_amd import _native as N
from torchfdtd_amd.models import Region

pytestmark = pytest.mark.gpu


def grids(shapes, seed):
    from torchfdtd_amd.grid import YeeGrid
    out = []
    gen = torch.Generator(device='cpu').manual_seed(seed)
    for shape in shapes:
        r = Region(dimension='3d', size=tuple(n*.05 for n in shape), mesh=.05, pml_cells=4, backend='cuda')
        g = YeeGrid(r)
        g.E.copy_(torch.randn(g.E.shape, generator=gen, dtype=torch.float64).to(g.dtype))
        g.H.copy_(torch.randn(g.H.shape, generator=gen, dtype=torch.float64).to(g.dtype))
        g.inverse_permittivity.copy_((torch.rand(g.inverse_permittivity.shape, generator=gen, dtype=torch.float64)*.9+.1)
                                     .to(g.dtype))
        for psi in g.psi_buffers.values():
            psi.copy_(torch.randn(psi.shape, generator=gen, dtype=torch.float64).to(g.dtype))
        out.append(g)
    return out


def both(norms):
    """(workspace result, no-allocation result) of one StateNorms table."""
    lib, stream = N.library(), N.stream_handle()
    count = len(norms.grids)
    N.check(lib.tfdtd_state_norm(norms.table.data_ptr(), count, norms.tasks, norms.dtype, norms.partial.data_ptr(),
                                 norms.out.data_ptr(), stream), 'tfdtd_state_norm')
    expected = norms.out.cpu().numpy().copy()
    out = torch.full((count, 3), float('nan'), device=norms.out.device, dtype=torch.float64)
    torch.cuda.synchronize()
    free_before = torch.cuda.mem_get_info()[0]
    N.check(lib.tfdtd_state_norm_noalloc(norms.table.data_ptr(), count, norms.tasks, norms.dtype, out.data_ptr(), stream),
            'tfdtd_state_norm_noalloc')
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free_before
    return expected, out.cpu().numpy()


def test_no_alloc_norm_equals_the_workspace_entry_point(gpu):
    from torchfdtd_amd.diagnostics import StateNorms
    big = StateNorms(grids([(96, 80, 72), (40, 33, 29)], seed=1))
    small = StateNorms(grids([(23, 19, 17)], seed=2))
    assert big.tasks > 3*small.tasks and big.tasks <= 262144
    for norms in (big, small):       # the second call is smaller: rows of the first call stay behind
        expected, got = both(norms)
        np.testing.assert_array_equal(got, expected)
        assert np.all(expected[:, 0] > 0) and np.all(expected[:, 2] == 0)
    # A non-finite CPML value is flagged by both.
    small.grids[0].psi_buffers[True].view(-1)[5] = float('inf')
    expected, got = both(small)
    np.testing.assert_array_equal(got, expected)
    assert got[0, 2] == 1


def test_no_alloc_norm_rejects_inputs_beyond_its_buffer(gpu):
    from torchfdtd_amd.diagnostics import StateNorms
    norms = StateNorms(grids([(16, 15, 14)], seed=3))
    lib = N.library()
    out = torch.zeros((1, 3), device=gpu, dtype=torch.float64)
    status = lib.tfdtd_state_norm_noalloc(norms.table.data_ptr(), 1, 262145, norms.dtype, out.data_ptr(), N.stream_handle())
    assert status != 0 and 'exceed' in N.last_error()
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0
