"""FP64 state norm, field peak and finiteness of grids on the GPU (run_control.py:127-159).

StateNorms measures any number of grids with one pair of HIP launches
(tfdtd_state_norm): the energy sum_c vol E_c^2/inv_eps_c + vol H_c^2, the peak of
|E| and |H|, and whether E, H or the CPML memory holds a non-finite value
(the reduction of torchfdtd/cuda_diagnostics.py:63-135).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N


def cell_widths(region):
    """Per-axis cell widths over the reference step; None for a uniform mesh (every width 1)."""
    widths = [np.diff(nodes)/region.reference_step if n > 1 else np.ones(1)
              for nodes, n in zip(region.mesh_nodes, region.shape)]
    if all(np.all(w == 1) for w in widths):
        return None
    return widths


class StateNorms:
    """Energy, peak and non-finite flag of every grid in `grids`, in one pair of launches."""

    def __init__(self, grids):
        self.grids = list(grids)
        self.lib = lib = N.library()
        count = len(self.grids)
        descriptors = (N.NormGrid*count)()
        self._keep = []
        for d, g in zip(descriptors, self.grids):
            d.dtype = N.F64 if g.dtype == torch.float64 else N.F32
            d.nx, d.ny, d.nz = g.Nx, g.Ny, g.Nz
            d.E, d.H = g.E.data_ptr(), g.H.data_ptr()
            d.inverse_permittivity = g.inverse_permittivity.data_ptr()
            d.eps_stride = g.inverse_permittivity.shape[-1]
            widths = cell_widths(g.region)
            if widths is not None:
                for a, w in enumerate(widths):
                    t = torch.as_tensor(np.ascontiguousarray(w, dtype=np.float64), device=g.device)
                    self._keep.append(t)
                    d.widths[a] = t.data_ptr()
            for j, forward in enumerate((False, True)):
                psi = g.psi_buffers.get(forward)
                if psi is not None and psi.numel():
                    d.psi[j], d.psi_count[j] = psi.data_ptr(), psi.numel()
        self.tasks = lib.tfdtd_norm_tasks(descriptors, count)
        nbytes = lib.tfdtd_norm_table_bytes(descriptors, count)
        host = (ctypes.c_ubyte*nbytes)()
        N.check(lib.tfdtd_norm_pack(descriptors, count, host, nbytes), 'tfdtd_norm_pack')
        device = self.grids[0].device
        self.dtype = descriptors[0].dtype
        self.table = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.partial = torch.empty((self.tasks, 3), device=device, dtype=torch.float64)
        self.out = torch.empty((count, 3), device=device, dtype=torch.float64)

    def launch(self):
        N.check(self.lib.tfdtd_state_norm(self.table.data_ptr(), len(self.grids), self.tasks, self.dtype,
                                          self.partial.data_ptr(), self.out.data_ptr(), N.stream_handle()),
                'tfdtd_state_norm')

    def values(self):
        """(count, 3) FP64: energy, peak, non-finite flag (1.0) of each grid."""
        self.launch()
        return self.out.cpu().numpy()


def state_norm(grid):
    """(energy, peak) of one grid; raises FloatingPointError on any non-finite state value."""
    energy, peak, bad = StateNorms([grid]).values()[0]
    if bad:
        raise FloatingPointError('Non-finite full-domain or CPML state detected.')
    return float(energy), float(peak)
