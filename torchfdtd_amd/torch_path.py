"""The reference's unfused torch update (cuda_kernel='torch'), on ROCm tensors.

Restates torchfdtd/boundaries.py:363-438 with torch operators: one small kernel
per term, a full-volume curl temporary per half step.  It is selected only by an
explicit ``Region(cuda_kernel='torch')`` and serves as the reference's PyTorch FP32
path on the same GPU (the ~150-launch baseline the fused kernels replace).
"""
from __future__ import annotations

import torch

from .planning import CURL_TERMS


def _sl(axis, value, component=None):
    parts = [slice(None)]*3
    parts[axis] = value
    if component is not None:
        parts.append(component)
    return tuple(parts)


def torch_curl(grid, field, forward):
    result = torch.zeros_like(field)
    for axis, comp, out, sign in CURL_TERMS:
        if field.shape[axis] == 1:
            continue
        derivative = field[_sl(axis, slice(1, None), comp)]-field[_sl(axis, slice(None, -1), comp)]
        start = 0 if forward else 1
        for seg in grid.cpml.get((forward, axis, comp), []):
            shape = [1, 1, 1]
            shape[axis] = seg['hi']-seg['lo']
            data = derivative[_sl(axis, slice(seg['lo']-start, seg['hi']-start))]
            psi = seg['psi']
            psi *= seg['b'].reshape(shape)
            psi += seg['c'].reshape(shape)*data
            data *= seg['inv_k'].reshape(shape)
            data += psi
        result[_sl(axis, slice(None, -1) if forward else slice(1, None), out)] += sign*derivative
        if axis in grid.wrap:
            edge = field[_sl(axis, 0, comp)]-field[_sl(axis, -1, comp)]
            result[_sl(axis, -1 if forward else 0, out)] += sign*edge
        elif forward and axis in grid.pec_upper:
            edge = -grid.pec_upper[axis]*field[_sl(axis, -1, comp)]
            result[_sl(axis, -1, out)] += sign*edge
    return result


def configure_torch_kernel(grid):
    def update_E():
        grid.E += grid.courant_number*grid.inverse_permittivity*torch_curl(grid, grid.H, False)

    def update_H():
        grid.H -= grid.courant_number*grid.inverse_permeability*torch_curl(grid, grid.E, True)
    grid.update_E = update_E
    grid.update_H = update_H
