"""FP64 state norm, field peak and CPML finiteness of a grid (run_control.py:127-159)."""
from __future__ import annotations

import torch


def state_norm(grid, volume):
    """(energy, peak) in FP64; raises FloatingPointError on a non-finite CPML memory."""
    energy = torch.zeros((), device=grid.E.device, dtype=torch.float64)
    inverse = grid.inverse_permittivity.to(torch.float64)
    for j in range(3):
        e = grid.E[..., j].to(torch.float64)
        inv = inverse[..., j if inverse.shape[-1] == 3 else 0]
        energy = energy+(e*e*(volume/inv)).sum()
        h = grid.H[..., j].to(torch.float64)
        energy = energy+(h*h*volume).sum()
    peak = torch.maximum(grid.E.abs().amax().to(torch.float64), grid.H.abs().amax().to(torch.float64))
    memories = [s['psi'] for segments in grid.cpml.values() for s in segments]
    valid = torch.stack([torch.isfinite(a).all() for a in memories]).all().to(torch.float64) if memories else torch.ones_like(energy)
    values = torch.stack((energy, peak, valid)).cpu().tolist()
    if not values[2]:
        raise FloatingPointError('Non-finite CPML or incident-line memory detected.')
    return values[:2]
