"""Density -> permittivity maps for inverse design (torch operations; the gradient flows to rho).

* ``bounded_density_layer``: an xy density extruded over [bottom, top] inside an axis-aligned
  box of a non-periodic domain; every E component gets the volume average, over a cell-sized
  box centred at its own Yee location, of eps_bg + rho (eps_design - eps_bg) inside the design
  box and of the base permittivity outside it (torchfdtd/density_layer.py:75-126).
* ``periodic_density_layer``: the same over a periodic xy cell (density_layer.py:26-61).
* ``DensityParameterization``: trainable xy density -- sigmoid logits or clamped densities,
  fixed pixels, normalised conic filter (1 - d/r)+ with periodic or truncated edges, symmetry
  orbit averaging and the tanh projection with (beta, eta)
  (torchfdtd/design_parameterization.py:25-215).
* ``cell_density_map``: rho on a box of cells, eps = eps_bg + (eps_design - eps_bg) rho, the
  per-cell map of the benchmark's 50x50x6 design region.

The overlap and projection formulas match the reference's because they define the map.
"""
from __future__ import annotations

import math

import numpy as np
import torch
from torch import nn
from torch.nn import functional as F

from .planning import field_axes


def _edges(lower, upper, count, like):
    return torch.linspace(lower, upper, count+1, device=like.device, dtype=like.dtype)


def _clipped_overlap(centers, width, edges, like):
    """(centres, pixels): length of [c - w/2, c + w/2] inside each pixel, over w."""
    c = torch.as_tensor(np.array(centers, dtype=np.float64), device=like.device, dtype=like.dtype)[:, None]
    lo, hi = c-width/2, c+width/2
    return (torch.minimum(hi, edges[None, 1:])-torch.maximum(lo, edges[None, :-1])).clamp_min(0)/width


def _z_fraction(region, component, bottom, top, like):
    z = torch.as_tensor(np.array(field_axes(region, component)[2], dtype=np.float64), device=like.device, dtype=like.dtype)
    width = region.axis_steps[2]
    return (torch.minimum(z+width/2, z.new_tensor(top))-torch.maximum(z-width/2, z.new_tensor(bottom))).clamp_min(0)/width


def _check_density(density):
    if not isinstance(density, torch.Tensor) or density.dtype not in (torch.float32, torch.float64) or density.ndim != 2 \
            or min(density.shape) < 1:
        raise ValueError('Density must be a nonempty real floating x,y tensor.')
    if not bool(torch.isfinite(density).all()) or bool(((density < 0) | (density > 1)).any()):
        raise ValueError('Density must be finite in [0,1].')


def _check_yee_region(region):
    if region.dimension != '3d' or region.mesh_type != 'uniform' or region.material_sampling != 'yee':
        raise ValueError('Density transfer requires a uniform 3D Yee-material grid.')


def bounded_density_layer(density, base_epsilon, region, *, bounds_um, bottom_um, top_um,
                          background_epsilon, design_epsilon):
    """Replace one box of `base_epsilon` (Nx,Ny,Nz,3) by the extruded density layer."""
    _check_density(density)
    if not isinstance(base_epsilon, torch.Tensor) or tuple(base_epsilon.shape) != tuple(region.shape)+(3,) \
            or base_epsilon.dtype != density.dtype:
        raise ValueError('base_epsilon must be the explicit Yee diagonal epsilon of the region in the density precision.')
    _check_yee_region(region)
    values = (*bounds_um, bottom_um, top_um, background_epsilon, design_epsilon)
    if len(bounds_um) != 4 or any(isinstance(v, torch.Tensor) or not math.isfinite(v) for v in values):
        raise ValueError('Box bounds, layer interfaces and material values must be fixed finite scalars.')
    x0, x1, y0, y1 = bounds_um
    if x0 >= x1 or y0 >= y1 or bottom_um >= top_um or min(background_epsilon, design_epsilon) < 1:
        raise ValueError('Box spans and layer thickness must be positive and permittivities at least one.')
    for axis, (lo, hi) in enumerate(((x0, x1), (y0, y1), (bottom_um, top_um))):
        inner = region.interior_bounds(axis)
        if lo < inner[0]-1e-9 or hi > inner[1]+1e-9:
            raise ValueError('The design box must lie inside the non-PML region.')
    step = region.axis_steps
    z_periodic = all(face.kind == 'periodic' for face in region.boundaries.pair(2))
    inner_z = region.interior_bounds(2)
    invariant = z_periodic and math.isclose(bottom_um, inner_z[0], abs_tol=1e-9) and math.isclose(top_um, inner_z[1], abs_tol=1e-9)
    ex, ey = _edges(x0, x1, density.shape[0], density), _edges(y0, y1, density.shape[1], density)
    fields = []
    for c, component in enumerate(('Ex', 'Ey', 'Ez')):
        x, y, _ = field_axes(region, component)
        wx, wy = _clipped_overlap(x, step[0], ex, density), _clipped_overlap(y, step[1], ey, density)
        weighted = wx@density@wy.T
        inside = wx.sum(1)[:, None]*wy.sum(1)[None, :]
        fz = density.new_ones(region.shape[2]) if invariant else _z_fraction(region, component, bottom_um, top_um, density)
        design = background_epsilon*inside+(design_epsilon-background_epsilon)*weighted
        base = base_epsilon[..., c]
        fields.append(base+fz[None, None, :]*(design[:, :, None]-base*inside[:, :, None]))
    return torch.stack(fields, -1)


def periodic_density_layer(density, region, *, bottom_um, top_um, background_epsilon, design_epsilon,
                           pixel_origin='cell_edges'):
    """Density pixels tiling a periodic xy cell, extruded over [bottom, top] (Nx,Ny,Nz,3)."""
    _check_density(density)
    if pixel_origin not in ('cell_edges', 'sample_centers'):
        raise ValueError('Unknown density pixel origin.')
    _check_yee_region(region)
    if any(face.kind != 'periodic' for axis in (0, 1) for face in region.boundaries.pair(axis)):
        raise ValueError('Density transfer requires periodic x and y boundaries.')
    values = (bottom_um, top_um, background_epsilon, design_epsilon)
    if any(isinstance(v, torch.Tensor) or not math.isfinite(v) for v in values):
        raise ValueError('Layer interfaces and material values must be fixed finite scalars.')
    if bottom_um >= top_um or min(background_epsilon, design_epsilon) < 1:
        raise ValueError('Layer thickness must be positive and permittivities at least one.')
    lower, upper = region.interior_bounds(2)
    if bottom_um < lower or top_um > upper:
        raise ValueError('Patterned layer must lie inside the non-PML region.')
    step, period = region.axis_steps, region.actual_size

    def wrapped(centers, width, count, length):
        edges = _edges(-length/2, length/2, count, density)
        if pixel_origin == 'sample_centers':
            edges = edges-length/(2*count)
        total = 0
        for shift in (-length, 0., length):
            total = total+_clipped_overlap(centers, width, edges+shift, density)
        return total
    fields = []
    for component in ('Ex', 'Ey', 'Ez'):
        x, y, _ = field_axes(region, component)
        xy = wrapped(x, step[0], density.shape[0], period[0])@density@wrapped(y, step[1], density.shape[1], period[1]).T
        fraction = _z_fraction(region, component, bottom_um, top_um, density)
        fields.append(background_epsilon+(design_epsilon-background_epsilon)*xy[:, :, None]*fraction)
    return torch.stack(fields, -1)


def cell_density_map(density, base_epsilon, box, *, background_epsilon, design_epsilon):
    """base_epsilon with cells `box` ((x0,x1),(y0,y1),(z0,z1)) set to eps_bg + (eps_d - eps_bg) rho."""
    if not isinstance(density, torch.Tensor) or density.shape != tuple(b-a for a, b in box):
        raise ValueError('density must have the shape of the box')
    sl = tuple(slice(a, b) for a, b in box)
    eps = base_epsilon.detach().to(density.dtype).clone()
    eps[sl] = background_epsilon+(design_epsilon-background_epsilon)*density
    return eps


_SYMMETRIES = ('none', 'mirror_x', 'mirror_y', 'mirror_xy', 'rotate180', 'rotate90', 'dihedral4')


def _images(value, symmetry):
    if symmetry == 'none':
        return (value,)
    if symmetry == 'mirror_x':
        return value, value.flip(0)
    if symmetry == 'mirror_y':
        return value, value.flip(1)
    if symmetry == 'mirror_xy':
        return value, value.flip(0), value.flip(1), value.flip((0, 1))
    if symmetry == 'rotate180':
        return value, value.flip((0, 1))
    turns = tuple(torch.rot90(value, k, (0, 1)) for k in range(4))
    return turns if symmetry == 'rotate90' else turns+tuple(v.flip(0) for v in turns)


class DensityParameterization(nn.Module):
    """Trainable xy density: raw parameters -> bounded density (design_parameterization.py:25-215)."""

    def __init__(self, shape, *, spacing_um, initial=None, mode='logits', filter_radius_um=0., boundary='periodic',
                 symmetry='none', beta=1., eta=.5, fixed_mask=None, fixed_values=0., dtype=torch.float32, device=None):
        super().__init__()
        shape = tuple(shape)
        if len(shape) != 2 or any(isinstance(n, bool) or not isinstance(n, int) or n < 1 for n in shape):
            raise ValueError('shape must contain two positive integer xy counts.')
        spacing = (float(spacing_um),)*2 if isinstance(spacing_um, (int, float)) else tuple(float(v) for v in spacing_um)
        if len(spacing) != 2 or any(not math.isfinite(v) or v <= 0 for v in spacing):
            raise ValueError('spacing_um must contain two fixed positive finite lengths.')
        if mode not in ('logits', 'density') or boundary not in ('periodic', 'truncate'):
            raise ValueError('Use logits/density mode and periodic/truncate boundary.')
        if symmetry not in _SYMMETRIES:
            raise ValueError('Unknown reflection or rotation symmetry.')
        if symmetry in ('rotate90', 'dihedral4') and (shape[0] != shape[1] or spacing[0] != spacing[1]):
            raise ValueError('Quarter-turn symmetry requires a square grid with equal physical spacing.')
        if not math.isfinite(filter_radius_um) or filter_radius_um < 0:
            raise ValueError('filter_radius_um must be a fixed finite nonnegative length.')
        self._check_projection(beta, eta)
        self.shape, self.mode, self.boundary, self.symmetry = shape, mode, boundary, symmetry
        where = device if device is not None else initial.device if isinstance(initial, torch.Tensor) else 'cpu'
        value = torch.as_tensor((0. if mode == 'logits' else .5) if initial is None else initial, dtype=dtype, device=where)
        if value.ndim == 0:
            value = value.expand(shape)
        if tuple(value.shape) != shape or not bool(torch.isfinite(value).all()):
            raise ValueError('initial must be finite, scalar or shaped like the xy grid.')
        if mode == 'density' and bool(((value < 0) | (value > 1)).any()):
            raise ValueError('Initial density must lie in [0,1].')
        self.design = nn.Parameter(value.detach().clone())
        mask = torch.zeros(shape, dtype=torch.bool, device=where) if fixed_mask is None else torch.as_tensor(fixed_mask, device=where)
        fixed = torch.as_tensor(fixed_values, dtype=dtype, device=where)
        fixed = fixed.expand(shape) if fixed.ndim == 0 else fixed
        if mask.dtype != torch.bool or tuple(mask.shape) != shape or tuple(fixed.shape) != shape \
                or bool(((fixed < 0) | (fixed > 1)).any()):
            raise ValueError('fixed_mask must be a boolean xy array and fixed_values lie in [0,1].')
        active = torch.where(mask, fixed, torch.zeros_like(fixed))
        if any(not torch.equal(mask, m) for m in _images(mask, symmetry)) or \
                any(not torch.equal(active, a) for a in _images(active, symmetry)):
            raise ValueError('Fixed masks and their values must respect the requested symmetry.')
        self.register_buffer('fixed_mask', mask.clone())
        self.register_buffer('fixed_values', fixed.clone())
        self.register_buffer('beta', self.design.new_tensor(float(beta)))
        self.register_buffer('eta', self.design.new_tensor(float(eta)))
        # Symmetry orbits: every pixel averages over the images of its position.
        index = torch.arange(math.prod(shape), device=where).reshape(shape)
        representative = torch.stack(_images(index, symmetry)).amin(0).reshape(-1)
        _, orbit = torch.unique(representative, sorted=True, return_inverse=True)
        self.register_buffer('_orbit', orbit)
        self.register_buffer('_orbit_size', torch.bincount(orbit).to(dtype))
        # Conic kernel (1 - d/r)+ normalised to unit sum; periodic grids keep one image per offset.
        radius = float(filter_radius_um)
        ranges = []
        for n, h in zip(shape, spacing):
            reach = min(n-1, math.ceil(radius/h))
            lo, hi = (-min(n//2, reach), min((n-1)//2, reach)) if boundary == 'periodic' else (-reach, reach)
            ranges.append((lo, hi, torch.arange(lo, hi+1, dtype=dtype, device=where)*h))
        d = (ranges[0][2][:, None]**2+ranges[1][2][None, :]**2).sqrt()
        kernel = (1-d/radius).clamp_min(0) if radius else torch.ones_like(d)
        self.register_buffer('_kernel', (kernel/kernel.sum())[None, None])
        self._padding = (-ranges[1][0], ranges[1][1], -ranges[0][0], ranges[0][1])
        ones = torch.ones((1, 1, *shape), dtype=dtype, device=where)
        norm = F.conv2d(F.pad(ones, self._padding), self._kernel) if boundary == 'truncate' else ones.new_ones((1, 1, 1, 1))
        self.register_buffer('_norm', norm)

    @staticmethod
    def _check_projection(beta, eta):
        if any(isinstance(v, torch.Tensor) or not math.isfinite(v) for v in (beta, eta)) or beta < 0 or not 0 < eta < 1:
            raise ValueError('beta must be nonnegative and eta strictly inside (0,1).')

    @torch.no_grad()
    def set_beta(self, value):
        self._check_projection(value, float(self.eta))
        self.beta.fill_(float(value))

    def forward(self, *, hard=False, straight_through=False):
        if straight_through and not hard:
            raise ValueError('straight_through requires hard=True.')
        density = self.design.sigmoid() if self.mode == 'logits' else self.design.clamp(0, 1)
        density = torch.where(self.fixed_mask, self.fixed_values, density)
        padded = F.pad(density[None, None], self._padding, mode='circular' if self.boundary == 'periodic' else 'constant')
        filtered = (F.conv2d(padded, self._kernel)/self._norm)[0, 0].clamp(0, 1)
        sums = filtered.new_zeros(self._orbit_size.shape).scatter_add(0, self._orbit, filtered.reshape(-1))
        symmetric = (sums/self._orbit_size)[self._orbit].reshape(self.shape)
        beta, eta = float(self.beta), self.eta
        if beta == 0:
            projected = symmetric
        else:
            low = torch.tanh(beta*eta)
            projected = ((low+torch.tanh(beta*(symmetric-eta)))/(low+torch.tanh(beta*(1-eta)))).clamp(0, 1)
        density = torch.where(self.fixed_mask, self.fixed_values, projected)
        if hard:
            binary = (density >= .5).to(density.dtype)
            density = density+(binary-density).detach() if straight_through else binary
        return density
