"""Density -> permittivity maps for inverse design (torch operations; the gradient flows to rho).

* ``bounded_density_layer``: an xy density extruded over [bottom, top] inside an axis-aligned
  box of a non-periodic domain; every E component gets the volume average, over a cell-sized
  box centred at its own Yee location, of eps_bg + rho (eps_design - eps_bg) inside the design
  box and of the base permittivity outside it (torchfdtd/density_layer.py:75-126).
* ``periodic_density_layer``: the same over a periodic xy cell (density_layer.py:26-61).
* ``DensityParameterization``: trainable xy density -- sigmoid logits or clamped densities,
  fixed pixels, the normalised conic filter (1 - d/r)+ with periodic or truncated edges (an FFT
  correlation), the mean over a symmetry group's images and the tanh projection with (beta, eta)
  (torchfdtd/design_parameterization.py:25-215).
* ``cell_density_map``: rho on a box of cells, eps = eps_bg + (eps_design - eps_bg) rho, the
  per-cell map of the benchmark's 50x50x6 design region.

The box-overlap integrals, the conic kernel and the projection formula match the reference's
because they define the maps; the rest is this module's own construction.
"""
from __future__ import annotations

import math

import numpy as np
import torch
from torch import nn

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


class _Checks:
    """Argument checks of the density maps: each failed requirement raises ValueError with its text."""

    @staticmethod
    def require(condition, text):
        if not condition:
            raise ValueError(text)

    @classmethod
    def scalars(cls, *values):
        cls.require(all(not isinstance(v, torch.Tensor) and isinstance(v, (int, float)) and math.isfinite(v) for v in values),
                    'geometry and material values of a density map are fixed finite numbers')

    @classmethod
    def density(cls, density):
        cls.require(isinstance(density, torch.Tensor) and density.dtype in (torch.float32, torch.float64)
                    and density.ndim == 2 and density.numel() > 0, 'density is a non-empty real (x, y) float tensor')
        cls.require(bool(torch.isfinite(density).all()) and not bool(((density < 0) | (density > 1)).any()),
                    'density values lie in [0, 1]')

    @classmethod
    def yee_region(cls, region):
        cls.require(region.dimension == '3d' and region.mesh_type == 'uniform' and region.material_sampling == 'yee',
                    'density layers need a uniform 3D region with Yee material sampling')


def bounded_density_layer(density, base_epsilon, region, *, bounds_um, bottom_um, top_um,
                          background_epsilon, design_epsilon):
    """Replace one box of `base_epsilon` (Nx,Ny,Nz,3) by the extruded density layer."""
    check = _Checks
    check.density(density)
    check.yee_region(region)
    check.require(isinstance(base_epsilon, torch.Tensor) and base_epsilon.dtype == density.dtype
                  and tuple(base_epsilon.shape) == tuple(region.shape)+(3,),
                  'base_epsilon is the (Nx, Ny, Nz, 3) Yee permittivity of the region, in the density dtype')
    check.require(len(bounds_um) == 4, 'bounds_um is (x0, x1, y0, y1)')
    check.scalars(*bounds_um, bottom_um, top_um, background_epsilon, design_epsilon)
    x0, x1, y0, y1 = bounds_um
    spans = ((x0, x1), (y0, y1), (bottom_um, top_um))
    check.require(all(lo < hi for lo, hi in spans), 'the box and the layer have positive extent')
    check.require(min(background_epsilon, design_epsilon) >= 1, 'permittivities are at least 1')
    for axis, (lo, hi) in enumerate(spans):
        inner_lo, inner_hi = region.interior_bounds(axis)
        check.require(inner_lo-1e-9 <= lo and hi <= inner_hi+1e-9, 'the design box stays out of the CPML')
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
    check = _Checks
    check.density(density)
    check.yee_region(region)
    check.require(pixel_origin in ('cell_edges', 'sample_centers'), "pixel_origin is 'cell_edges' or 'sample_centers'")
    check.require(all(face.kind == 'periodic' for axis in (0, 1) for face in region.boundaries.pair(axis)),
                  'a periodic density layer needs periodic x and y faces')
    check.scalars(bottom_um, top_um, background_epsilon, design_epsilon)
    check.require(bottom_um < top_um, 'the layer has positive thickness')
    check.require(min(background_epsilon, design_epsilon) >= 1, 'permittivities are at least 1')
    lower, upper = region.interior_bounds(2)
    check.require(lower <= bottom_um and top_um <= upper, 'the layer stays out of the CPML')
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


def _group_images(value, symmetry):
    """The images of an (x, y) array under every element of a symmetry group."""
    flips = {'none': ((),), 'mirror_x': ((), (0,)), 'mirror_y': ((), (1,)), 'mirror_xy': ((), (0,), (1,), (0, 1)),
             'rotate180': ((), (0, 1))}
    if symmetry in flips:
        return [value.flip(axes) if axes else value for axes in flips[symmetry]]
    turns = [torch.rot90(value, k, (0, 1)) for k in range(4)]
    return turns if symmetry == 'rotate90' else turns+[t.flip(0) for t in turns]


SYMMETRIES = ('none', 'mirror_x', 'mirror_y', 'mirror_xy', 'rotate180', 'rotate90', 'dihedral4')


class ConicFilter(nn.Module):
    """Normalised conic filter w(d) = (1 - d/r)+ / sum w on an (x, y) pixel grid of given spacing
    (design_parameterization.py:100-115), applied as an FFT correlation.

    'periodic': each torus offset counts once (minimum image; for an even count the half-period
    neighbour counts once).  'truncate': the grid is zero padded and each output divided by the
    filtered grid of ones, so pixels outside the domain drop out of the average.  Radius 0 keeps
    the density unchanged."""

    def __init__(self, shape, spacing, radius, boundary, dtype, device):
        super().__init__()
        self.shape, self.boundary = tuple(shape), boundary
        self.identity = radius == 0
        if self.identity:
            return
        axes = []
        for n, h in zip(self.shape, spacing):
            reach = min(n-1, math.ceil(radius/h))
            lo, hi = (-min(n//2, reach), min((n-1)//2, reach)) if boundary == 'periodic' else (-reach, reach)
            axes.append(torch.arange(lo, hi+1, dtype=torch.float64))
        offsets = [a*h for a, h in zip(axes, spacing)]
        weight = (1-torch.sqrt(offsets[0][:, None]**2+offsets[1][None, :]**2)/radius).clamp_min(0)
        weight = weight/weight.sum()
        if boundary == 'periodic':
            self.padded = self.shape
            torus = torch.zeros(self.shape, dtype=torch.float64)
            torus[(axes[0].long() % self.shape[0])[:, None], (axes[1].long() % self.shape[1])[None, :]] = weight
        else:
            self.reach = [int(a[-1]) for a in axes]
            self.padded = tuple(n+2*r for n, r in zip(self.shape, self.reach))
            torus = torch.zeros(self.padded, dtype=torch.float64)
            torus[(axes[0].long() % self.padded[0])[:, None], (axes[1].long() % self.padded[1])[None, :]] = weight
        # Correlation with w equals multiplication by the conjugate spectrum of w on the torus.
        self.register_buffer('spectrum', torch.fft.rfft2(torus).conj().to(device))
        self.register_buffer('weight', weight.to(dtype=dtype, device=device))
        if boundary == 'truncate':
            self.register_buffer('coverage', self._correlate(torch.ones(self.shape, dtype=torch.float64,
                                                                        device=device)).to(dtype))

    def _correlate(self, values):
        if self.boundary == 'truncate':
            values = torch.nn.functional.pad(values, (0, 2*self.reach[1], 0, 2*self.reach[0]))
            values = values.roll((self.reach[0], self.reach[1]), (0, 1))
        spectrum = self.spectrum if values.dtype == torch.float64 else self.spectrum.to(torch.complex64)
        out = torch.fft.irfft2(torch.fft.rfft2(values)*spectrum, s=self.padded)
        if self.boundary == 'truncate':
            out = out[self.reach[0]:self.reach[0]+self.shape[0], self.reach[1]:self.reach[1]+self.shape[1]]
        return out

    def forward(self, density):
        if self.identity:
            return density
        out = self._correlate(density)
        return out/self.coverage if self.boundary == 'truncate' else out


class DensityParameterization(nn.Module):
    """Trainable (x, y) density for inverse design (design_parameterization.py:25-215):

        raw --bound--> [0,1] --pin fixed--> --conic filter--> clip [0,1] --group mean-->
            --tanh projection (beta, eta)--> --pin fixed--> density  [--> hard 0/1]

    `mode` 'logits' bounds by a sigmoid, 'density' by a clamp (zero derivative outside [0,1]).
    The group mean averages the density over the images of the symmetry group, i.e. over each
    pixel's orbit.  beta = 0 skips the projection.  `hard` thresholds at 1/2; with
    `straight_through` the forward value is the threshold and the derivative the smooth map's."""

    def __init__(self, shape, *, spacing_um, initial=None, mode='logits', filter_radius_um=0., boundary='periodic',
                 symmetry='none', beta=1., eta=.5, fixed_mask=None, fixed_values=0., dtype=torch.float32, device=None):
        super().__init__()
        need = _Checks.require
        shape = tuple(shape)
        need(len(shape) == 2 and all(type(n) is int and n > 0 for n in shape), 'shape is two positive integers (nx, ny)')
        spacing = (spacing_um, spacing_um) if isinstance(spacing_um, (int, float)) else tuple(spacing_um)
        need(len(spacing) == 2, 'spacing_um is one length or (dx, dy)')
        _Checks.scalars(*spacing, filter_radius_um)
        spacing = tuple(float(v) for v in spacing)
        need(min(spacing) > 0 and filter_radius_um >= 0, 'spacings are positive and the filter radius non-negative')
        need(mode in ('logits', 'density'), "mode is 'logits' or 'density'")
        need(boundary in ('periodic', 'truncate'), "boundary is 'periodic' or 'truncate'")
        need(symmetry in SYMMETRIES, f'symmetry is one of {SYMMETRIES}')
        need(symmetry not in ('rotate90', 'dihedral4') or (shape[0] == shape[1] and spacing[0] == spacing[1]),
             'quarter turns need a square grid of equal spacings')
        self._check_projection(beta, eta)
        self.shape, self.mode, self.symmetry = shape, mode, symmetry
        if device is None:
            device = initial.device if isinstance(initial, torch.Tensor) else 'cpu'
        start = torch.as_tensor((0. if mode == 'logits' else .5) if initial is None else initial, dtype=dtype, device=device)
        start = start.expand(shape) if start.ndim == 0 else start
        need(tuple(start.shape) == shape and bool(torch.isfinite(start).all()), 'initial is finite: a scalar or an (nx, ny) array')
        need(mode == 'logits' or not bool(((start < 0) | (start > 1)).any()), 'an initial density lies in [0, 1]')
        self.design = nn.Parameter(start.detach().clone())
        mask = torch.zeros(shape, dtype=torch.bool, device=device) if fixed_mask is None else torch.as_tensor(fixed_mask, device=device)
        pinned = torch.as_tensor(fixed_values, dtype=dtype, device=device)
        pinned = pinned.expand(shape) if pinned.ndim == 0 else pinned
        need(mask.dtype == torch.bool and tuple(mask.shape) == shape, 'fixed_mask is a boolean (nx, ny) array')
        need(tuple(pinned.shape) == shape and not bool(((pinned < 0) | (pinned > 1)).any()), 'fixed_values lie in [0, 1]')
        held = pinned*mask
        need(all(torch.equal(image, mask) for image in _group_images(mask, symmetry))
             and all(torch.equal(image, held) for image in _group_images(held, symmetry)),
             'the fixed pixels and their values are invariant under the symmetry')
        self.register_buffer('fixed_mask', mask.clone())
        self.register_buffer('fixed_values', pinned.clone())
        self.register_buffer('beta', self.design.new_tensor(float(beta)))
        self.register_buffer('eta', self.design.new_tensor(float(eta)))
        self.filter = ConicFilter(shape, spacing, float(filter_radius_um), boundary, dtype, device)

    @staticmethod
    def _check_projection(beta, eta):
        _Checks.scalars(beta, eta)
        _Checks.require(beta >= 0 and 0 < eta < 1, 'beta >= 0 and 0 < eta < 1')

    @torch.no_grad()
    def set_beta(self, value):
        self._check_projection(value, float(self.eta))
        self.beta.fill_(float(value))

    def project(self, density):
        """(tanh(beta eta) + tanh(beta (rho - eta))) / (tanh(beta eta) + tanh(beta (1 - eta))), clipped to [0,1]."""
        beta, eta = float(self.beta), float(self.eta)
        if beta == 0:
            return density
        low, high = math.tanh(beta*eta), math.tanh(beta*(1-eta))
        return ((low+torch.tanh(beta*(density-eta)))/(low+high)).clamp(0, 1)

    def forward(self, *, hard=False, straight_through=False):
        _Checks.require(hard or not straight_through, 'straight_through needs hard=True')
        bounded = self.design.sigmoid() if self.mode == 'logits' else self.design.clamp(0, 1)
        bounded = torch.where(self.fixed_mask, self.fixed_values, bounded)
        filtered = self.filter(bounded).clamp(0, 1)
        symmetric = torch.stack(_group_images(filtered, self.symmetry)).mean(0)
        density = torch.where(self.fixed_mask, self.fixed_values, self.project(symmetric))
        if not hard:
            return density
        binary = (density >= .5).to(density.dtype)
        return density+(binary-density).detach() if straight_through else binary
