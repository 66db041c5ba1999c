"""Density -> permittivity maps (CPU): closed forms of the box overlaps, the conic filter and the
projection, symmetry, and autograd gradients (torch gradcheck in FP64)."""
import math

import numpy as np
import pytest
import torch

from torchfdtd_amd.density import (DensityParameterization, bounded_density_layer, cell_density_map,
                                   periodic_density_layer)
from torchfdtd_amd.models import BoundaryFace, Region


def yee_region(periodic=False):
    r = Region(dimension='3d', size=(2.0, 2.0, 1.6), mesh=.1, pml_cells=4, material_sampling='yee', precision='float64')
    if periodic:
        for a in 'xy':
            setattr(r.boundaries, a+'_min', BoundaryFace(kind='periodic'))
            setattr(r.boundaries, a+'_max', BoundaryFace(kind='periodic'))
    return r


def test_bounded_layer_closed_form():
    r = yee_region()
    base = torch.full(tuple(r.shape)+(3,), 1.5, dtype=torch.float64)
    ones = torch.ones((4, 4), dtype=torch.float64)
    eps = bounded_density_layer(ones, base, r, bounds_um=(-.4, .4, -.4, .4), bottom_um=-.2, top_um=.2,
                                background_epsilon=2., design_epsilon=12.)
    # Ez sits at cell centres in x and y: cell (10,10) spans [0, .1]^2 fully inside the box, z centre 0.05.
    assert eps[10, 10, 8, 2].item() == pytest.approx(12.)
    assert eps[2, 2, 8, 2].item() == 1.5 and eps[10, 10, 14, 2].item() == 1.5
    zero = bounded_density_layer(0*ones, base, r, bounds_um=(-.4, .4, -.4, .4), bottom_um=-.2, top_um=.2,
                                 background_epsilon=2., design_epsilon=12.)
    assert zero[10, 10, 8, 2].item() == pytest.approx(2.)
    # Ez sits at x nodes: the node on the box edge x=-.4 sees half of its cell box inside.
    x = np.round(np.asarray(r.mesh_nodes[0][:-1]), 9)
    edge = int(np.flatnonzero(np.isclose(x, -.4))[0])
    assert eps[edge, 10, 8, 2].item() == pytest.approx(1.5+.5*(12-1.5))


def test_density_maps_have_exact_gradients():
    r = yee_region()
    base = torch.full(tuple(r.shape)+(3,), 1.5, dtype=torch.float64)
    rho = torch.rand((3, 5), dtype=torch.float64, requires_grad=True)
    torch.autograd.gradcheck(lambda d: bounded_density_layer(d, base, r, bounds_um=(-.33, .41, -.3, .25), bottom_um=-.17,
                                                               top_um=.23, background_epsilon=2., design_epsilon=9.).sum(-1)[8:12, 8:12, 6:10],
                             (rho,))
    rp = yee_region(periodic=True)
    torch.autograd.gradcheck(lambda d: periodic_density_layer(d, rp, bottom_um=-.1, top_um=.15, background_epsilon=1.,
                                                                design_epsilon=4.)[::3, ::3, 6:10], (rho,))
    cells = torch.rand((2, 3, 2), dtype=torch.float64, requires_grad=True)
    out = cell_density_map(cells, torch.ones((6, 6, 6), dtype=torch.float64), ((1, 3), (2, 5), (3, 5)),
                           background_epsilon=2.085, design_epsilon=12.11)
    out.sum().backward()
    assert torch.allclose(cells.grad, torch.full_like(cells, 12.11-2.085))
    assert out[0, 0, 0].item() == 1 and out[1, 2, 3].item() == pytest.approx(2.085+10.025*cells[0, 0, 0].item())


def test_periodic_layer_uniform_density_is_uniform():
    r = yee_region(periodic=True)
    eps = periodic_density_layer(torch.full((5, 7), .3, dtype=torch.float64), r, bottom_um=-.2, top_um=.2,
                                 background_epsilon=1., design_epsilon=3.)
    inside = eps[:, :, 8]
    assert torch.allclose(inside, torch.full_like(inside, 1+.3*2))


def test_parameterization_filter_projection_and_symmetry():
    p = DensityParameterization((9, 9), spacing_um=.05, mode='density', initial=.37, filter_radius_um=.12, beta=0,
                                dtype=torch.float64)
    assert torch.allclose(p(), torch.full((9, 9), .37, dtype=torch.float64))      # unit-sum kernel, periodic edges
    assert p.filter.weight.sum().item() == pytest.approx(1)
    q = DensityParameterization((8, 8), spacing_um=.05, mode='density', initial=torch.rand(8, 8, dtype=torch.float64),
                                filter_radius_um=.1, symmetry='dihedral4', beta=8, eta=.4, dtype=torch.float64)
    out = q()
    assert torch.allclose(out, out.flip(0)) and torch.allclose(out, out.T)
    eta, beta = .4, 8.
    proj = lambda x: (math.tanh(beta*eta)+math.tanh(beta*(x-eta)))/(math.tanh(beta*eta)+math.tanh(beta*(1-eta)))
    assert proj(0) == pytest.approx(0, abs=1e-15) and proj(1) == pytest.approx(1)
    hard = q(hard=True)
    assert set(hard.unique().tolist()) <= {0., 1.}
    st = q(hard=True, straight_through=True)
    st.sum().backward()
    assert q.design.grad is not None and torch.isfinite(q.design.grad).all()
    # Exact derivative of the smooth map: central differences of the raw parameters.
    r = DensityParameterization((6, 5), spacing_um=.05, mode='logits', initial=torch.randn(6, 5, dtype=torch.float64),
                                filter_radius_um=.08, boundary='truncate', beta=2., dtype=torch.float64)
    w = torch.randn(6, 5, dtype=torch.float64)
    (r()*w).sum().backward()
    for cell in [(0, 0), (2, 3), (5, 4)]:
        with torch.no_grad():
            r.design[cell] += 1e-6
            up = (r()*w).sum().item()
            r.design[cell] -= 2e-6
            down = (r()*w).sum().item()
            r.design[cell] += 1e-6
        assert r.design.grad[cell].item() == pytest.approx((up-down)/2e-6, rel=1e-6, abs=1e-9)
    with pytest.raises(ValueError):
        DensityParameterization((4, 5), spacing_um=.05, symmetry='rotate90')


def brute_filter(density, spacing, radius, boundary):
    """(1 - d/r)+ weighted mean by direct summation: periodic over minimum-image torus offsets (one
    per residue), truncate over in-domain pixels only."""
    nx, ny = density.shape
    out = np.zeros((nx, ny))
    offsets = []
    for n in (nx, ny):
        reach = min(n-1, math.ceil(radius/spacing))
        offsets.append(range(-min(n//2, reach), min((n-1)//2, reach)+1) if boundary == 'periodic' else range(-reach, reach+1))
    for i in range(nx):
        for j in range(ny):
            total = norm = 0.
            for a in offsets[0]:
                for b in offsets[1]:
                    w = max(0., 1-math.hypot(a*spacing, b*spacing)/radius)
                    if boundary == 'periodic':
                        total += w*density[(i+a) % nx, (j+b) % ny]
                        norm += w
                    elif 0 <= i+a < nx and 0 <= j+b < ny:
                        total += w*density[i+a, j+b]
                        norm += w
            out[i, j] = total/norm
    return out


@pytest.mark.parametrize('boundary,shape,radius', [('periodic', (7, 6), .12), ('periodic', (4, 5), .4),
                                                   ('truncate', (7, 6), .12), ('truncate', (5, 3), .2)])
def test_conic_filter_equals_direct_summation(boundary, shape, radius):
    rng = np.random.default_rng(7)
    rho = rng.random(shape)
    p = DensityParameterization(shape, spacing_um=.05, mode='density', initial=torch.as_tensor(rho), filter_radius_um=radius,
                                boundary=boundary, beta=0, dtype=torch.float64)
    np.testing.assert_allclose(p().detach().numpy(), brute_filter(rho, .05, radius, boundary), rtol=1e-12, atol=1e-14)


def test_symmetry_mean_and_fixed_pixels():
    rng = np.random.default_rng(8)
    rho = torch.as_tensor(rng.random((6, 6)))
    p = DensityParameterization((6, 6), spacing_um=.05, mode='density', initial=rho, symmetry='mirror_x', beta=0,
                                dtype=torch.float64)
    np.testing.assert_allclose(p().detach().numpy(), ((rho+rho.flip(0))/2).numpy(), rtol=0, atol=1e-15)
    mask = torch.zeros((6, 6), dtype=torch.bool)
    mask[0, :] = mask[-1, :] = True
    q = DensityParameterization((6, 6), spacing_um=.05, mode='density', initial=rho, symmetry='mirror_x', beta=4.,
                                fixed_mask=mask, fixed_values=1., dtype=torch.float64)
    out = q()
    assert torch.all(out[mask] == 1) and torch.allclose(out, out.flip(0))
    with pytest.raises(ValueError):
        mask[0, 0] = False
        DensityParameterization((6, 6), spacing_um=.05, symmetry='mirror_x', fixed_mask=mask)
