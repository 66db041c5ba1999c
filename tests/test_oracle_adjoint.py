"""Pin the oracle's discrete adjoint (CPU only).

* dot-product identity <A x, y> = <x, A^T y> for one full Yee/CPML step in FP64 on random
  states, with CPML on some faces, a periodic axis and PEC upper faces (pattern of the
  reference's tests/test_cuda_adjoint.py:10-35);
* the gradient of a weighted-observation objective equals central finite differences of
  FP64 forward runs (pattern of tests/test_differentiable.py:70-89, 218-270);
* a planted mistake in the CPML memory transpose breaks the identity.
"""
import numpy as np
import pytest

import oracle.adjoint as A
from oracle.scene import Scene
from oracle.yee import Grid


def mixed_project(precision='float64', dimension='3d'):
    return {'region': {'dimension': dimension, 'size': (1.0, .9, .8), 'mesh': .1, 'precision': precision, 'pml_cells': 2,
                       'boundaries': {'y_min': {'kind': 'periodic'}, 'y_max': {'kind': 'periodic'},
                                      'z_max': {'kind': 'pec'}, 'x_min': {'kind': 'pml', 'layers': 3, 'kappa': 2,
                                                                          'alpha': .05, 'alpha_polynomial': 1}}}}


def random_state(grid, rng):
    E = rng.normal(size=grid.E.shape)
    H = rng.normal(size=grid.H.shape)
    psi = {key: [rng.normal(size=s['psi'].shape) for s in segs] for key, segs in grid.cpml.items()}
    return E, H, psi


def apply_step(grid, E, H, psi):
    grid.E[...], grid.H[...] = E, H
    for key, segs in grid.cpml.items():
        for s, v in zip(segs, psi[key]):
            s['psi'][...] = v
    grid.update_E()
    grid.update_H()
    return grid.E.copy(), grid.H.copy(), {key: [s['psi'].copy() for s in segs] for key, segs in grid.cpml.items()}


def inner(a, b):
    return float(np.sum(a[0]*b[0])+np.sum(a[1]*b[1])+sum(np.sum(x*y) for key in a[2] for x, y in zip(a[2][key], b[2][key])))


@pytest.mark.parametrize('dimension', ['2d', '3d'])
def test_dot_product_identity(dimension):
    scene = Scene(mixed_project(dimension=dimension))
    rng = np.random.default_rng(11)
    eps = rng.uniform(1, 4, (*scene.shape, 3))
    grid = Grid(scene, 1/eps)
    assert any(grid.cpml.values()) and grid.wrap and (dimension == '2d' or grid.pec_upper)
    x = random_state(grid, rng)
    y = random_state(grid, rng)
    ax = apply_step(grid, *x)
    # The transpose: state of time n in the grid (only the gradient term reads it).
    for key, segs in grid.cpml.items():
        for s, v in zip(segs, x[2][key]):
            s['psi'][...] = v
    e_bar, h_bar, psi_bar, _ = A.transpose_step(grid, x[1], y[0], y[1], y[2], eps)
    left, right = inner(ax, y), inner(x, (e_bar, h_bar, psi_bar))
    assert left == pytest.approx(right, rel=1e-12)


def test_planted_memory_mistake_breaks_the_identity(monkeypatch):
    scene = Scene(mixed_project())
    rng = np.random.default_rng(12)
    grid = Grid(scene)
    x, y = random_state(grid, rng), random_state(grid, rng)
    ax = apply_step(grid, *x)
    original = A.curl_transpose

    def wrong(grid, bar, psi_bar, forward):      # drops the c*m coupling into the derivative
        saved = {key: [s['c'] for s in segs] for key, segs in grid.cpml.items()}
        for segs in grid.cpml.values():
            for s in segs:
                s['c'] = 0*s['c']
        try:
            return original(grid, bar, psi_bar, forward)
        finally:
            for key, segs in grid.cpml.items():
                for s, c in zip(segs, saved[key]):
                    s['c'] = c
    monkeypatch.setattr(A, 'curl_transpose', wrong)
    e_bar, h_bar, psi_bar, _ = A.transpose_step(grid, x[1], y[0], y[1], y[2], np.ones(scene.shape))
    assert abs(inner(ax, y)-inner(x, (e_bar, h_bar, psi_bar))) > 1e-6*abs(inner(ax, y))


@pytest.mark.parametrize('per_component', [False, True])
def test_gradient_matches_central_differences(per_component):
    project = {'region': {'dimension': '3d', 'size': (.9, .8, .7), 'mesh': .1, 'precision': 'float64', 'pml_cells': 2}}
    scene = Scene(project)
    rng = np.random.default_rng(5)
    shape = scene.shape+((3,) if per_component else ())
    eps = rng.uniform(1.5, 3, shape)
    steps = 24
    n = int(np.prod(scene.shape))
    mid = tuple(n//2 for n in scene.shape)
    centre = ((mid[0]*scene.shape[1]+mid[1])*scene.shape[2]+mid[2])*3+2
    sources = [('Ez', mid, np.sin(.5*np.arange(steps))*np.exp(-((np.arange(steps)-6)/3.)**2))]
    observers = [('E', int(rng.integers(n))*3+1), ('H', int(rng.integers(n))*3+2), ('E', centre+3*scene.shape[2])]
    weights = rng.normal(size=(steps, len(observers)))
    J, gradient, _ = A.run_with_gradient(scene, eps, steps, observers, weights, sources)
    h = 1e-6
    cells = [tuple(int(rng.integers(s)) for s in shape) for _ in range(6)]
    cells.append(tuple(int(v) for v in np.unravel_index(np.argmax(np.abs(gradient)), shape)))
    for cell in cells:
        plus, minus = eps.copy(), eps.copy()
        plus[cell] += h
        minus[cell] -= h
        fd = (A.run_with_gradient(scene, plus, steps, observers, weights, sources)[0]
              - A.run_with_gradient(scene, minus, steps, observers, weights, sources)[0])/(2*h)
        assert gradient[cell] == pytest.approx(fd, rel=1e-6, abs=1e-9*np.abs(gradient).max())
    assert np.abs(gradient).max() > 0
