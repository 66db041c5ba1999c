"""Pin the C restatement of the oracle's half steps (oracle/native.py) to the NumPy oracle (CPU only).

The C grid must equal oracle/yee.py bit for bit -- fields and every CPML memory -- from random
fields and random CPML memory in every boundary layout the engine supports, in both precisions,
with 1/eps per Yee component, per cell and as one value, and must reproduce the reference-written
golden runs.  Only then do the configuration-size parity tests lean on it.
"""
import json

import numpy as np
import pytest

from oracle import simulate
from oracle.native import NativeGrid
from oracle.scene import Scene
from oracle.yee import Grid
from tests.conftest import GOLDEN

LAYOUTS = {
    'pml2d': ('2d', {}),
    'pml3d': ('3d', {'x_min': {'layers': 5, 'kappa': 3, 'alpha': .03, 'alpha_polynomial': 1}}),
    'periodic': ('3d', {'y_min': {'kind': 'periodic'}, 'y_max': {'kind': 'periodic'},
                        'x_min': {'kind': 'periodic'}, 'x_max': {'kind': 'periodic'}}),
    'pec': ('3d', {'z_min': {'kind': 'pec'}, 'z_max': {'kind': 'pec'}, 'y_max': {'kind': 'pec'}}),
    'xpec': ('3d', {'x_max': {'kind': 'pec'}, 'x_min': {'kind': 'pec'}}),
    'periodic2d': ('2d', {'y_min': {'kind': 'periodic'}, 'y_max': {'kind': 'periodic'}, 'x_max': {'kind': 'pec'}}),
}


def project_for(layout, precision):
    dimension, faces = LAYOUTS[layout]
    return {'region': {'dimension': dimension, 'size': (1.9, 1.7, 1.3), 'mesh': .1, 'pml_cells': 4,
                       'precision': precision, 'boundaries': faces}}


@pytest.mark.parametrize('precision', ['float32', 'float64'])
@pytest.mark.parametrize('layout', sorted(LAYOUTS))
@pytest.mark.parametrize('components', [3, 1, 0])
def test_native_half_steps_equal_numpy_bitwise(layout, precision, components):
    scene = Scene(project_for(layout, precision))
    real = scene.real
    rng = np.random.default_rng(2024)
    shape = (*scene.shape, components) if components else (1,)
    inv = rng.uniform(.25, 1, shape).astype(real)
    grids = [Grid(scene, inv.copy()), NativeGrid(scene, inv.copy())]
    E = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    H = (rng.normal(size=(*scene.shape, 3))*.02).astype(real)
    # a few subnormals and signed zeros: the rounding of every operation is compared
    E.reshape(-1)[::97] = np.finfo(real).tiny/8
    H.reshape(-1)[::89] = -0.
    states = [(rng.normal(size=s.shape)*.01).astype(real) for s in grids[0].memory_states]
    for g in grids:
        g.E[:], g.H[:] = E, H
        for mine, start in zip(g.memory_states, states):
            mine[...] = start
    assert len(states) > 0 or layout == 'periodic'
    for _ in range(9):
        for g in grids:
            g.update_E()
            g.update_H()
    ref, got = grids
    np.testing.assert_array_equal(got.E, ref.E)
    np.testing.assert_array_equal(got.H, ref.H)
    assert np.array_equal(np.signbit(got.E), np.signbit(ref.E))
    for a, b in zip(got.memory_states, ref.memory_states):
        np.testing.assert_array_equal(a, b)
    assert float(np.abs(ref.E).max()) > 0


@pytest.mark.parametrize('name,steps', [('tiny_2d_complete', 40), ('tiny_2d_cancelled', 20)])
def test_native_oracle_reproduces_reference_golden_runs_bitwise(name, steps):
    data = np.load(GOLDEN/f'{name}.npz', allow_pickle=False)
    project = json.loads(str(data['project']))
    out = simulate.run(project, steps=steps, native=True)
    assert isinstance(out['grid'], NativeGrid)
    np.testing.assert_array_equal(out['E'], data['E'])
    np.testing.assert_array_equal(out['H'], data['H'])
    np.testing.assert_array_equal(out['signals'], data['signals'])


def test_native_oracle_fails_for_a_planted_sign_error(monkeypatch):
    """The bitwise comparison is not vacuous: a flipped curl sign in the NumPy grid is detected."""
    import oracle.yee as yee
    flipped = list(yee.CURL_TERMS)
    flipped[4] = (2, 1, 0, 1)
    monkeypatch.setattr(yee, 'CURL_TERMS', tuple(flipped))
    with pytest.raises(AssertionError):
        test_native_half_steps_equal_numpy_bitwise('pml3d', 'float32', 3)
