"""Pins of the oracle's dispersive-material ADE (oracle/materials.py, a restatement of
torchfdtd/materials.py:29-87) by closed forms (CPU).

A uniform field in a fully periodic grid has zero curl, so each cell evolves by the ADE alone:
* the displacement eps_inf E + sum_j P_j is conserved step by step (exactly, up to FP64 rounding):
  a flipped sign of delta, of Q's update or of k_sum breaks it;
* with damping, a Lorentz medium relaxes to its static permittivity: E -> eps_inf E0 / (eps_inf +
  sum_j s_j / w0_j^2);
* a Drude medium (w0 = 0, infinite static susceptibility) drives E to 0.
And the reference's discrete permittivity (materials.py:15-26, the bilinear map
omega -> (2/dt) tan(omega dt/2)) is what a driven single cell reproduces: a time-harmonic E with
the trapezoidal P gives P/E = chi(omega~) to FP64 accuracy.
"""
import math

import numpy as np
import pytest

from oracle import simulate
from oracle.materials import MaterialADE, oscillators


def periodic_project(material, steps):
    faces = {k: {'kind': 'periodic'} for k in ('x_min', 'x_max', 'y_min', 'y_max', 'z_min', 'z_max')}
    return {'region': {'dimension': '3d', 'size': (.6, .6, .6), 'mesh': .1, 'steps': steps, 'precision': 'float64',
                       'boundaries': faces},
            'materials': [dict(name='metal', **material)],
            'structures': [{'kind': 'rectangle', 'size': (10, 10, 10), 'material': 'metal'}]}


def uniform_run(material, steps, e0=(1., -.5, .25)):
    project = periodic_project(material, steps)
    scene, eps, grid = simulate.prepare(project)
    grid.E[:] = np.asarray(e0)
    (state,) = grid.material_states
    displacement0 = state.eps*grid.E.copy()
    history = []
    for _ in range(steps):
        prepared = state.prepare(grid.E)
        grid.update_E()
        state.correct(grid.E, *prepared)
        grid.update_H()
        P = state.P.sum(axis=0) if state.multiple else state.P
        history.append((grid.E.reshape(-1, 3)[0].copy(), (state.eps*grid.E.reshape(-1, 3)+P)))
    return grid, state, displacement0.reshape(-1, 3), history


LORENTZ = dict(model='lorentz', epsilon_inf=2.0, resonance_rad_s=3e15, delta_epsilon=1.5, linewidth_rad_s=4e14)
TWO_POLES = dict(model='multipole', epsilon_inf=1.5, poles=[
    dict(resonance_rad_s=2e15, strength_rad_s_squared=8e30, damping_rad_s=6e14),
    dict(resonance_rad_s=5e15, strength_rad_s_squared=1e31, damping_rad_s=1e15)])
DRUDE = dict(model='drude', epsilon_inf=1.2, plasma_rad_s=1.2e16, collision_rad_s=2e15)


@pytest.mark.parametrize('material', [LORENTZ, TWO_POLES, DRUDE], ids=['lorentz', 'two_poles', 'drude'])
def test_uniform_field_conserves_displacement(material):
    grid, state, d0, history = uniform_run(material, 400)
    assert np.abs(grid.H).max() == 0                 # zero curl throughout
    for _, displacement in history:
        np.testing.assert_allclose(displacement, d0, rtol=1e-13, atol=1e-13)
    assert state.multiple == (material is TWO_POLES)


@pytest.mark.parametrize('material', [LORENTZ, TWO_POLES], ids=['lorentz', 'two_poles'])
def test_damped_lorentz_relaxes_to_the_static_permittivity(material):
    poles = oscillators(material)
    static = material['epsilon_inf']+sum(s/(w0*w0) for w0, s, _ in poles)
    _, _, _, history = uniform_run(material, 6000)
    np.testing.assert_allclose(history[-1][0], np.array([1., -.5, .25])*material['epsilon_inf']/static, rtol=1e-9)


def test_drude_metal_screens_a_static_field():
    _, _, _, history = uniform_run(DRUDE, 3000)
    assert np.abs(history[-1][0]).max() < 1e-9*1.0


def test_single_cell_follows_the_discrete_permittivity():
    """Drive one sample with E(t) = cos(w t) (the ADE's free field is the prescribed E each step):
    after the transient the polarization is Re(chi(w~) e^{-i w t}) with chi the reference's discrete
    susceptibility, w~ = (2/dt) tan(w dt/2)."""
    class Grid:
        E = np.zeros((1, 1, 1, 3))

        class scene:
            dt = 2e-17
    material = dict(model='lorentz', epsilon_inf=1.0, resonance_rad_s=2e15, delta_epsilon=2.0, linewidth_rad_s=3e14)
    state = MaterialADE(Grid, material, [0])
    dt = Grid.scene.dt
    omega = 1.3e15
    w0, s, g = oscillators(material)[0]
    wt = 2/dt*math.tan(omega*dt/2)
    chi = s/(w0*w0-wt*wt-1j*g*wt)
    field = Grid.E
    for n in range(40000):
        old, response = state.prepare(field)
        field[...] = math.cos(omega*(n+1)*dt)
        # correct() solves for E; here E is imposed, so apply the polarization update with that E.
        new = field.reshape(-1, 3)[[0]]
        delta = response+state.k*(new+old)
        state.P += delta
        state.Q *= -1
        state.Q += 2*delta
    t = 40000*dt
    expected = (chi*np.exp(-1j*omega*t)).real      # the exp(-i omega t) convention of materials.py:1-5
    np.testing.assert_allclose(state.P[0, 0], expected, rtol=2e-6, atol=2e-6*abs(chi))
