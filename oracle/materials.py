"""Oracle: passive Drude/Lorentz materials by the coupled trapezoidal ADE of
torchfdtd/materials.py:29-87 (MaterialADE), in NumPy.  Test infrastructure only.

Per owned sample and oscillator j (rates w0_j, strength s_j, damping g_j; P relative polarization,
Q = dt dP/dt):  d_j = 1 + g_j dt/2 + (w0_j dt)^2/4,  a_j = (w0_j dt)^2/2,  k_j = s_j dt^2/(4 d_j);
around the E update:
    old = E;  response_j = (Q_j - a_j P_j)/d_j          (before)
    free = E after the curl update
    new = (eps_inf free - k_sum old - sum_j response_j)/(eps_inf + k_sum)
    delta_j = response_j + k_j (new + old);  P_j += delta_j;  Q_j = -Q_j + 2 delta_j;  E = new
A single oscillator keeps the reference's scalar coefficients (Python floats, so eps_inf + k is
formed in double and rounded once); several hold them as arrays in the field precision
(materials.py:41-60), which changes the rounding of k_sum and eps_inf + k_sum.
"""
from __future__ import annotations

import numpy as np


def oscillators(material):
    """(w0, strength, damping) per pole of a material dict (models.py:161-174)."""
    model = material.get('model', 'dielectric')
    if model == 'drude':
        return [(0., material.get('plasma_rad_s', 2e15)**2, material.get('collision_rad_s', 1e14))]
    if model == 'lorentz':
        w0 = material.get('resonance_rad_s', 2e15)
        return [(w0, material.get('delta_epsilon', 1.)*w0**2, 2*material.get('linewidth_rad_s', 1e14))]
    if model == 'multipole':
        return [(p['resonance_rad_s'], p['strength_rad_s_squared'], p['damping_rad_s']) for p in material.get('poles', [])]
    return []


class MaterialADE:
    def __init__(self, grid, material, indices, components=False):
        self.indices = np.asarray(indices, dtype=np.int64)
        self.components = components
        real = grid.E.dtype
        sample_shape = (len(indices),) if components else (len(indices), 3)
        self.oscillators = oscillators(material)
        if not self.oscillators:
            raise ValueError('MaterialADE requires at least one oscillator.')
        self.multiple = len(self.oscillators) > 1
        shape = (len(self.oscillators), *sample_shape) if self.multiple else sample_shape
        self.P = np.zeros(shape, real)
        self.Q = np.zeros(shape, real)
        rates = np.asarray(self.oscillators, dtype=float)
        if self.multiple:
            rates = rates.reshape(len(rates), 3, *([1]*len(sample_shape)))
            w0, strength, gamma = (rates[:, j] for j in range(3))
        else:
            w0, strength, gamma = rates[0]
        dt = grid.scene.dt
        denominator = 1+.5*gamma*dt+.25*(w0*dt)**2

        def coefficient(value):
            return np.asarray(value, dtype=real) if self.multiple else float(value)
        self.a = coefficient(.5*(w0*dt)**2)
        self.d = coefficient(denominator)
        self.k = coefficient(strength*dt*dt/(4*denominator))
        self.k_sum = self.k.sum(axis=0) if self.multiple else self.k
        self.eps = float(material.get('epsilon_inf', 1.))

    def take(self, field):
        flat = field.reshape(-1) if self.components else field.reshape(-1, 3)
        return flat[self.indices]

    def prepare(self, field):
        return self.take(field), (self.Q-self.a*self.P)/self.d

    def correct(self, field, old, response):
        free = self.take(field)
        total = response.sum(axis=0) if self.multiple else response
        new = (self.eps*free-self.k_sum*old-total)/(self.eps+self.k_sum)
        delta = response+self.k*(new+old)
        self.P += delta
        self.Q *= -1
        self.Q += 2*delta
        flat = field.reshape(-1) if self.components else field.reshape(-1, 3)
        flat[self.indices] = new


def configure(grid, scene, ownership):
    """One ADE state per dispersive material that owns samples ('ade' PML mode: every owned sample,
    PML cells included; materials.py:124-200 without the frozen and absorber modes)."""
    listed = scene.project.get('materials', [])
    states = []
    for i, m in enumerate(listed):
        if not oscillators(m) or not np.any(ownership == i):
            continue
        states.append(MaterialADE(grid, m, np.flatnonzero(ownership.reshape(-1) == i), ownership.ndim == 4))
    return states
