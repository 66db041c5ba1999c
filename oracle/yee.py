"""Oracle: the Yee/CPML half steps in NumPy, restating torchfdtd/boundaries.py:363-438.

Fields are (Nx, Ny, Nz, 3) arrays, z fastest with interleaved components, exactly
the reference layout.  Every elementwise operation is a separate NumPy ufunc, so
each product and sum is rounded to the field precision in the reference's order:
  derivative = f[high] - f[low]
  psi = psi*b; psi = psi + c*derivative; derivative = derivative*inv_k + psi (in two roundings)
  curl[out] += sign*derivative           (terms in CURL_TERMS order)
  E += (S*inv_eps)*curl ;  H -= (S*inv_mu)*curl
Periodic wrap (real fields) and the PEC upper ghost node follow boundaries.py:384-393.
Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

from .scene import CURL_TERMS


def _sl(axis, value, component=None):
    parts = [slice(None)]*3
    parts[axis] = value
    if component is not None:
        parts.append(component)
    return tuple(parts)


class Grid:
    def __init__(self, scene, inverse_permittivity=None):
        self.scene = scene
        real = scene.real
        shape = scene.shape
        self.E = np.zeros((*shape, 3), real)
        self.H = np.zeros((*shape, 3), real)
        self.inverse_permittivity = np.ones((*shape, 3), real) if inverse_permittivity is None else inverse_permittivity
        self.inverse_permeability = np.ones((1,), real)
        self.courant = scene.courant
        self.cpml = scene.cpml_segments()
        self.wrap = {a for a in range(3) if shape[a] > 1 and scene.faces[a, 0].kind == 'periodic'}
        # boundaries.py:262-266: reference_step over the last cell width
        self.pec_upper = {a: float(scene.mesh/(scene.nodes[a][-1]-scene.nodes[a][-2]))
                          for a in range(3) if shape[a] > 1 and scene.faces[a, 1].kind == 'pec'}
        for (a, s), f in scene.faces.items():
            if shape[a] > 1 and f.kind not in ('pml', 'periodic', 'pec'):
                raise NotImplementedError('oracle: pml, periodic and pec faces only')

    @property
    def memory_states(self):
        return [s['psi'] for segs in self.cpml.values() for s in segs]

    def curl(self, field, forward):
        result = np.zeros_like(field)
        for axis, comp, out, sign in CURL_TERMS:
            if field.shape[axis] == 1:
                continue
            derivative = field[_sl(axis, slice(1, None), comp)]-field[_sl(axis, slice(None, -1), comp)]
            for seg in self.cpml[forward, axis, comp]:
                data = derivative[_sl(axis, slice(seg['lo']-seg['start'], seg['hi']-seg['start']))]
                psi = seg['psi']
                psi *= seg['b']
                psi += seg['c']*data
                data *= seg['inv_k']
                data += psi
            target = _sl(axis, slice(None, -1) if forward else slice(1, None), out)
            result[target] += sign*derivative
            if axis in self.wrap:
                first, last = _sl(axis, 0, comp), _sl(axis, -1, comp)
                edge = field[first]-field[last]
                result[_sl(axis, -1 if forward else 0, out)] += sign*edge
            elif forward and axis in self.pec_upper:
                edge = -self.pec_upper[axis]*field[_sl(axis, -1, comp)]
                result[_sl(axis, -1, out)] += sign*edge
        return result

    def update_E(self):
        curl = self.curl(self.H, False)
        self.E += self.courant*self.inverse_permittivity*curl

    def update_H(self):
        self.H -= self.courant*self.inverse_permeability*self.curl(self.E, True)
