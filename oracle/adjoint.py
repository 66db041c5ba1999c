"""Oracle: the discrete adjoint of the Yee/CPML step in NumPy -- TEST INFRASTRUCTURE ONLY.

Restates the reference's exact transpose (torchfdtd/differentiable.py:395-447,
``curl_faces_transpose``, and 496-519, ``transpose_step``) for real fields without PMC
faces, on the oracle grid of ``oracle/yee.py``:

  curl transpose, per curl term in CURL_TERMS order:
      d = sign * bar[out rows]
      in a CPML slab:  m = psi_bar + d;  psi_bar_prev = b*m;  d = d*inv_k + c*m
      input[high rows] += d;  input[low rows] -= d
      periodic edge / PEC-upper edge transposed likewise
  step transpose (E update, then H update, run backwards):
      e_bar += curl_T(forward)(-S*inv_mu*h_bar)          (H update)
      gradient = -S * e_bar * curl_H(n) / eps^2             (E update, d/d eps)
      h_bar += curl_T(backward)(S*inv_eps*e_bar)            (E update)

The forward step here is E += (S*inv_eps)*curl: its exact transpose uses S*inv_eps, the
reference's torch transpose writes S/eps (the same operator).  Pinned by the dot-product
identity <A x, y> = <x, A^T y> on random states (FP64) and by central finite differences
of FP64 forward runs (tests/test_oracle_adjoint.py).
"""
from __future__ import annotations

import numpy as np

from .scene import CURL_TERMS
from .yee import _sl


def state(grid):
    """(E, H, [psi of every segment in cpml key order]) -- the step's state, views."""
    return grid.E, grid.H, [s['psi'] for key in sorted(grid.cpml) for s in grid.cpml[key]]


def zero_adjoint(grid):
    return (np.zeros_like(grid.E), np.zeros_like(grid.H),
            {key: [np.zeros_like(s['psi']) for s in segs] for key, segs in grid.cpml.items()})


def curl_transpose(grid, bar, psi_bar, forward):
    """Adjoint of Grid.curl(., forward) for the output adjoint `bar` (differentiable.py:395-447).

    psi_bar[(forward, axis, comp)][i] is the adjoint of the segment's updated memory; returns
    (adjoint of the curl's input field, {key: previous psi adjoints}) for this family."""
    result = np.zeros_like(bar)
    previous = {key: [p.copy() for p in values] for key, values in psi_bar.items()}
    for axis, comp, out, sign in CURL_TERMS:
        n = bar.shape[axis]
        if n == 1:
            continue
        derivative = sign*bar[_sl(axis, slice(0, n-1) if forward else slice(1, n), out)]
        for i, seg in enumerate(grid.cpml[forward, axis, comp]):
            sl = _sl(axis, slice(seg['lo']-seg['start'], seg['hi']-seg['start']))
            value = derivative[sl]
            memory_bar = psi_bar[forward, axis, comp][i]+value
            previous[forward, axis, comp][i] = seg['b']*memory_bar
            derivative = derivative.copy()
            derivative[sl] = value*seg['inv_k']+seg['c']*memory_bar
        result[_sl(axis, slice(1, n), comp)] += derivative
        result[_sl(axis, slice(0, n-1), comp)] -= derivative
        if axis in grid.wrap:
            edge = sign*bar[_sl(axis, n-1 if forward else 0, out)]
            result[_sl(axis, 0, comp)] += edge
            result[_sl(axis, n-1, comp)] -= edge
        elif forward and axis in grid.pec_upper:
            result[_sl(axis, n-1, comp)] -= sign*grid.pec_upper[axis]*bar[_sl(axis, n-1, out)]
    return result, previous


def curl_with_memory(grid, field, forward):
    """Grid.curl without touching the grid's CPML memory (the value the step used)."""
    saved = [s['psi'].copy() for segs in grid.cpml.values() for s in segs]
    try:
        return grid.curl(field, forward)
    finally:
        for s, value in zip((s for segs in grid.cpml.values() for s in segs), saved):
            s['psi'][...] = value


def transpose_step(grid, H_n, e_bar, h_bar, psi_bar, epsilon):
    """One step backwards (differentiable.py:496-519): returns (e_bar, h_bar, psi_bar, gradient).

    The grid must hold the CPML memory of time n (the state the E update of step n read);
    H_n is the magnetic field of time n.  epsilon has the shape of inverse_permittivity's
    cells: (Nx,Ny,Nz) or (Nx,Ny,Nz,3); the gradient has the same shape."""
    S = grid.courant
    contribution, psi_bar = curl_transpose(grid, -S*grid.inverse_permeability*h_bar, psi_bar, True)
    e_bar = e_bar+contribution
    curl_e = curl_with_memory(grid, H_n, False)
    eps4 = epsilon[..., None] if epsilon.ndim == 3 else epsilon
    full = -S*(e_bar*curl_e)/(eps4*eps4)
    gradient = full.sum(axis=-1) if epsilon.ndim == 3 else full
    contribution, psi_bar = curl_transpose(grid, S*grid.inverse_permittivity*e_bar, psi_bar, False)
    return e_bar, h_bar+contribution, psi_bar, gradient


def run_with_gradient(scene, epsilon, steps, observers, weights, sources=None):
    """J = sum_t sum_k weights[t, k] * field_k(t) and dJ/d epsilon by the reverse sweep.

    observers: list of (family 'E'|'H', flat element index); field_k(t) is sampled after
    step t (both half steps and their sources).  sources: soft sources as
    Scene.source_terms() gives them, [(component, location, waveform[steps])].  Returns
    (J, gradient, signals).  Forward states are stored whole (tiny grids only)."""
    from .yee import Grid
    eps4 = epsilon[..., None] if epsilon.ndim == 3 else epsilon
    real = scene.real
    inverse = np.broadcast_to((1/eps4).astype(real), (*scene.shape, 3)).copy()
    grid = Grid(scene, inverse)
    sources = sources or []
    history, signals = [], np.zeros((steps, len(observers)), real)
    for t in range(steps):
        history.append((grid.H.copy(), [s['psi'].copy() for segs in grid.cpml.values() for s in segs]))
        grid.update_E()
        for comp, loc, wave in sources:
            if comp[0] == 'E':
                grid.E[loc+('xyz'.index(comp[1].lower()),)] += wave[t]
        grid.update_H()
        for comp, loc, wave in sources:
            if comp[0] == 'H':
                grid.H[loc+('xyz'.index(comp[1].lower()),)] += wave[t]
        for k, (family, index) in enumerate(observers):
            signals[t, k] = (grid.E if family == 'E' else grid.H).reshape(-1)[index]
    J = float(np.sum(weights*signals))
    e_bar, h_bar, psi_bar = zero_adjoint(grid)
    gradient = np.zeros(epsilon.shape, np.float64)
    for t in range(steps-1, -1, -1):
        for k, (family, index) in enumerate(observers):
            (e_bar if family == 'E' else h_bar).reshape(-1)[index] += weights[t, k]
        H_n, psis = history[t]
        for s, value in zip((s for segs in grid.cpml.values() for s in segs), psis):
            s['psi'][...] = value
        e_bar, h_bar, psi_bar, part = transpose_step(grid, H_n, e_bar, h_bar, psi_bar, epsilon)
        gradient += part
    return J, gradient, signals
