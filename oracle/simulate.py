"""Oracle: the forward time loop of torchfdtd/solver.py:785-797 and 844-907 on NumPy.

Per step: update_E, inject E, update_H, inject H, plane DFT update, trace record,
counter += 1.  Frames follow the reference's barrier schedule (every
max(snapshot_interval, ceil(steps/100)) steps and the last).  Test infrastructure only.
"""
from __future__ import annotations

import math

import numpy as np

from .monitors import FrequencyPlane, point_spectrum, spectrum_settings
from .scene import Scene
from .yee import Grid


def prepare(project, native=False):
    """(scene, eps, grid); native=True steps the grid with the C restatement (oracle/native.py).
    Dispersive materials get their ADE states in grid.material_states (oracle/materials.py)."""
    from .materials import configure
    scene = Scene(project)
    eps, ownership, _ = scene.voxelize(with_ownership=True)
    inverse = np.ones((*scene.shape, 3), scene.real)
    inverse[:] = 1/(eps if eps.ndim == 4 else eps[..., None])
    if native:
        from .native import NativeGrid
        grid = NativeGrid(scene, inverse)
    else:
        grid = Grid(scene, inverse)
    grid.material_states = configure(grid, scene, ownership)
    return scene, eps, grid


def run(project, steps=None, grid=None, record_frames=True, native=False, rule=None, check_interval=50):
    """Step the project; with `rule` (oracle.run_control.Rule) the FP64 state norm is judged every
    `check_interval` steps and at the last step, stopping the loop where the rule stops the run
    (solver.py:844-881)."""
    if grid is None:
        scene, eps, g = prepare(project, native=native)
    else:
        scene, g = grid.scene, grid
        eps = scene.voxelize()
    steps = scene.steps if steps is None else steps
    terms = scene.source_terms()
    sources = {'E': [], 'H': []}
    for comp, loc, values in terms:
        sources[comp[0]].append((values, loc, 'xyz'.index(comp[1].lower())))
    monitors = [m for m in project.get('monitors', []) if m.get('enabled', True)]
    points = [m for m in monitors if m.get('kind', 'point') == 'point']
    planes = [FrequencyPlane(scene, m) for m in monitors if m.get('kind', 'point') == 'field']
    locs = [(scene.index_at(m.get('center', (0, 0, 0)), m.get('component', 'Ez')), m.get('component', 'Ez')) for m in points]
    traces = np.zeros((steps, len(points)), scene.real)
    axis = 'xyz'.index(scene.slice_axis)
    comp = 'xyz'.index(scene.field[1].lower())
    plane_index = scene.index_at(tuple(scene.slice_position if i == axis else 0 for i in range(3)), scene.field)[axis]
    interval = max(scene.snapshot_interval, math.ceil(scene.steps/100))
    frames, frame_steps = [], []

    def inject(family, n):
        field = g.E if family == 'E' else g.H
        for values, loc, c in sources[family]:
            field[loc+(c,)] += values[n]

    states = getattr(g, 'material_states', [])
    for n in range(steps):
        # materials.py / cuda_kernels.py:355-362: the ADE wraps the E update, before the E sources.
        prepared = [state.prepare(g.E) for state in states]
        g.update_E()
        for state, (old, response) in zip(states, prepared):
            state.correct(g.E, old, response)
        inject('E', n)
        g.update_H()
        inject('H', n)
        for p in planes:
            p.update(g.E, g.H, n)
        for k, (loc, c) in enumerate(locs):
            traces[n, k] = (g.E if c[0] == 'E' else g.H)[loc+('xyz'.index(c[1].lower()),)]
        done = n+1
        stop = False
        if rule is not None and (done % check_interval == 0 or done == steps):
            from .run_control import state_norm
            energy, peak, finite = state_norm(g.E, g.H, g.inverse_permittivity, g.memory_states, states, scene.dt)
            if not finite:
                raise FloatingPointError('Non-finite CPML or incident-line memory detected.')
            stop = rule.update(done, energy, peak)
        if record_frames and (done % interval == 0 or done == steps or stop):
            sl = [slice(None)]*3
            sl[axis] = plane_index
            data = (g.E if scene.field[0] == 'E' else g.H)[..., comp][tuple(sl)]
            frames.append(data[::max(1, math.ceil(data.shape[0]/256)), ::max(1, math.ceil(data.shape[1]/256))].copy())
            frame_steps.append(done)
        if stop:
            steps = done
            traces = traces[:done]
            break
    times = np.arange(1, steps+1)*scene.dt
    spectra = []
    for k, m in enumerate(points):
        shift = scene.dt/2 if m.get('component', 'Ez').startswith('H') else 0
        ds = m.get('time_downsample', 1)
        spectra.append(point_spectrum((times+shift)[::ds], traces[::ds, k], spectrum_settings(m.get('spectrum'))))
    return dict(scene=scene, grid=g, epsilon=eps, E=g.E, H=g.H, signals=traces, times=times,
                frames=np.array(frames), frame_steps=np.array(frame_steps), spectra=spectra, planes=planes)
