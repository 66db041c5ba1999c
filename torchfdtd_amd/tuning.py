"""Measured launch shapes for one grid (the reference's select_launch / tune_yee,
torchfdtd/reversible_cuda_tuning.py:84-229).

`tune_launch(stepper)` times every candidate launch shape of the grid's step path -- the two-step
pass (wavefronts per tile, planes per x chunk) for float32 grids without per-step observers, else
the one-pass step (rows per tile, planes per x chunk) -- and installs the fastest.  As in the
reference:

* every candidate must give bitwise-identical physics: each advances the same saved state by one
  unit (four steps of the two-step pass, two of the one-pass step) and its fields and CPML memory
  must equal the first candidate's bit for bit, or it is rejected;
* timing runs on the grid's own buffers from a saved copy of the state, which is restored
  afterwards, so the run continues from where it was; sweeps alternate the candidate order and the
  median of each candidate's samples decides;
* tuning must precede graph capture (graphs hold the descriptors of one shape);
* results are cached (LRU, 128 entries) by device, grid shape, precision, CPML layout, permittivity
  layout, source count and path, and a cached shape is installed without timing.

The z extent of a tile is not a candidate: a tile's z column is one wavefront of 64 lanes (61 owned
cells in the two-step pass), the layout every variant's LDS staging is built around.
"""
from __future__ import annotations

import statistics
from collections import OrderedDict

import torch

TWO_STEP = tuple((w, c) for w in (12, 16) for c in (12, 16, 20, 28))
ONE_PASS = tuple((t, c) for t in (8, 16) for c in (4, 6, 8))
_CACHE: OrderedDict = OrderedDict()
_CACHE_SIZE = 128


def _key(stepper):
    g = stepper.grid
    faces = tuple(sorted((f, a, c, e['lo'], e['hi']) for (f, a, c), entries in g.cpml.items() for e in entries))
    return (torch.cuda.get_device_name(g.device), tuple(g.E.shape), str(g.dtype), g.inverse_permittivity.shape[-1],
            faces, len(stepper.sources['E']), len(stepper.sources['H']), stepper.two_step, stepper.one_pass,
            stepper.materials is not None and stepper.materials.count <= 256)


def _unit(stepper):
    """One timing unit: a quad of the two-step pass or a pair of one-pass steps (the state returns to
    the grid's own buffers)."""
    return (stepper.quad, 4) if stepper.two_step else (stepper.step_pair, 2)


def _state(stepper):
    g = stepper.grid
    return [g.E, g.H, stepper.counter, stepper.traces, *g.memory_states]


def tune_launch(stepper, candidates=None, sweeps=5, units=2, report=None):
    """Install the fastest bitwise-identical launch shape of `stepper`'s path and return its record:
    {'path', 'chosen', 'median_ms' per candidate (per step), 'rejected', 'cached'}."""
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError('launch tuning must precede graph capture')
    if not stepper.one_pass:
        record = dict(path='split', chosen=None, median_ms={}, rejected=[], cached=False)
        if report is not None:
            report['launch_tuning'] = record
        return record
    path = 'two_step' if stepper.two_step else 'one_pass'
    names = ('waves2', 'chunk_x2') if stepper.two_step else ('tile_y', 'chunk_x')
    key = _key(stepper)
    if candidates is None and key in _CACHE:
        _CACHE.move_to_end(key)
        record = dict(_CACHE[key], cached=True)
        stepper.retile(**record['chosen'])
        if report is not None:
            report['launch_tuning'] = record
        return record
    shapes = [dict(zip(names, c)) for c in (candidates or (TWO_STEP if stepper.two_step else ONE_PASS))]
    saved = [t.clone() for t in _state(stepper)]
    original = stepper.launch_shape()
    # Candidates start from a seeded random state (fields and CPML memory), so that the identity check
    # sees every cell and slab row at work; the run's own state is restored at the end.
    gen = torch.Generator(device=stepper.grid.device).manual_seed(20260514)
    g = stepper.grid
    trial = [torch.randn(t.shape, generator=gen, device=t.device, dtype=t.dtype)*.01 if t.is_floating_point() else t.clone()
             for t in (g.E, g.H, *g.psi_buffers.values())]
    trial_targets = [g.E, g.H, *g.psi_buffers.values()]

    def start():
        for target, value in zip(trial_targets, trial):
            target.copy_(value)

    def restore():
        for target, value in zip(_state(stepper), saved):
            target.copy_(value)

    run, steps = _unit(stepper)
    reference, accepted, rejected = None, [], []
    for shape in shapes:
        stepper.retile(**shape)
        restore()
        start()
        run()
        torch.cuda.synchronize()
        result = [t.clone() for t in (stepper.grid.E, stepper.grid.H, *stepper.grid.memory_states)]
        if reference is None:
            reference = result
            accepted.append(shape)
        elif all(torch.equal(a, b) for a, b in zip(result, reference)):
            accepted.append(shape)
        else:
            rejected.append(shape)
        del result
    samples = {i: [] for i in range(len(accepted))}
    for sweep in range(sweeps):
        order = list(range(len(accepted)))
        if sweep % 2:
            order.reverse()
        for i in order:
            stepper.retile(**accepted[i])
            restore()
            start()
            run()      # first launch of this shape outside the timed region
            begin, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            begin.record()
            for _ in range(units):
                run()
            end.record()
            end.synchronize()
            samples[i].append(begin.elapsed_time(end)/(units*steps))
    medians = {i: statistics.median(v) for i, v in samples.items()}
    best = min(medians, key=medians.get)
    chosen = {**original, **accepted[best]}
    stepper.retile(**chosen)
    restore()
    torch.cuda.synchronize()
    del saved, trial
    record = dict(path=path, chosen=accepted[best], cached=False, rejected=rejected,
                  median_ms={','.join(f'{k}={v}' for k, v in shape.items()): round(medians[i], 4)
                             for i, shape in enumerate(accepted)})
    if candidates is None:
        _CACHE[key] = {k: v for k, v in record.items() if k != 'cached'}
        while len(_CACHE) > _CACHE_SIZE:
            _CACHE.popitem(last=False)
    if report is not None:
        report['launch_tuning'] = record
    return record
