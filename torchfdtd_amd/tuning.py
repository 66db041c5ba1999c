"""Measured step path and launch shape for one grid (the reference's select_launch / tune_yee,
torchfdtd/reversible_cuda_tuning.py:84-229).

`tune_launch(stepper)` times every candidate the grid can take -- the split step, the one-pass step
(rows per tile, planes per x chunk) on one-pass grids, the two-step pass (wavefronts per tile, planes
per x chunk) on float32 one-pass grids without per-step observers -- and installs the fastest.  As
in the reference:

* every candidate must give bitwise-identical physics: each advances the same saved state by one
  unit of four steps (four split steps, two one-pass pairs or one two-step quad) and its fields and
  CPML memory must equal the first candidate's bit for bit, or it is rejected;
* timing runs on the grid's own buffers from a saved copy of the state, which is restored
  afterwards, so the run continues from where it was; sweeps alternate the candidate order and the
  median of each candidate's samples decides;
* tuning must precede graph capture (graphs hold the descriptors of one path and shape);
* results are cached (LRU, 128 entries) by device, grid shape, precision, CPML layout, permittivity
  layout, source count and the paths available, and a cached choice is installed without timing.

The z extent of a tile is not a candidate: a tile's z column is one wavefront of 64 lanes (61 owned
cells in the two-step pass), the layout every variant's LDS staging is built around.
"""
from __future__ import annotations

import statistics
from collections import OrderedDict

import torch

# Two-step shapes (wavefronts, planes per chunk).  24-plane chunks time 0.4% faster than 20 in the
# tuner's eager units on the 512^3 cube but ran 0.9% slower in its graphed bench (1.386 against
# 1.374 ms per step), so they are not offered.
TWO_STEP = tuple((w, c) for w in (12, 16) for c in (12, 16, 20, 28))
ONE_PASS = tuple((t, c) for t in (8, 16) for c in (2, 3, 4, 6, 8))
NAMES = {'one_pass': ('tile_y', 'chunk_x'), 'two_step': ('waves2', 'chunk_x2'), 'split': ()}
UNIT_STEPS = 4
_CACHE: OrderedDict = OrderedDict()
_CACHE_SIZE = 128


def _key(stepper):
    g = stepper.grid
    faces = tuple(sorted((f, a, c, e['lo'], e['hi']) for (f, a, c), entries in g.cpml.items() for e in entries))
    return (torch.cuda.get_device_name(g.device), tuple(g.E.shape), str(g.dtype), g.inverse_permittivity.shape[-1],
            faces, len(stepper.sources['E']), len(stepper.sources['H']), tuple(stepper.paths()),
            len(stepper.planes), len(stepper.trace_points) > 0)


def candidates_for(stepper):
    """Every (path, shape) the grid can take, split first."""
    paths = stepper.paths()
    out = [('split', {})]
    if 'one_pass' in paths:
        out += [('one_pass', dict(zip(NAMES['one_pass'], c))) for c in ONE_PASS]
    if 'two_step' in paths:
        out += [('two_step', dict(zip(NAMES['two_step'], c))) for c in TWO_STEP]
    return out


def _normalise(stepper, candidates):
    """Candidates as (path, shape) pairs: plain pairs of numbers are shapes of the stepper's current path."""
    out = []
    for c in candidates:
        if isinstance(c[0], str):
            out.append((c[0], dict(c[1]) if len(c) > 1 else {}))
        else:
            out.append((stepper.path, dict(zip(NAMES[stepper.path], c))))
    return out


def _unit(stepper):
    """Four steps on the stepper's current path; the state returns to the grid's own buffers."""
    if stepper.two_step:
        stepper.quad()
    elif stepper.one_pass:
        stepper.step_pair()
        stepper.step_pair()
    else:
        for _ in range(UNIT_STEPS):
            stepper.step()


def _install(stepper, path, shape):
    if stepper.path != path:
        stepper.set_path(path)
    if shape:
        stepper.retile(**shape)


def _state(stepper):
    g = stepper.grid
    return [g.E, g.H, stepper.counter, stepper.traces, *g.memory_states]


def tune_launch(stepper, candidates=None, sweeps=5, units=2, report=None):
    """Install the fastest bitwise-identical (path, launch shape) of `stepper` and return its record:
    {'path', 'chosen', 'median_ms' per candidate (per step), 'rejected', 'cached'}."""
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError('launch tuning must precede graph capture')
    if stepper.dispersive:
        record = dict(path='split', chosen={}, median_ms={}, rejected=[], cached=False)
        if report is not None:
            report['launch_tuning'] = record
        return record
    key = _key(stepper)
    if candidates is None and key in _CACHE:
        _CACHE.move_to_end(key)
        record = dict(_CACHE[key], cached=True)
        _install(stepper, record['path'], record['chosen'])
        if report is not None:
            report['launch_tuning'] = record
        return record
    trials = _normalise(stepper, candidates) if candidates is not None else candidates_for(stepper)
    saved = [t.clone() for t in _state(stepper)]
    original_shape = stepper.launch_shape()
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

    reference, accepted, rejected = None, [], []
    for path, shape in trials:
        _install(stepper, path, shape)
        restore()
        start()
        _unit(stepper)
        torch.cuda.synchronize()
        result = [t.clone() for t in (stepper.grid.E, stepper.grid.H, stepper.traces, *stepper.grid.memory_states)]
        if reference is None:
            reference = result
            accepted.append((path, shape))
        elif all(torch.equal(a, b) for a, b in zip(result, reference)):
            accepted.append((path, shape))
        else:
            rejected.append(dict(path=path, **shape))
        del result
    samples = {i: [] for i in range(len(accepted))}
    for sweep in range(sweeps):
        order = list(range(len(accepted)))
        if sweep % 2:
            order.reverse()
        for i in order:
            _install(stepper, *accepted[i])
            restore()
            start()
            _unit(stepper)      # first launch of this path and shape outside the timed region
            begin, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            begin.record()
            for _ in range(units):
                _unit(stepper)
            end.record()
            end.synchronize()
            samples[i].append(begin.elapsed_time(end)/(units*UNIT_STEPS))
    medians = {i: statistics.median(v) for i, v in samples.items()}
    best = min(medians, key=medians.get)
    path, shape = accepted[best]
    stepper.retile(**original_shape)
    _install(stepper, path, shape)
    restore()
    torch.cuda.synchronize()
    del saved, trial

    def label(p, sh):
        return ','.join([p, *(f'{k}={v}' for k, v in sh.items())])
    record = dict(path=path, chosen=shape, cached=False, rejected=rejected,
                  median_ms={label(*accepted[i]): round(medians[i], 4) for i in range(len(accepted))})
    if candidates is None:
        _CACHE[key] = {k: v for k, v in record.items() if k != 'cached'}
        while len(_CACHE) > _CACHE_SIZE:
            _CACHE.popitem(last=False)
    if report is not None:
        report['launch_tuning'] = record
    return record
