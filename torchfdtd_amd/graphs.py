"""Whole-step HIP graphs and the host observation schedule (torchfdtd/cuda_graph.py:12-63).

A step (E, H, plane DFTs, traces, counter) is captured once for 1 step and once for
`width` steps; replays advance the device counter, so sources and monitors read
the right sample in every replay.  Captures record kernel launches only.
"""
from __future__ import annotations

import torch


def validate_graph_steps(value, enabled=True):
    if isinstance(value, bool) or not isinstance(value, int) or not 1 <= value <= 64:
        raise ValueError('cuda_graph_steps must be an integer from 1 to 64.')
    if value != 1 and not enabled:
        raise ValueError('cuda_graph_steps > 1 requires CUDA graph execution.')
    return value


def observation_schedule(steps, width, barriers=()):
    """Yield (completed steps, replay width), stopping exactly at every host barrier."""
    completed = 0
    for end in sorted(set(barriers) | {steps}):
        if not 0 < end <= steps:
            raise ValueError('Observation barriers must lie within the run.')
        while completed < end:
            advance = width if end-completed >= width else 1
            completed += advance
            yield completed, advance


class StepGraphs:
    """Capture `step` for 1 and `width` steps on a side stream; `states` are zeroed after warm-up and capture.

    With `pair` (two steps that return the state to its own buffers) the `width` graph
    is built from pairs, plus one `step` when `width` is odd; with `quad` as well (four steps
    that return the state to its own buffers) from quads first, and a graph of one quad is
    captured too (`replay(4)`).  `prepare`, when given, runs
    before every replay (e.g. Stepper.refresh_materials, which follows permittivity changes).
    """

    def __init__(self, step, states, steps, width=1, warmup=3, pair=None, prepare=None, quad=None, widths=()):
        """`widths`: further step counts to capture a graph of each (<= 64), for callers that advance
        by counts other than 1 and `width` (a checkpoint segment and its remainder)."""
        self.prepare = prepare
        self.width = min(validate_graph_steps(width), steps)
        states = list(states)
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            for _ in range(min(warmup, steps)):
                step()
            if pair is not None and steps >= 2:
                pair()      # the pair's first launches happen outside capture too
            if quad is not None and steps >= 4:
                quad()
            for value in states:
                value.zero_()
        torch.cuda.current_stream().wait_stream(stream)
        torch.cuda.synchronize()
        self.graphs = {}
        self.kernel_steps = dict(two_step=0, one_pass=0, split=0)     # steps replayed through each path
        self.composition = {}
        counts = {1, self.width} | {min(int(w), steps) for w in widths if 1 <= int(w) <= 64}
        if quad is not None and self.width > 4:
            counts.add(4)       # one four: replays that stop at any multiple of four stay on the quad path
        for count in sorted(counts):
            graph = torch.cuda.CUDAGraph()
            made = dict(two_step=0, one_pass=0, split=0)
            with torch.cuda.stream(stream):
                graph.capture_begin()
                if pair is not None and count > 1:
                    rest = count
                    if quad is not None:
                        for _ in range(rest // 4):
                            quad()
                        made['two_step'] = 4*(rest // 4)
                        rest %= 4
                    for _ in range(rest // 2):
                        pair()
                    made['one_pass'] = 2*(rest // 2)
                    if count % 2:
                        step()
                        made['split'] = 1
                else:
                    for _ in range(count):
                        step()
                    made['split'] = count
                graph.capture_end()
            self.composition[count] = made
            self.graphs[count] = graph
            for value in states:
                value.zero_()
        torch.cuda.synchronize()

    def advance(self, count):
        """`count` steps by the widest captured graphs first (greedy over the captured widths)."""
        widths = sorted(self.graphs, reverse=True)
        while count > 0:
            width = next(w for w in widths if w <= count)
            self.replay(width)
            count -= width

    def replay(self, width=1):
        if self.prepare is not None:
            self.prepare()
        self.graphs[width].replay()
        for path, steps in self.composition[width].items():
            self.kernel_steps[path] += steps
