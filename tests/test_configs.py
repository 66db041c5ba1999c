"""The benchmark configurations at their own sizes against the oracle (GPU).

The C restatement of the oracle (oracle/native.py, pinned bit for bit to the NumPy oracle and to
the reference-written golden runs by tests/test_oracle_native.py) steps the same seeded inputs on
the host, so the engine's fields and every CPML memory are compared bit for bit:

* the 512^3 roofline cube's workload (16-cell CPML, eps = 1 + 11 s from seeded smoothed noise, Ez
  dipole) through the two-step pass and its graphs, at 128^3 from rest with the source and at the
  full 512^3 from random fields and random CPML memory;
* the 160x120x40 waveguide (12-cell CPML, Ey dipole, point trace) over its full 1000 steps;
* the 64-grid uneven batch (64x64x32 .. 192x160x64, 10-cell CPML, binary design blocks, float64
  DFT planes) as one cohort with the tuned one-pass launch, its smallest, median and largest grid
  against the oracle (fields bit for bit, planes to 1e-12 of their largest entry).
"""
import math

import numpy as np
import pytest
import torch

from oracle import simulate
from oracle.native import NativeGrid
from oracle.scene import Scene

pytestmark = pytest.mark.gpu


def cube_oracle(project, mask):
    """NativeGrid of the cube project with its per-cell 1/eps (1 or the FP32 reciprocal of 12)."""
    scene = Scene(project.model_dump())
    inv = np.where(mask, np.float32(1)/np.float32(12), np.float32(1)).astype(np.float32)[..., None]
    return NativeGrid(scene, np.ascontiguousarray(inv))


def test_cube_128_two_step_graphs_match_oracle_bitwise(gpu):
    """The cube workload at 128^3 from rest: 44 steps of graph replays on the two-step pass (graphs of
    20 steps, one four, singles) against the oracle, source included."""
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.solver import Stepper
    from torchfdtd_amd.workloads import cube, cube_inverse_permittivity, cube_noise_mask
    n, steps = 128, 44
    project = cube(n=n, steps=steps)
    mask = cube_noise_mask(n)
    run = Stepper(project, two_step=True)
    assert run.two_step and run.one_pass
    run.grid.inverse_permittivity.copy_(cube_inverse_permittivity(mask)[..., None])
    graphs = StepGraphs(run.step, run.states(), steps, 20, warmup=1, pair=run.step_pair, prepare=run.refresh_materials,
                        quad=run.quad)
    for width in (20, 20, 4):
        graphs.replay(width)
    torch.cuda.synchronize()
    assert int(run.counter.item()) == steps
    ref = simulate.run(project.model_dump(), steps=steps, grid=cube_oracle(project, mask.cpu().numpy()),
                       record_frames=False)
    g = run.grid
    np.testing.assert_array_equal(g.E.cpu().numpy(), ref['E'])
    np.testing.assert_array_equal(g.H.cpu().numpy(), ref['H'])
    for got, expected in zip(g.memory_states, ref['grid'].memory_states):
        np.testing.assert_array_equal(got.cpu().numpy(), expected)
    assert float(np.abs(ref['E']).max()) > 0 and run.materials.count == 2


def test_cube_512_two_step_matches_oracle_bitwise(gpu):
    """The full 512^3 cube from random fields and random CPML memory of both families: two graph
    replays of four steps on the two-step pass against the oracle (every cell, every CPML row)."""
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.solver import Stepper
    from torchfdtd_amd.workloads import cube, cube_inverse_permittivity, cube_noise_mask
    n, steps = 512, 8
    project = cube(n=n, steps=40)
    mask = cube_noise_mask(n)
    run = Stepper(project)
    assert run.two_step
    g = run.grid
    g.inverse_permittivity.copy_(cube_inverse_permittivity(mask)[..., None])
    graphs = StepGraphs(run.step, run.states(), 40, 4, warmup=1, pair=run.step_pair, prepare=run.refresh_materials,
                        quad=run.quad)
    ref = cube_oracle(project, mask.cpu().numpy())
    del mask
    rng = np.random.default_rng(512)
    for name in ('E', 'H'):
        value = (rng.standard_normal((n, n, n, 3), dtype=np.float32)*np.float32(.02))
        getattr(ref, name)[...] = value
        getattr(g, name).copy_(torch.from_numpy(value))
        del value
    assert len(ref.memory_states) == len(g.memory_states) == 24
    for got, expected in zip(g.memory_states, ref.memory_states):
        expected[...] = rng.standard_normal(expected.shape, dtype=np.float32)*np.float32(.01)
        got.copy_(torch.from_numpy(expected))
    for _ in range(steps // 4):
        graphs.replay(4)
    simulate.run(project.model_dump(), steps=steps, grid=ref, record_frames=False)   # the Ez dipole included
    torch.cuda.synchronize()
    assert int(run.counter.item()) == steps
    for name in ('E', 'H'):
        np.testing.assert_array_equal(getattr(g, name).cpu().numpy(), getattr(ref, name))
    for got, expected in zip(g.memory_states, ref.memory_states):
        np.testing.assert_array_equal(got.cpu().numpy(), expected)


def test_waveguide_3d_1000_steps_matches_oracle_bitwise(gpu):
    """The 160x120x40 waveguide (12-cell CPML, Ey dipole, point trace of Ey) for its full 1000 steps."""
    from torchfdtd_amd import Simulation
    from torchfdtd_amd.workloads import waveguide_3d
    project = waveguide_3d()
    got = Simulation(project).run(cuda_graph_steps=50)
    ref = simulate.run(project.model_dump(), native=True)
    np.testing.assert_array_equal(got.electric, ref['E'])
    np.testing.assert_array_equal(got.magnetic, ref['H'])
    np.testing.assert_array_equal(got.signals, ref['signals'])
    np.testing.assert_array_equal(got.frames, ref['frames'])
    assert got.summary['steps'] == 1000 and np.abs(got.signals).max() > 1e-3


def test_uneven_batch_grids_match_oracle(gpu):
    """All 64 grids of the uneven batch as one cohort (tuned one-pass launch: 8-row tiles, 6-plane chunks)
    for 600 steps; the smallest, a median and the largest grid against the oracle."""
    from torchfdtd_amd import workloads as W
    from torchfdtd_amd.batch import Launch, run_tensor_batch
    steps = 600
    cases = W.uneven_batch(64, steps)
    cells = [math.prod(c.project.region.shape) for c in cases]
    order = np.argsort(cells, kind='stable')
    picks = [int(order[0]), int(order[len(order)//2]), int(order[-1])]
    report = run_tensor_batch(cases, launch=Launch('march', 8, 6), cuda_graph_steps=50)
    assert report.plan['launch'] == Launch('march', 8, 6).as_dict()
    for i in picks:
        case, item = cases[i], report.items[i]
        assert item.status == 'completed', item.id
        project = case.project.model_dump()
        scene = Scene(project)
        inv = (1/case.epsilon.astype(np.float32))[..., None]
        grid = NativeGrid(scene, np.ascontiguousarray(np.broadcast_to(inv, (*scene.shape, 3))))
        ref = simulate.run(project, grid=grid, record_frames=False)
        res = item.result
        np.testing.assert_array_equal(res.electric, ref['E'], err_msg=item.id)
        np.testing.assert_array_equal(res.magnetic, ref['H'], err_msg=item.id)
        (plane,), (expected,) = res.frequency_fields, ref['planes']
        names = list(expected.components)
        values = expected.value[..., [names.index(c) for c in plane['components']]]
        scale = np.abs(values).max()
        assert scale > 0, item.id
        np.testing.assert_allclose(plane['fields']/scale, values/scale, rtol=1e-12, atol=1e-12, err_msg=item.id)
