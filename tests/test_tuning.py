"""Measured launch shapes of one grid (tuning.tune_launch; the reference's select_launch / tune_yee,
torchfdtd/reversible_cuda_tuning.py:84-229): candidates are timed from a saved state that is
restored afterwards, every accepted shape gives the same physics bit for bit, the choice is cached,
and a tuned run equals an untuned one."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def cube_stepper(n=96, steps=40, **kw):
    from torchfdtd_amd.solver import Stepper
    from torchfdtd_amd.workloads import cube, cube_inverse_permittivity, cube_noise_mask
    project = cube(n=n, steps=steps)
    run = Stepper(project, **kw)
    run.grid.inverse_permittivity.copy_(cube_inverse_permittivity(cube_noise_mask(n))[..., None])
    return run


def advance(run, fours):
    for _ in range(fours):
        if run.two_step:
            run.quad()
        elif run.one_pass:
            run.step_pair()
            run.step_pair()
        else:
            for _ in range(4):
                run.step()
    torch.cuda.synchronize()


@pytest.mark.parametrize('path', ['two_step', 'one_pass'])
def test_tuned_launch_keeps_the_state_and_the_physics(gpu, path):
    from torchfdtd_amd import tuning
    kw = dict(one_pass=True) if path == 'two_step' else dict(one_pass=True, two_step=False)
    tuned, plain = cube_stepper(**kw), cube_stepper(**kw)
    assert tuned.two_step == (path == 'two_step')
    advance(tuned, 2)
    advance(plain, 2)
    before = tuned.grid.E.clone()
    candidates = ((12, 12), (12, 20), (16, 28)) if path == 'two_step' else ((8, 4), (8, 6), (16, 6))
    record = tuning.tune_launch(tuned, candidates=candidates, sweeps=3, units=1)
    assert record['path'] == path and not record['rejected'] and len(record['median_ms']) == 3
    names = ('waves2', 'chunk_x2') if path == 'two_step' else ('tile_y', 'chunk_x')
    assert tuple(record['chosen'][k] for k in names) in candidates
    assert all(tuned.launch_shape()[k] == record['chosen'][k] for k in names)
    assert torch.equal(tuned.grid.E, before) and int(tuned.counter.item()) == 8    # the run's state is back
    advance(tuned, 3)
    advance(plain, 3)
    for a, b in zip((tuned.grid.E, tuned.grid.H, *tuned.grid.memory_states), (plain.grid.E, plain.grid.H,
                                                                           *plain.grid.memory_states)):
        assert torch.equal(a, b)


def test_choice_is_cached_and_simulation_accepts_launch_auto(gpu):
    from torchfdtd_amd import Simulation, tuning
    from torchfdtd_amd.workloads import cube
    run = cube_stepper(n=64)
    assert run.path == 'split' and run.paths() == ['split', 'one_pass', 'two_step']    # small grids: split by default
    first = tuning.tune_launch(run, sweeps=2, units=1)
    assert not first['cached'] and len(first['median_ms']) == len(tuning.candidates_for(run))
    assert len(first['median_ms']) == 1+len(tuning.ONE_PASS)+len(tuning.TWO_STEP) and not first['rejected']
    assert run.path == first['path']
    again = tuning.tune_launch(cube_stepper(n=64))
    assert again['cached'] and again['chosen'] == first['chosen'] and again['path'] == first['path']
    project = cube(n=48, steps=24)
    auto = Simulation(project).run(cuda_graph_steps=8, launch='auto')
    default = Simulation(project).run(cuda_graph_steps=8)
    np.testing.assert_array_equal(auto.electric, default.electric)
    np.testing.assert_array_equal(auto.magnetic, default.magnetic)
    per_pass = {'split': 0, 'one_pass': 1, 'two_step': 2}
    assert auto.summary['steps_per_pass'] == per_pass[auto.summary['launch_tuning']['path']]
    assert default.summary['launch_tuning'] is None and default.summary['steps_per_pass'] == 0


def test_every_path_continues_a_run_bitwise(gpu):
    """set_path between whole units: a run that changes path every few steps equals one that never does."""
    mixed, plain = cube_stepper(n=40, two_step=False), cube_stepper(n=40, two_step=False)
    assert mixed.path == plain.path == 'split'
    for path in ('one_pass', 'two_step', 'split', 'two_step', 'one_pass'):
        mixed.set_path(path)
        advance(mixed, 1)
        for _ in range(4):
            plain.step()
    torch.cuda.synchronize()
    assert int(mixed.counter.item()) == int(plain.counter.item()) == 20
    for a, b in zip((mixed.grid.E, mixed.grid.H, *mixed.grid.memory_states),
                    (plain.grid.E, plain.grid.H, *plain.grid.memory_states)):
        assert torch.equal(a, b)
