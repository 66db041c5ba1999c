"""The default step path by grid size (solver.default_step_path), CPU: the thresholds sit at the
measured crossovers of the split step, the one-pass step and the two-step pass."""
from torchfdtd_amd.solver import ONE_PASS_MIN_CELLS, TWO_STEP_MIN_CELLS, default_step_path


def test_default_step_path_follows_the_measured_crossovers():
    assert TWO_STEP_MIN_CELLS < ONE_PASS_MIN_CELLS
    assert default_step_path(96**3, True, True) == 'split'
    assert default_step_path(192 * 160 * 64, False, True) == 'split'
    assert default_step_path(256**3, True, True) == 'split'        # 16.8 M cells: split and two-step tie
    assert default_step_path(320**3, True, True) == 'two_step'     # 32.8 M: the two-step pass 21% faster
    assert default_step_path(320**3, False, True) == 'split'       # observed: one-pass and split tie
    assert default_step_path(512**3, False, True) == 'one_pass'
    assert default_step_path(512**3, True, True) == 'two_step'
    assert default_step_path(512**3, True, False) == 'split'       # no one-pass kernel for this grid
