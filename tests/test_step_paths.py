"""The default step path by grid size (solver.default_step_path), CPU: the thresholds sit at the
measured crossovers of the split step, the one-pass step and the two-step pass."""
from torchfdtd_amd.solver import ONE_PASS_MIN_CELLS, TWO_STEP_MIN_CELLS, default_step_path, merge_ranges


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


def test_snapshot_ranges_merge_to_at_most_four():
    """The x ranges of a step n+1 snapshot: overlapping and touching spans joined, then the nearest
    neighbours until four remain; every span lies inside one range."""
    assert merge_ranges([]) == []
    assert merge_ranges([(64, 67), (65, 66), (66, 67)]) == [(64, 67)]
    assert merge_ranges([(10, 12), (12, 13), (64, 67)]) == [(10, 13), (64, 67)]
    spans = [(0, 1), (10, 11), (20, 22), (30, 31), (33, 34), (60, 62)]
    merged = merge_ranges(spans)      # gaps 9, 9, 8, 2, 26: the 2-plane gap closes, then the 8-plane one
    assert merged == [(0, 1), (10, 11), (20, 34), (60, 62)]
    assert merge_ranges(spans, most=5) == [(0, 1), (10, 11), (20, 22), (30, 34), (60, 62)]
    for lo, hi in spans:
        assert any(a <= lo and hi <= b for a, b in merged)
    assert all(a[1] < b[0] for a, b in zip(merged, merged[1:]))
