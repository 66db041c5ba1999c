"""The default step path by grid size (solver.default_step_path), CPU: the thresholds sit at the
measured crossovers of the split step, the one-pass step and the two-step pass."""
from torchfdtd_amd.solver import ONE_PASS_MIN_CELLS, TWO_STEP_MIN_CELLS, default_step_path, merge_boxes


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


def test_snapshot_boxes_merge_to_at_most_four():
    """The boxes of a step n+1 snapshot: overlapping and touching boxes replaced by their bounding box,
    then the pair that adds the fewest cells, until four remain; every observer box lies inside one."""
    assert merge_boxes([]) == []
    plane = ((64, 10, 9), (67, 31, 22))
    traces = [((65, 22, 15), (66, 23, 16)), ((66, 20, 16), (67, 21, 17))]
    assert merge_boxes([plane, *traces]) == [plane]                     # traces inside the plane's box
    far = ((17, 10, 9), (19, 31, 22))
    assert merge_boxes([plane, far]) == [far, plane]                    # two disjoint boxes stay two
    assert merge_boxes([((0, 0, 0), (1, 1, 1)), ((1, 0, 0), (2, 1, 1))]) == [((0, 0, 0), (2, 1, 1))]   # touching
    points = [((x, y, z), (x + 1, y + 1, z + 1)) for x, y, z in ((0, 0, 0), (0, 0, 9), (50, 0, 0), (50, 30, 0), (50, 30, 9))]
    merged = merge_boxes(points)
    assert len(merged) == 4 and ((0, 0, 0), (1, 1, 10)) in merged      # the pair 9 cells apart joins first
    for lo, hi in points:
        assert any(all(a[k] <= lo[k] and hi[k] <= b[k] for k in range(3)) for a, b in merged)
