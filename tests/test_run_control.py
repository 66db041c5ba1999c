"""Run control decided on the GPU (tfdtd_decay_update) against the oracle's rule
(torchfdtd/run_control.py:58-85).

CPU: the oracle's rule on energy sequences whose stop or failure step is worked out by hand.
GPU: the device rule on synthetic norms of several grids at once equals the oracle rule barrier by
barrier (verdicts, steps, every history value); whole runs stop, or fail, at the oracle's step.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import simulate
from oracle.run_control import Rule


# ---------------------------------------------------------------- the oracle's rule (CPU)

def test_rule_stops_after_the_required_run_of_decayed_checks():
    """Checks every 50 steps, sources over at t = 0, eligible from step 100.  Energy 1 at step 50, 4 at
    step 100 (the frozen baseline), then divided by 10 per check: ratios at steps 150, 200, 250 ...
    are .1, .01, 1e-3, 1e-4 ...  With threshold 2e-3 the ratio is under it from step 250; three
    such checks in a row end the run at step 350."""
    rule = Rule(eligible_time=0., dt=1e-17, decay_threshold=2e-3, min_steps=100, consecutive_checks=3)
    energies = [1., 4., .4, .04, .004, 4e-4, 4e-5, 4e-6]
    stops = [rule.update(50*(k+1), e, math.sqrt(e)) for k, e in enumerate(energies)]
    assert stops == [False]*6+[True, True]
    assert stops.index(True) == 6 and 50*(6+1) == 350
    assert [h['source_finished'] for h in rule.history][:2] == [False, True]
    assert rule.history[1]['relative_norm'] == 1 and rule.history[4]['relative_norm'] == pytest.approx(1e-3, rel=1e-12)


def test_rule_waits_for_the_sources_and_resets_on_a_rise():
    """eligible_time 2.4e-15 s at dt 1e-17 s: eligible from step 250 (the baseline freezes at the
    peak energy 10 there), so the low energies of steps 100-200 count for nothing.  Step 250 is
    decayed (ratio 1e-4), step 300 rises (ratio .5) and resets the run, steps 350 and 400 decay:
    two in a row stop the run at step 400."""
    rule = Rule(eligible_time=2.4e-15, dt=1e-17, decay_threshold=.1, min_steps=0, consecutive_checks=2)
    energies = [10., 1e-3, 1e-3, 1e-3, 1e-3, 5., 1e-3, 1e-3]   # steps 50 .. 400
    out = [rule.update(50*(k+1), e, 1.) for k, e in enumerate(energies)]
    assert out == [False]*7+[True]
    assert [h['source_finished'] for h in rule.history] == [False]*4+[True]*4
    rule = Rule(eligible_time=0., dt=1e-17, decay_threshold=.1, min_steps=0, consecutive_checks=2)
    out = [rule.update(50*(k+1), e, 1.) for k, e in enumerate([10., 1e-3, 5., 1e-3, 1e-3])]
    assert out == [False, False, False, False, True]


def test_rule_raises_on_growth_field_limit_and_non_finite_values():
    rule = Rule(eligible_time=0., dt=1e-17, growth_limit=100., min_steps=100)
    rule.update(50, 1., 1.)
    rule.update(100, 2., 1.)                 # baseline frozen at 2
    rule.update(150, 200., 1.)               # ratio 100: at the limit, not over
    with pytest.raises(FloatingPointError, match='grew by 100.5 at step 200'):
        rule.update(200, 201., 1.)
    rule = Rule(eligible_time=0., dt=1e-17, field_limit=3.)
    rule.update(50, 1., 3.)
    with pytest.raises(FloatingPointError, match='limit exceeded at step 100: 3.5'):
        rule.update(100, 1., 3.5)
    with pytest.raises(FloatingPointError, match='Non-finite full-domain or auxiliary state at step 150'):
        Rule(eligible_time=0., dt=1e-17).update(150, math.inf, 1.)
    # Without divergence checks growth and field limits pass.
    rule = Rule(eligible_time=0., dt=1e-17, growth_limit=1., field_limit=1., divergence_check=False, min_steps=0)
    assert rule.update(50, 1., 5.) is False and rule.update(100, 50., 5.) is False


def test_oracle_state_norm_is_the_weighted_sum():
    rng = np.random.default_rng(4)
    E, H = rng.standard_normal((2, 4, 3, 2, 3))
    inv = rng.random((4, 3, 2, 3))+.5
    from oracle.run_control import state_norm
    energy, peak, finite = state_norm(E, H, inv)
    assert energy == pytest.approx(float((E*E/inv).sum()+(H*H).sum()), rel=1e-14)
    assert peak == max(np.abs(E).max(), np.abs(H).max()) and finite


# ---------------------------------------------------------------- the device rule (GPU)

def _rule_args(r):
    return dict(eligible_time=r.eligible_time, dt=r.time_step, decay_threshold=r.decay_threshold,
                growth_limit=r.growth_limit, field_limit=r.field_limit or None, min_steps=r.min_steps,
                consecutive_checks=r.consecutive_checks, divergence_check=bool(r.divergence_check),
                auto_shutoff=bool(r.auto_shutoff))


@pytest.mark.gpu
def test_device_rule_equals_the_oracle_rule_on_synthetic_norms(gpu):
    import torch

    from torchfdtd_amd import _native as N
    lib = N.library()
    rng = np.random.default_rng(11)
    count, barriers, rows = 6, 40, 40
    rules = (N.DecayRule*count)()
    for k, r in enumerate(rules):
        r.eligible_time, r.time_step = [0., 3e-15, 0., 1e-15, 0., 0.][k], 1e-17
        r.decay_threshold = [1e-3, 1e-2, 1e-9, .5, 1e-3, 1e-3][k]
        r.growth_limit = [1e6, 1e6, 10., 1e6, 1e6, 1e6][k]
        r.field_limit = [0., 0., 0., 0., 2.5, 0.][k]
        r.min_steps, r.consecutive_checks = [100, 0, 100, 50, 0, 100][k], [3, 2, 3, 1, 3, 3][k]
        r.divergence_check, r.auto_shutoff = [1, 1, 1, 1, 1, 1][k], [1, 1, 1, 1, 1, 0][k]
    energy = np.empty((barriers, count))
    for k in range(count):
        rise = np.exp(np.linspace(0, 3, 8))
        fall = np.exp(-rng.uniform(.3, 1.2)*np.arange(barriers-8))
        energy[:, k] = np.r_[rise, rise[-1]*fall]
    energy[20:, 2] *= np.exp(np.arange(barriers-20))        # grid 2 grows past its limit
    peak = np.sqrt(energy)
    flag = np.zeros_like(energy)
    flag[30, 5] = 1                                          # grid 5: non-finite CPML memory at barrier 30
    device = torch.device('cuda')
    rule_table = torch.frombuffer(bytearray(bytes(rules)), dtype=torch.uint8).to(device)
    states = torch.zeros(ctypes.sizeof(N.DecayState)*count, dtype=torch.uint8, device=device)
    history = torch.zeros((count, rows, 6), dtype=torch.float64, device=device)
    norms = torch.zeros((count, 3), dtype=torch.float64, device=device)
    oracle = [Rule(**_rule_args(r)) for r in rules]
    expected = [None]*count          # (verdict, step)
    for b in range(barriers):
        step = 50*(b+1)
        norms.copy_(torch.as_tensor(np.stack([energy[b], peak[b], flag[b]], axis=1)))
        N.check(lib.tfdtd_decay_update(rule_table.data_ptr(), states.data_ptr(), norms.data_ptr(), count, step,
                                       history.data_ptr(), rows, N.stream_handle()), 'tfdtd_decay_update')
        for k in range(count):
            if expected[k] is not None:
                continue
            if flag[b, k]:
                expected[k] = (N.RUN_NONFINITE, step)
                continue
            try:
                if oracle[k].update(step, energy[b, k], peak[b, k]):
                    expected[k] = (N.RUN_DECAYED, step)
            except FloatingPointError as exc:
                kind = N.RUN_GROWTH if 'grew' in str(exc) else N.RUN_FIELD_LIMIT
                expected[k] = (kind, step)
    torch.cuda.synchronize()
    got = (N.DecayState*count).from_buffer_copy(states.cpu().numpy().tobytes())
    table = history.cpu().numpy()
    kinds = set()
    for k in range(count):
        verdict = (got[k].verdict, got[k].verdict_step) if got[k].verdict else None
        assert verdict == expected[k], k
        kinds.add(verdict[0] if verdict else 0)
        rows_k = oracle[k].history
        assert got[k].checks == len(rows_k)
        for row, h in zip(table[k], rows_k):
            assert tuple(row) == (h['step'], h['time_s'], h['state_norm'], h['relative_norm'], h['field_peak'],
                                  float(h['source_finished'])), k
    assert kinds >= {N.RUN_DECAYED, N.RUN_GROWTH, N.RUN_FIELD_LIMIT, N.RUN_NONFINITE}


def _shutoff_project(**control):
    from torchfdtd_amd.workloads import waveguide_2d
    project = waveguide_2d(steps=2000)
    rc = project.region.run_control
    rc.auto_shutoff, rc.check_interval, rc.decay_threshold, rc.consecutive_checks = True, 50, 1e-4, 2
    for key, value in control.items():
        setattr(rc, key, value)
    return project


def _oracle_rule(project):
    from torchfdtd_amd.run_control import decay_rule
    rule = decay_rule(project, project.region.time_step)
    return Rule(**_rule_args(rule))


@pytest.mark.gpu
def test_auto_shutoff_stops_at_the_oracle_step(gpu):
    from torchfdtd_amd import Simulation
    project = _shutoff_project()
    got = Simulation(project).run(cuda_graph_steps=25)
    rule = _oracle_rule(project)
    ref = simulate.run(project.model_dump(), rule=rule, check_interval=50, record_frames=False)
    stop = len(ref['signals'])
    assert got.summary['termination_reason'] == 'decayed' and got.summary['steps'] == stop < 2000
    np.testing.assert_array_equal(got.electric, ref['E'])
    np.testing.assert_array_equal(got.signals, ref['signals'])
    assert got.frame_steps[-1] == stop
    history = got.summary['diagnostics']
    assert [h['step'] for h in history] == [h['step'] for h in rule.history]
    np.testing.assert_allclose([h['state_norm'] for h in history], [h['state_norm'] for h in rule.history], rtol=1e-12)
    assert [h['source_finished'] for h in history] == [h['source_finished'] for h in rule.history]


@pytest.mark.gpu
def test_field_limit_fails_at_the_oracle_step(gpu):
    from torchfdtd_amd import Simulation
    project = _shutoff_project(field_limit=.05, auto_shutoff=False)
    rule = _oracle_rule(project)
    with pytest.raises(FloatingPointError) as expected:
        simulate.run(project.model_dump(), rule=rule, check_interval=50, record_frames=False)
    with pytest.raises(FloatingPointError) as failure:
        Simulation(project).run(cuda_graph_steps=25)
    assert str(failure.value).startswith(str(expected.value))
    assert 'Field magnitude limit exceeded at step' in str(failure.value)


@pytest.mark.gpu
def test_progress_reports_the_latest_diagnostics(gpu):
    from torchfdtd_amd import Simulation
    project = _shutoff_project(auto_shutoff=False)
    project.region.steps = 300
    reports = []
    got = Simulation(project).run(progress=reports.append, cuda_graph_steps=10)
    assert [r['step'] for r in reports] == list(got.frame_steps)
    with_diagnostics = [r for r in reports if r['diagnostics'] is not None]
    assert with_diagnostics and all(r['diagnostics']['step'] <= r['step'] for r in with_diagnostics)
    assert len(got.summary['diagnostics']) == 300//50
