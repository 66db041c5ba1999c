"""Uneven batches against independent runs, bit for bit.

The reference's own check (tests/test_tensor_batch.py:25-41) runs a cohort and every
case alone and requires equal arrays.  Here the cohorts mix shapes, CPML widths,
source and monitor positions, plane precisions and permittivity layouts, and every
launch configuration must give exactly the single-run result.
"""
import numpy as np
import pytest
import torch

from torchfdtd_amd import workloads as W
from torchfdtd_amd.batch import (BatchCase, Launch, TensorBatchTuning, plan_grouped_batch, resolve_launch,
                                 run_grouped_batch, run_tensor_batch, tune_tensor_batch)
from torchfdtd_amd.models import Monitor, Project, Region, Source

SHAPES = [(24, 20, 16), (30, 18, 22), (20, 27, 18), (33, 24, 17)]


def uneven_cases(precision='float32', steps=64):
    rng = np.random.default_rng(7)
    cases = []
    for i, shape in enumerate(SHAPES):
        p = W.uneven_grid(shape, steps=steps, pml=4+i % 2, frequencies=3+i, precision=precision, inset_cells=6, name=f't{i}')
        p.monitors.append(Monitor(id='probe', center=(.04*(i-1), 0, 0), component=('Ez', 'Hy')[i % 2]))
        if i == 2:
            p.monitors[0].dft_precision = 'field'
        design = W.binary_design(rng, shape=(8, 8))
        cases.append(BatchCase(f'c{i}', p, epsilon=W.design_epsilon(p, design, depth=4)))
    return cases


def plain_cases(dimension='2d'):
    cases = []
    for i, (sx, sy) in enumerate([(1.6, 1.2), (1.2, 1.4), (2.0, 1.0)]):
        p = Project(name=f'p{i}', region=Region(dimension=dimension, size=(sx, sy, 1.2 if dimension == '3d' else .05), mesh=.05,
                                                steps=50, pml_cells=5+i, backend='cuda', snapshot_interval=10),
                    sources=[Source(id='s', center=(-.2, .05*i, 0), component='Ez')],
                    monitors=[Monitor(id='m', center=(.2, 0, 0), component='Ez')])
        cases.append(BatchCase(f'p{i}', p))
    return cases


def solo(cases):
    from torchfdtd_amd.solver import Simulation
    return [Simulation(c.project, epsilon=c.epsilon).run() for c in cases]


def assert_same(batch, singles):
    for item, ref in zip(batch.items, singles):
        res = item.result
        assert item.status == 'completed'
        assert np.array_equal(res.electric, ref.electric), item.id
        assert np.array_equal(res.magnetic, ref.magnetic), item.id
        assert np.array_equal(res.signals, ref.signals), item.id
        assert np.array_equal(res.frames, ref.frames), item.id
        assert len(res.frequency_fields) == len(ref.frequency_fields)
        for a, b in zip(res.frequency_fields, ref.frequency_fields):
            for key, value in b.items():
                if isinstance(value, np.ndarray):
                    assert np.array_equal(a[key], value), (item.id, key)


@pytest.fixture(scope='module')
def uneven():
    cases = uneven_cases()
    return cases, solo(cases)


@pytest.mark.gpu
@pytest.mark.parametrize('launch', [Launch('split'), Launch('march', 8, 8), Launch('march', 16, 0), Launch('march', 8, 3)],
                         ids=lambda l: f'{l.kernel}-{l.tile_y}-{l.chunk_x}')
def test_uneven_cohort_matches_independent_runs(gpu, uneven, launch):
    cases, singles = uneven
    report = run_tensor_batch(cases, launch=launch, cuda_graph_steps=8)
    assert report.plan['launch'] == launch.as_dict()
    assert_same(report, singles)


@pytest.mark.gpu
def test_eager_cohort_and_cohort_splitting(gpu, uneven):
    cases, singles = uneven
    report = run_tensor_batch(cases, cuda_graph=False, cohort_size=3)
    assert report.plan['cohorts_completed'] == 2
    assert_same(report, singles)


@pytest.mark.gpu
def test_mixed_permittivity_layouts_share_one_pass_launch(gpu):
    cases = uneven_cases(steps=40)
    # Per-component sampling for one case: the cohort widens the others' per-cell arrays.
    p = cases[1].project
    p.region.material_sampling = 'yee'
    cases[1] = BatchCase('yee', Project.model_validate(p.model_dump()))
    singles = solo(cases)
    report = run_tensor_batch(cases, launch=Launch('march', 8, 8))
    assert_same(report, singles)


@pytest.mark.gpu
def test_two_dimensional_cohort_runs_split(gpu):
    cases = plain_cases('2d')
    assert resolve_launch([c.project for c in cases], 'auto') == Launch('split')
    report = run_tensor_batch(cases, cuda_graph_steps=4)
    assert_same(report, solo(cases))
    with pytest.raises(ValueError, match='one-pass'):
        run_tensor_batch(cases, launch=Launch('march'))


@pytest.mark.gpu
def test_float64_cohort(gpu):
    cases = uneven_cases('float64', steps=30)
    report = run_tensor_batch(cases, launch=Launch('march', 16, 4))
    assert_same(report, solo(cases))


@pytest.mark.gpu
def test_autotuner_checks_bitwise_identity_and_rejects_inapplicable_candidates(gpu):
    cases = uneven_cases(steps=30)
    candidates = (Launch('split'), Launch('march', 8, 8), Launch('march', 16, 0))
    report = tune_tensor_batch(cases, candidates=candidates, repeats=1, trial_steps=12)
    assert isinstance(report, TensorBatchTuning)
    assert report.launch in candidates and not report.rejected
    assert len(report.timings) == 3
    assert tune_tensor_batch(cases, candidates=candidates, repeats=1, trial_steps=12) is report   # cached
    flat = tune_tensor_batch(plain_cases('2d'), candidates=candidates, repeats=1, trial_steps=8)
    assert flat.launch == Launch('split') and len(flat.rejected) == 2
    out = run_tensor_batch(cases, launch=report)
    assert out.plan['launch'] == report.launch.as_dict()


@pytest.mark.gpu
def test_grouped_batch_restores_input_order(gpu):
    cases = uneven_cases(steps=20) + plain_cases('3d')
    report = run_grouped_batch(cases, cohort_size=3)
    assert [item.id for item in report.items] == [c.id for c in cases]
    assert_same(report, solo(cases))


# ------------------------------------------------------------------ host logic (no GPU)

def test_plan_groups_by_steps_and_precision_and_caps_cohorts():
    cases = uneven_cases(steps=20) + uneven_cases('float64', steps=20)[:2] + plain_cases('3d')
    cases = [BatchCase(f'k{i}', c.project) for i, c in enumerate(cases)]
    plan = plan_grouped_batch(cases, cohort_size=2)
    members = sorted(i for cohort in plan['cohorts'] for i in cohort)
    assert members == list(range(len(cases)))
    for cohort in plan['cohorts']:
        assert 1 <= len(cohort) <= 2
        keys = {(cases[i].project.region.steps, cases[i].project.region.precision) for i in cohort}
        assert len(keys) == 1
    limited = plan_grouped_batch(cases, cohort_size=64, memory_limit_bytes=1)
    assert limited['cohort_count'] == len(cases)


def test_case_and_launch_validation():
    with pytest.raises(ValueError):
        BatchCase('a/b', plain_cases()[0].project)
    with pytest.raises(ValueError):
        Launch('tiles')
    with pytest.raises(ValueError):
        Launch('march', 12)
    with pytest.raises(ValueError):
        plan_grouped_batch(plain_cases(), cohort_size=65)
    assert resolve_launch([c.project for c in uneven_cases()], 'auto') == Launch('march')


def test_uneven_workload_is_seeded_and_within_bounds():
    shapes = W.uneven_shapes()
    assert shapes == W.uneven_shapes() and len(shapes) == 64
    assert all(64 <= x <= 192 and 64 <= y <= 160 and 32 <= z <= 64 for x, y, z in shapes)
    design = W.binary_design(np.random.default_rng(0))
    assert design.shape == (32, 32) and .2 < design.mean() < .8
    # The conic filter is normalised: a constant density passes unchanged.
    assert np.allclose(W.conic_filter(np.full((9, 9), .3), 2.5), .3)


@pytest.mark.gpu
def test_cohort_reset_reproduces_fresh_cohorts(gpu):
    from torchfdtd_amd.batch import Cohort
    cases = uneven_cases(steps=30)
    projects = [c.project for c in cases]
    cohort = Cohort(projects, Launch('march', 8, 8), epsilons=[c.epsilon for c in cases])
    graphs = cohort.graphs(8)

    def run():
        for _ in range(3):
            graphs.replay(8)
        for _ in range(6):
            graphs.replay(1)
        return cohort.digest()
    first = run()
    cohort.reset([c.epsilon for c in cases])
    assert run() == first
    other = [np.flip(c.epsilon, 0).copy() for c in cases]
    cohort.reset(other)
    changed = run()
    fresh = Cohort(projects, Launch('march', 8, 8), epsilons=other)
    fresh.advance(30)
    assert changed != first and fresh.digest() == changed
