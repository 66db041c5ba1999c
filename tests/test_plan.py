"""The resolved plan and its hash (torchfdtd_amd.plan; the reference's resolve_plan / plan_hash,
torchfdtd/plan.py:497-539), CPU.  The invariants are the reference's (tests/test_plan.py:66-146):
placement, ids and names stay outside the hash; a physical change shows in exactly the dotted keys
it touches; pml_dispersion enters the exterior section; run control enters the time section only
through the automatic shutoff.  Arrays are checked against the planners that the engine runs.
Parity unpinned: the hash values themselves (the reference's serialization is not reproduced)."""
import json

import numpy as np
import pytest

from torchfdtd_amd import BoundaryFace, FieldMonitor, Material, Monitor, Project, Region, Source, Structure
from torchfdtd_amd.models import RunControl
from torchfdtd_amd.plan import PLACEMENT_FIELDS, SAMPLE_TIME_STEPS, SimulationPlan, resolve_plan


def scene(monitors='point', **region):
    settings = dict(dimension='2d', size=(1.6, 1.5, 1.4), mesh=.1, pml_cells=3, steps=24, backend='cpu',
                    material_sampling='yee')
    settings.update(region)
    r = Region(**settings)
    r.boundaries.x_min = BoundaryFace(layers=4, kappa=2, alpha=.03, alpha_polynomial=1)
    observers = ([Monitor(id='ez', component='Ez', center=(.1, 0, 0)), Monitor(id='hy', component='Hy', center=(0, .1, 0))]
                 if monitors == 'point' else
                 [FieldMonitor(id='plane', center=(.12, 0, 0), size=(0, .65, .55), normal='x', downsample=2)])
    return Project(region=r, structures=[Structure(id='block', center=(-.3, .1, 0), size=(.3, .4, .2))],
                   sources=[Source(id='src', center=(-.2, 0, 0), pulse='continuous', wavelength=1.1)], monitors=observers)


def rebuilt(p):
    return Project.model_validate(p.model_dump())


def test_plan_is_frozen_and_placement_ids_and_names_stay_outside_the_hash():
    p = scene()
    plan = resolve_plan(p)
    assert isinstance(plan, SimulationPlan) and len(plan.plan_hash) == 64
    with pytest.raises((AttributeError, TypeError)):
        plan.time_step = 1.
    for array in (*plan.nodes, *plan.axes['Ex'], plan.source_terms[0][2], plan.boundaries.cpml[0].b,
                  plan.material.epsilon):
        assert not array.flags.writeable
    assert plan.sample_time_steps == SAMPLE_TIME_STEPS == {'E': 1., 'H': 1.5}
    assert plan.plan_hash == resolve_plan(p).plan_hash
    placed = p.model_copy(deep=True)
    r = placed.region
    r.backend, r.cuda_kernel, r.cuda_monitor_kernel = 'cuda', 'hip', 'hip'
    r.memory_mode, r.precision, r.execution_mode = 'streamed', 'float64', 'streamed_host'
    r.field, r.snapshot_interval, placed.name = 'Hy', 7, 'renamed'
    for item in placed.structures+placed.sources+placed.monitors:
        item.id, item.name = 'new-'+item.id, 'renamed'
    other = resolve_plan(rebuilt(placed))
    assert other.plan_hash == plan.plan_hash
    assert set(other.diff(plan)) == {'placement.'+k for k in ('backend', 'cuda_kernel', 'cuda_monitor_kernel', 'memory_mode',
                                                               'precision', 'execution_mode')}
    assert set(plan.placement) == set(PLACEMENT_FIELDS)
    np.testing.assert_array_equal(plan.material.ownership, resolve_plan(p).material.ownership)


def test_plan_arrays_are_the_engine_planners():
    """The frozen arrays equal what the engine's planners give (the plan wraps them, it restates nothing)."""
    from torchfdtd_amd.planning import cpml_slabs, source_terms, voxelize
    p = scene('field')
    plan = resolve_plan(p)
    r = p.region
    np.testing.assert_array_equal(plan.nodes[0], r.mesh_nodes[0])
    assert plan.time_step == r.time_step and plan.shape == tuple(r.shape)
    slabs = [s for v in cpml_slabs(r, dtype=np.float64).values() for s in v]
    assert len(slabs) == len(plan.boundaries.cpml) > 0
    for s, seg in zip(slabs, plan.boundaries.cpml):
        assert (s.lo, s.hi, s.axis, s.side) == (seg.start, seg.stop, seg.axis, seg.side)
        np.testing.assert_array_equal(s.b, seg.b)
        assert seg.b.dtype == np.float64
        assert (seg.kappa.max() > 1) == (seg.axis == 0 and seg.side == 0)     # only the x_min face has kappa 2
    (term,), = [s.terms for s in plan.sources]
    expected = source_terms(p, dtype=np.float64)[0].values
    np.testing.assert_array_equal(term.samples, expected)
    np.testing.assert_allclose(term.sample_times, np.arange(1, 25)*r.time_step, rtol=0, atol=0)
    np.testing.assert_array_equal(plan.material.epsilon, voxelize(p))
    (plane,) = plan.monitors
    assert plane.kind == 'field' and plane.maps and plane.frequency_hz is not None


def test_plan_json_and_diff_name_the_changed_keys():
    p = scene()
    plan = resolve_plan(p)
    payload = json.loads(json.dumps(plan.to_json(), allow_nan=False))
    assert payload['plan_hash'] == plan.plan_hash and payload['steps'] == 24
    assert len(payload['nodes'][0]) == plan.shape[0]+1 and len(payload['sources'][0]['terms'][0]['samples']) == 24
    assert payload['boundaries']['cpml'][0]['b'] and payload['placement']['backend'] == 'cpu'
    assert payload['resources']['cells'] == int(np.prod(plan.shape))
    moved = p.model_copy(deep=True)
    moved.structures[0].center = (-.2, .1, 0)
    assert resolve_plan(rebuilt(moved)).diff(plan) == ['material.structures[0].center[0]']
    refined = p.model_copy(deep=True)
    refined.region.mesh = .08
    keys = resolve_plan(rebuilt(refined)).diff(plan)
    for prefix in ('mesh.nodes', 'time.time_step', 'exterior.boundaries.cpml', 'sources[0].terms[0].samples'):
        assert any(k.startswith(prefix) for k in keys), prefix
    pulsed = p.model_copy(deep=True)
    pulsed.sources[0].phase = 30
    assert resolve_plan(rebuilt(pulsed)).diff(plan) == ['sources[0].terms[0].samples']
    # A setting the continuous waveform does not read is carried but not hashed.
    inert = p.model_copy(deep=True)
    inert.sources[0].wavelength_stop = 1.9
    again = resolve_plan(rebuilt(inert))
    assert again.diff(plan) == [] and again.sources[0].settings['wavelength_stop'] == 1.9
    with pytest.raises(TypeError):
        plan.diff(object())


def test_pml_dispersion_enters_the_exterior_section():
    plan = resolve_plan(scene())
    frozen = resolve_plan(scene(pml_dispersion='frozen'))
    assert plan.pml_dispersion == 'ade' and frozen.pml_dispersion == 'frozen'
    assert frozen.plan_hash != plan.plan_hash and frozen.diff(plan) == ['exterior.pml_dispersion']
    assert frozen.sections['exterior']['pml_dispersion'] == 'frozen' and frozen.to_json()['pml_dispersion'] == 'frozen'


def test_run_control_enters_the_time_section_only_through_the_shutoff():
    plan = resolve_plan(scene())
    assert plan.run_control == {'auto_shutoff': False} == plan.sections['time']['run_control']
    inert = resolve_plan(scene(run_control=RunControl(min_steps=20, check_interval=5, field_limit=1e3, growth_limit=10)))
    assert inert.plan_hash == plan.plan_hash and inert.diff(plan) == []
    on = resolve_plan(scene(run_control=RunControl(auto_shutoff=True)))
    keys = on.diff(plan)
    assert on.plan_hash != plan.plan_hash and 'time.run_control.auto_shutoff' in keys
    assert all(k.startswith('time.run_control.') for k in keys)
    assert on.to_json()['run_control']['min_steps'] == 100
    tuned = resolve_plan(scene(run_control=RunControl(auto_shutoff=True, min_steps=20)))
    assert tuned.diff(on) == ['time.run_control.min_steps']
    divergence = resolve_plan(scene(run_control=RunControl(auto_shutoff=True, field_limit=1e3, divergence_check=False)))
    assert divergence.plan_hash == on.plan_hash and divergence.diff(on) == []


def test_dispersive_materials_enter_through_their_ade_coefficients():
    p = scene()
    p.materials = [Material(name='metal', model='drude', epsilon_inf=2., plasma_rad_s=1e16, collision_rad_s=1e14)]
    p.structures[0].material = 'metal'
    p = rebuilt(p)
    plan = resolve_plan(p)
    (ade,) = plan.ade
    dt = p.region.time_step
    d = 1+.5*1e14*dt
    np.testing.assert_array_equal(ade.d, [d])
    np.testing.assert_array_equal(ade.k, [1e32*dt*dt/(4*d)])
    assert ade.a[0] == 0 and plan.materials[0]['epsilon'] == 2.
    shifted = p.model_copy(deep=True)
    shifted.materials[0].collision_rad_s = 2e14
    keys = resolve_plan(rebuilt(shifted)).diff(plan)
    assert 'material.ade[0].d' in keys and 'material.materials[0].oscillators[0][2]' in keys
    unused = p.model_copy(deep=True)
    unused.materials.append(Material(name='spare', model='lorentz'))
    assert resolve_plan(rebuilt(unused)).diff(plan) == []      # materials no structure uses are not hashed


@pytest.mark.gpu
def test_entry_points_share_the_plan_hash(gpu):
    """Simulation.run, the differentiable models and the batch report the plan_hash of resolve_plan."""
    import torch
    from torchfdtd_amd import Simulation
    from torchfdtd_amd.adjoint import DifferentiablePlaneSimulation, DifferentiableSimulation
    from torchfdtd_amd.batch import BatchCase, run_tensor_batch
    p = scene(dimension='3d', backend='cuda')
    expected = resolve_plan(p).plan_hash
    assert Simulation(p).run(cuda_graph_steps=4).summary['plan_hash'] == expected
    assert DifferentiableSimulation(p).plan_hash == expected
    (item,) = run_tensor_batch([BatchCase('a', p)]).items
    assert item.summary['plan_hash'] == expected
    planes = scene('field', dimension='3d', backend='cuda')
    planes.monitors[0].spectrum.apodization = 'none'
    planes = rebuilt(planes)
    assert DifferentiablePlaneSimulation(planes).plan_hash == resolve_plan(planes).plan_hash != expected
    torch.cuda.synchronize()
