"""Host planning of the product against the oracle (CPU only): integer maps bit-exact,
coefficients, permittivity and waveforms bitwise in the field precision."""
import json
import math

import numpy as np
import pytest

from oracle.monitors import interpolation_map as oracle_map
from oracle.monitors import plane_plan as oracle_plane
from oracle.scene import Scene
from tests.conftest import GOLDEN
from torchfdtd_amd import FieldMonitor, Material, Monitor, Project, Region, Source, SpectrumSettings, Structure
from torchfdtd_amd import planning
from torchfdtd_amd.monitors import interpolation_map, plane_plan


def configs():
    golden = Project.model_validate_json(str(np.load(GOLDEN/'tiny_2d_complete.npz')['project']))
    core = Material(name='core', index=math.sqrt(12.11))
    waveguide2d = Project(region=Region(dimension='2d', size=(8, 6, .05), mesh=.05, steps=1000, pml_cells=10,
                                        background_index=math.sqrt(2.085)),
                          materials=[core], structures=[Structure(size=(8, .65, 1), material='core')],
                          sources=[Source(center=(-2.5, 0, 0), pulse='broadband', time_definition='wavelength',
                                          wavelength_start=1.5, wavelength_stop=1.6)])
    waveguide3d = Project(region=Region(dimension='3d', size=(8, 6, 2), mesh=.05, steps=1000, pml_cells=12,
                                        background_index=math.sqrt(2.085)),
                          materials=[core], structures=[Structure(size=(8, .65, .4), material='core')],
                          sources=[Source(center=(-2.5, 0, 0), component='Ey', pulse='broadband', time_definition='wavelength',
                                          wavelength_start=1.5, wavelength_stop=1.6)])
    yee = Project(region=Region(dimension='3d', size=(3.2, 2.8, 2.4), mesh=.1, pml_cells=4, material_sampling='yee',
                                boundaries={'y_min': {'kind': 'periodic'}, 'y_max': {'kind': 'periodic'},
                                            'x_min': {'layers': 5, 'kappa': 3, 'alpha': .03, 'alpha_polynomial': 1}}),
                  structures=[Structure(kind='sphere', radius=.4, material='SiO2 (constant n)'),
                              Structure(kind='circle', radius=.3, size=(1, 1, .6), center=(.4, .2, 0), mesh_order=1)],
                  sources=[Source(center=(-.8, 0, 0)), Source(kind='plane', center=(-1, 0, 0), size=(0, 2, 1.2), component='Hz')])
    return {'golden': golden, 'waveguide2d': waveguide2d, 'waveguide3d': waveguide3d, 'yee': yee}


@pytest.mark.parametrize('name', ['golden', 'waveguide2d', 'waveguide3d', 'yee'])
def test_planning_matches_oracle(name):
    p = configs()[name]
    scene = Scene(p.model_dump())
    r = p.region
    assert r.shape == scene.shape
    assert r.time_step == scene.dt and r.rectangular_courant == scene.courant
    for a in range(3):
        np.testing.assert_array_equal(r.mesh_nodes[a], scene.nodes[a])
    ours, theirs = planning.cpml_slabs(r), scene.cpml_segments()
    assert ours.keys() == theirs.keys()
    for key in ours:
        assert len(ours[key]) == len(theirs[key])
        for s, t in zip(ours[key], theirs[key]):
            assert (s.lo, s.hi) == (t['lo'], t['hi'])
            for field in ('b', 'c', 'inv_k'):
                np.testing.assert_array_equal(getattr(s, field), t[field].ravel())
    np.testing.assert_array_equal(planning.voxelize(p), scene.voxelize())
    terms = planning.source_terms(p)
    expected = scene.source_terms()
    assert len(terms) == len(expected)
    for t, (comp, loc, values) in zip(terms, expected):
        assert t.family+'xyz'[t.comp] == comp[0]+comp[1].lower()
        assert tuple(t.box) == tuple(slice(v, v+1) if isinstance(v, int) else v for v in loc)
        np.testing.assert_array_equal(t.values, values)


def test_cpml_slab_rows_for_the_2d_configuration():
    """10-cell CPML on 160 cells: E (backward) rows 1..9 and 150..159, H (forward) rows 0..9 and 150..158."""
    r = configs()['waveguide2d'].region
    slabs = planning.cpml_slabs(r)
    assert [(s.lo, s.hi) for s in slabs[False, 0, 2]] == [(1, 10), (150, 160)]
    assert [(s.lo, s.hi) for s in slabs[True, 0, 2]] == [(0, 10), (150, 159)]
    assert (False, 2, 0) not in slabs   # the invariant z axis has no derivative


def test_psi_count_of_the_512_cube():
    r = Region(dimension='3d', size=(51.2, 51.2, 51.2), mesh=.1, pml_cells=16)
    assert r.shape == (512, 512, 512)
    per_cell = planning.psi_elements(r)/512**3
    assert abs(per_cell-0.727) < 2e-3


def test_plane_maps_match_oracle():
    p = configs()['yee']
    scene = Scene(p.model_dump())
    m = FieldMonitor(normal='x', center=(.7, 0, 0), size=(0, 1.03, 1), spectrum=SpectrumSettings(sampling='frequency', frequency_points=3, apodization='none'))
    plan = plane_plan(p.region, m)
    points, area = oracle_plane(scene, m.model_dump())
    np.testing.assert_array_equal(plan['points_um'], points)
    np.testing.assert_array_equal(plan['weights'], area)
    for comp in m.required_fields:
        i, w = interpolation_map(p.region, comp, points)
        oi, ow = oracle_map(scene, comp, points)
        np.testing.assert_array_equal(i, oi)
        np.testing.assert_array_equal(w, ow)


def test_reference_project_files_load():
    project = Project.model_validate_json((GOLDEN/'tiny_result.json').read_text())
    assert project.region.shape == (20, 20, 1) and project.region.cuda_kernel == 'torch'
    dumped = json.loads(project.model_dump_json())
    assert dumped['region']['pml_cells'] == 3


def test_unsupported_physics_is_refused_before_allocation():
    from torchfdtd_amd.grid import YeeGrid
    r = Region(mesh_type='graded', material_sampling='yee')
    with pytest.raises(ValueError, match='graded'):
        YeeGrid(r)
