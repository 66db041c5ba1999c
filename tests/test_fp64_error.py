"""FP32 error of the engine against an FP64 run of the reference CPU path (the oracle in float64).

The engine's FP32 fields equal the reference's FP32 CPU path bit for bit (the parity tests), so
their distance from the FP64 oracle is the FP32 rounding error of the reference algorithm itself.
Relative L2 errors (fields over the whole grid, the point trace, the point spectrum, the float64
DFT plane) are printed by each test and held to the north-star bounds of the FP32 path, 1e-5 for
fields and traces and 1e-4 for spectra.  Measured on MI355X:

  160x120 2D TM waveguide, 1000 steps:   Ez 2.1e-6, H 2.3e-6, trace 1.3e-6, spectrum 1.4e-6
  160x120x40 3D waveguide, 1000 steps:   E 4.1e-6, H 4.2e-6, trace 3.9e-6, spectrum 4.0e-6
  128^3 cube (16-cell CPML, eps 1 or 12), 200 steps:  E 1.2e-7, H 3.3e-7
  uneven-batch grid 96x80x40, 600 steps, 16-frequency float64 plane: E 5.1e-6, H 3.6e-6, plane 2.2e-7

(The FP64 runs of the engine equal the FP64 oracle bit for bit, so the same figures come from an
FP64 run of the engine, which is how the benchmark reports them.)
"""
import numpy as np
import pytest

from oracle import simulate
from oracle.monitors import point_spectrum, spectrum_settings

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a-b).ravel())/np.linalg.norm(b.ravel()))


def fp64(project):
    raw = project.model_dump()
    raw['region']['precision'] = 'float64'
    return raw


def test_2d_waveguide_fp32_error_against_fp64(gpu):
    from torchfdtd_amd import Simulation
    from torchfdtd_amd.workloads import waveguide_2d
    project = waveguide_2d()
    got = Simulation(project).run(cuda_graph_steps=50)
    ref = simulate.run(fp64(project), record_frames=False)
    f, v = ref['spectra'][0]
    errors = dict(Ez=rel(got.electric[..., 2], ref['E'][..., 2]), H=rel(got.magnetic, ref['H']),
                  trace=rel(got.signals, ref['signals']), spectrum=rel(got.spectra[0]['value'], v))
    print('2D waveguide FP32 vs FP64 relative L2:', errors)
    np.testing.assert_array_equal(got.spectra[0]['frequency_hz'], f)
    assert errors['Ez'] < 1e-5 and errors['H'] < 1e-5 and errors['trace'] < 1e-5 and errors['spectrum'] < 1e-4


def test_3d_waveguide_fp32_error_against_fp64(gpu):
    from torchfdtd_amd import Simulation
    from torchfdtd_amd.workloads import waveguide_3d
    project = waveguide_3d()
    got = Simulation(project).run(cuda_graph_steps=50)
    ref = simulate.run(fp64(project), record_frames=False, native=True)
    settings = spectrum_settings(project.monitors[0].spectrum.model_dump())
    settings.update(sampling='wavelength', frequency_points=32, wavelength_start=1.5, wavelength_stop=1.6, apodization='none')
    times = ref['times']
    _, expected = point_spectrum(times, ref['signals'][:, 0], settings)
    _, spectrum = point_spectrum(times, got.signals[:, 0].astype(np.float64), settings)
    errors = dict(E=rel(got.electric, ref['E']), H=rel(got.magnetic, ref['H']), trace=rel(got.signals, ref['signals']),
                  spectrum=rel(spectrum, expected))
    print('3D waveguide FP32 vs FP64 relative L2:', errors)
    assert errors['E'] < 1e-5 and errors['H'] < 1e-5 and errors['trace'] < 1e-5 and errors['spectrum'] < 1e-4


def test_cube_fp32_error_against_fp64(gpu):
    from oracle.native import NativeGrid
    from oracle.scene import Scene
    from torchfdtd_amd import Simulation
    from torchfdtd_amd.workloads import cube, cube_noise_mask
    n, steps = 128, 200
    project = cube(n=n, steps=steps)
    mask = cube_noise_mask(n)
    eps = (1+11*mask.double()).cpu().numpy()
    del mask
    got = Simulation(project, epsilon=eps.astype(np.float32)).run(cuda_graph_steps=20)
    raw = fp64(project)
    scene = Scene(raw)
    inv = (1/eps)[..., None].repeat(3, axis=-1)
    ref = simulate.run(raw, grid=NativeGrid(scene, np.ascontiguousarray(inv)), record_frames=False)
    errors = dict(E=rel(got.electric, ref['E']), H=rel(got.magnetic, ref['H']))
    print('128^3 cube FP32 vs FP64 relative L2:', errors)
    assert errors['E'] < 1e-5 and errors['H'] < 1e-5


def test_uneven_batch_grid_plane_fp32_error_against_fp64(gpu):
    from oracle.native import NativeGrid
    from oracle.scene import Scene
    from torchfdtd_amd import workloads as W
    from torchfdtd_amd.batch import run_tensor_batch
    cases = W.uneven_batch(1, 600, shapes=[(96, 80, 40)])
    report = run_tensor_batch(cases, cuda_graph_steps=50)
    item, case = report.items[0], cases[0]
    raw = fp64(case.project)
    scene = Scene(raw)
    # The FP32 case holds float32(12.11) and float32(2.085); the FP64 run takes the exact values.
    eps = np.where(case.epsilon == np.float32(W.EPS_HIGH), W.EPS_HIGH, W.EPS_LOW)
    assert np.array_equal(eps.astype(np.float32), case.epsilon)
    inv = (1/eps)[..., None]
    ref = simulate.run(raw, grid=NativeGrid(scene, np.ascontiguousarray(np.broadcast_to(inv, (*scene.shape, 3)))),
                       record_frames=False)
    res = item.result
    (plane,), (expected,) = res.frequency_fields, ref['planes']
    names = list(expected.components)
    values = expected.value[..., [names.index(c) for c in plane['components']]]
    errors = dict(E=rel(res.electric, ref['E']), H=rel(res.magnetic, ref['H']), plane=rel(plane['fields'], values))
    print('96x80x40 batch grid FP32 vs FP64 relative L2:', errors)
    assert errors['E'] < 1e-5 and errors['H'] < 1e-5 and errors['plane'] < 1e-4
