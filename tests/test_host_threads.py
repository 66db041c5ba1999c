"""Two host threads step two grids at once, each on a stream of its own, through the one-pass and
two-step launches (every call forks the device's side streams from its stream and joins them back):
fields and CPML memory equal those of the same runs made one after the other, bit for bit."""
import threading

import pytest
import torch

pytestmark = pytest.mark.gpu


def project(source_x):
    from torchfdtd_amd import Project, Region, Source, Structure
    region = Region(dimension='3d', size=(3.2, 1.6, 1.2), mesh=.04, pml_cells=8, steps=40, backend='cuda')
    return Project(region=region,
                   structures=[Structure(center=(0, 0, 0), size=(2.4, .4, .3), material='Si (constant n)')],
                   sources=[Source(center=(source_x, 0, 0), component='Ez', wavelength=1, pulse_cycles=1)],
                   monitors=[])


def snapshot(run):
    g = run.grid
    return [g.E.cpu(), g.H.cpu()] + [m.cpu() for m in g.psi_buffers.values()]


@pytest.mark.parametrize('path', ['one_pass', 'two_step'])
def test_two_host_threads_step_at_once(gpu, path):
    from torchfdtd_amd.solver import Stepper
    projects = [project(.8), project(-.6)]

    def make():
        runs = [Stepper(p, one_pass=True, two_step=path == 'two_step') for p in projects]
        assert all(r.path == path for r in runs), [r.path for r in runs]
        return runs

    serial = make()
    for run in serial:
        run.advance(40)
    torch.cuda.synchronize()
    expected = [snapshot(run) for run in serial]

    runs = make()
    streams = [torch.cuda.Stream() for _ in runs]
    torch.cuda.synchronize()
    errors = []
    start = threading.Barrier(len(runs))

    def work(run, stream):
        try:
            with torch.cuda.stream(stream):
                start.wait(timeout=60)
                for _ in range(10):
                    run.advance(4)
            stream.synchronize()
        except BaseException as e:   # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(run, s)) for run, s in zip(runs, streams)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads) and not errors, errors
    torch.cuda.synchronize()
    for run, want in zip(runs, expected):
        got = snapshot(run)
        assert got[0].abs().max() > 0
        for a, b in zip(got, want):
            assert torch.equal(a, b)
