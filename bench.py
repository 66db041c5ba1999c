"""Benchmark of the fused 3D Yee/CPML step on one MI355X: the 512^3 roofline cube.

Workload: 512^3 cells FP32, 16-cell CPML on all faces, per-cell eps = 1 + 11 s
(s = smoothed, thresholded seeded noise), Ez dipole at the centre.  One bench step is
one Yee time step (E and H of every cell, CPML, source, counter) replayed from a HIP
graph; the configuration runs 500 of them.  Inputs are resident in HBM before timing.

Prints one JSON line: Gcell-updates/s over the timed steps, the roofline of the
dominant kernel (algorithmic bytes per launch / its HIP-event-timed duration against
the 8 TB/s HBM peak) and the NumPy oracle of the reference CPU path timed on a bounded
slab of the same workload on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(region, real_bytes=4):
    """Per cell-update: E,H read+write (12 values), eps^-1 per cell, CPML psi read+write."""
    from torchfdtd_amd.planning import psi_elements
    cells = math.prod(region.shape)
    k_eps = 3 if region.material_sampling == 'yee' else 1
    psi = psi_elements(region)
    return real_bytes*(12*cells+k_eps*cells+2*psi), psi/cells


def cpu_baseline(seconds_budget, mask):
    """Oracle (NumPy, one thread) on a 512x512x64 slab of the cube with the same CPML: cell-updates/s."""
    from oracle.scene import Scene
    from oracle.yee import Grid
    nz = 64
    h = .04
    project = {'region': {'dimension': '3d', 'size': (512*h, 512*h, nz*h), 'mesh': h, 'pml_cells': 16, 'steps': 10}}
    scene = Scene(project)
    inv = np.where(mask[:, :, :nz].cpu().numpy(), np.float32(1)/np.float32(12), np.float32(1)).astype(np.float32)
    g = Grid(scene, np.repeat(inv[..., None], 3, axis=-1))
    rng = np.random.default_rng(1)
    g.E[:] = rng.normal(size=g.E.shape).astype(np.float32)*1e-2
    g.H[:] = rng.normal(size=g.H.shape).astype(np.float32)*1e-2
    g.update_E()
    g.update_H()
    steps, t0 = 0, time.perf_counter()
    while True:
        g.update_E()
        g.update_H()
        steps += 1
        elapsed = time.perf_counter()-t0
        if elapsed > seconds_budget or steps >= 20:
            break
    cells = math.prod(scene.shape)
    return dict(value=cells*steps/elapsed/1e9, unit='Gcell-updates/s', cores=1, kind='port',
                sample=f'NumPy oracle of the reference CPU path (FP32, one thread of {os.cpu_count()}), '
                       f'512x512x{nz} slab of the 512^3 cube with its 16-cell CPML, {steps} timed steps after 1 warm-up, '
                       f'{elapsed:.1f} s')


def counted_traffic(n, two_step=False):
    """HBM bytes per step of this cube (the two-step pass: half of one launch) from the committed
    counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a plain copy's counts), or None
    for another size."""
    path = os.path.join(ROOT, 'profiles', 'march_512cube_hbm_counters.md')
    if n != 512 or not os.path.exists(path):
        return None
    key = 'two_step_traffic_bytes_per_step:' if two_step else 'traffic_bytes_per_step:'
    for line in open(path):
        if line.startswith(key):
            return int(float(line.split(':')[1]))
    return None


def stream_copy_rate(torch, gib=4, reps=10):
    """GB/s (bytes read + written) of a plain device-to-device copy of `gib` GiB of float32 on this GPU:
    the practical HBM ceiling the step's achieved bandwidth is also compared with."""
    src = torch.ones(int(gib*2**30)//4, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    rate = 2*src.numel()*4*reps/(e0.elapsed_time(e1)/1e3)/1e9
    del src, dst
    torch.cuda.empty_cache()
    return rate


def batched_line(steps, count, loop_grids, launch='tune', width=50):
    """The uneven batch: 64 grids of 64x64x32 .. 192x160x64 cells, 2000 steps, autotuned launch,
    against the per-grid loop of Simulation.run (both with graphs of `width` steps)."""
    from torchfdtd_amd import workloads as W
    from torchfdtd_amd.batch import Launch, run_tensor_batch, tune_tensor_batch
    from torchfdtd_amd.solver import Simulation
    cases = W.uneven_batch(count, steps)
    cells = sum(math.prod(c.project.region.shape) for c in cases)
    print(f'bench: uneven batch, {count} grids, {cells} cells', file=sys.stderr, flush=True)
    if launch == 'tune':
        tuning = tune_tensor_batch(cases, candidates=(Launch('split'), Launch('march', 8, 2), Launch('march', 8, 3),
                                                      Launch('march', 8, 4), Launch('march', 8, 6), Launch('march', 16, 3),
                                                      Launch('march', 16, 6)), repeats=2, trial_steps=100)
        chosen, timings = tuning.launch, {k: round(v['median_seconds'], 5) for k, v in tuning.timings.items()}
    else:
        chosen, timings = Launch(launch), None
    report = run_tensor_batch(cases, launch=chosen, cuda_graph_steps=width, keep_results=False)
    batched = cells*steps/report.plan['loop_seconds']/1e9
    print(f'bench: batched {batched:.2f} Gcell/s with {chosen}; per-grid loop over {loop_grids} grids',
          file=sys.stderr, flush=True)
    loop_seconds, loop_cells = 0., 0
    for c in cases[:loop_grids]:
        result = Simulation(c.project, epsilon=c.epsilon).run(cuda_graph_steps=width)
        loop_seconds += result.summary['seconds']
        loop_cells += math.prod(c.project.region.shape)
    loop = loop_cells*steps/loop_seconds/1e9 if loop_grids else None
    step_bytes = sum(algorithmic_bytes(c.project.region)[0] for c in cases)
    plane_bytes = sum(plane_bytes_per_step(c.project) for c in cases)
    per_step = report.plan['loop_seconds']/steps
    achieved = (step_bytes+plane_bytes)/per_step/1e9
    # Counted HBM bytes per step of this cohort and launch, from the committed counter summaries.
    traffic, source = None, None
    summaries = {('split',): ('batch_uneven64_kernel_stats.md', 'split_traffic_bytes_per_step:'),
                 ('march', 8, 6): ('batch_uneven64_march_kernel_stats.md', 'march_8_6_traffic_bytes_per_step:')}
    name, key = summaries.get((chosen.kernel,) if chosen.kernel == 'split' else (chosen.kernel, chosen.tile_y, chosen.chunk_x),
                              (None, None))
    path = os.path.join(ROOT, 'profiles', name) if name else None
    if count == 64 and path and os.path.exists(path):
        for line in open(path):
            if line.startswith(key):
                traffic, source = int(line.split(':')[1]), f'profiles/{name}'
    roofline = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved/HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=(f'copied from {source} (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, '
                                    'scaled by a plain copy), not measured in this run' if traffic else None),
                    kernel=f'batched {chosen.kernel} step of all {count} grids plus their DFT planes; duration = '
                           'per-step span of the timed cohort loop',
                    bytes_per_cell_update=round((step_bytes+plane_bytes)/cells, 2),
                    field_bytes_per_cell_update=round(step_bytes/cells, 2))
    return dict(workload=f'{count} uneven 3D grids 64x64x32..192x160x64, 10-cell CPML, {steps} steps, Ez dipole, '
                         '16-frequency float64 DFT plane, binary design block', cells=cells,
                value=round(batched, 3), unit='Gcell-updates/s', loop_value=round(loop, 3) if loop else None,
                loop_grids=loop_grids, speedup_vs_loop=round(batched/loop, 3) if loop else None, launch=chosen.as_dict(),
                tuning=timings, roofline=roofline,
                loop_seconds=round(report.plan['loop_seconds'], 4))


def plane_bytes_per_step(project):
    """Algorithmic bytes per step of a project's frequency planes: each complex accumulator element
    read and written once per accumulation block, and 4 bytes per sampled interpolation corner."""
    from torchfdtd_amd.monitors import frequency_samples, plane_plan
    r = project.region
    total = 0.
    for raw in project.monitors:
        if not raw.enabled or raw.kind != 'field':
            continue
        m = project.resolved_monitor(raw)
        points = len(plane_plan(r, m)['weights'])
        comps = len(m.required_fields)
        nf = len(frequency_samples(m.spectrum))
        spec = 16 if (r.precision == 'float64' or m.dft_precision == 'float64') else 8
        block = max(1, min(32, 65536//(4*(spec//2)*nf)))
        corners = 8 if r.dimension == '3d' else 4
        total += 2*spec*nf*points*comps/block+4*corners*points*comps
    return total


def observed_cpu_baseline(project):
    """The reference's CPU path (the NumPy oracle, one thread) on an observed project: the whole run for
    a 2D guide, a bounded sample of its first 150 steps for a 3D one (about 15 s).  Returns (steps,
    seconds, Gcell-updates/s)."""
    from oracle import simulate
    r = project.region
    steps = r.steps if r.dimension == '2d' else min(r.steps, 150)
    t0 = time.perf_counter()
    simulate.run(project.model_dump(), steps=steps, record_frames=False)
    seconds = time.perf_counter()-t0
    return steps, seconds, math.prod(r.shape)*steps/seconds/1e9


def observed_cube_line(steps=96, width=48):
    """The 512^3 cube's workload (16-cell CPML, seeded permittivity, Ez dipole) with an x-normal
    16-frequency float64 DFT plane (E and H, 400 x 400 points) and an Ez and an Hy trace near high x:
    per-step time on the default path, the two-step pass (its observers sample step n+1 from a
    snapshot of the cells they read), and on the one-step kernel, both in graphs of `width`
    steps after a warm-up graph."""
    import torch

    from torchfdtd_amd import FieldMonitor, Monitor
    from torchfdtd_amd import workloads as W
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.models import SpectrumSettings
    from torchfdtd_amd.solver import Stepper
    project = W.cube(512, steps=steps+width)
    spectrum = SpectrumSettings(sampling='wavelength', frequency_points=16, wavelength_start=1.45, wavelength_stop=1.65,
                                apodization='none')
    project.monitors = [FieldMonitor(id='plane', center=(8., 0, 0), normal='x', size=(0, 16., 16.), dft_precision='float64',
                                     spectrum=spectrum),
                        Monitor(id='ez', component='Ez', center=(8., 1., 0)), Monitor(id='hy', component='Hy', center=(8.04, 0, 1.))]
    project = type(project).model_validate(project.model_dump())
    inverse = W.cube_inverse_permittivity(W.cube_noise_mask(512))[..., None]
    cells = math.prod(project.region.shape)
    out = dict(workload='512^3 cube, 16-cell CPML, Ez dipole, x-normal 16-frequency float64 DFT plane (400 x 400 points, '
                        'E and H) and two traces near high x', steps=steps)
    for name, kw in (('default', {}), ('one_step', dict(two_step=False))):
        print(f'bench: observed cube, {name} path', file=sys.stderr, flush=True)
        run = Stepper(project, **kw)
        run.grid.inverse_permittivity.copy_(inverse)
        graphs = StepGraphs(run.step, run.states(), project.region.steps, width, pair=run.step_pair,
                            prepare=run.refresh_materials, quad=run.quad if run.two_step else None)
        graphs.advance(width)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        graphs.advance(steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter()-t0)/steps*1e3
        out[name] = dict(path=run.path, ms_per_step=round(ms, 4), value=round(cells/ms/1e6, 3), unit='Gcell-updates/s',
                         snapshot_cells=int(run.snap_e.shape[0]) if run.two_step and run.snapshot_boxes else None)
        del run, graphs
        torch.cuda.empty_cache()
    out['speedup'] = round(out['one_step']['ms_per_step']/out['default']['ms_per_step'], 3)
    return out


def observed_lines(width=50, repeats=3):
    """The 160x120 2D TM waveguide (point trace, 32-frequency spectrum) and the 160x120x40 3D
    waveguide (point trace), 1000 steps each through Simulation.run (graphs of `width` steps):
    Gcell-updates/s (median of `repeats` runs after a warm-up), and the FP32 error against an FP64
    run of the same project (the FP64 engine equals the FP64 oracle of the reference CPU path bit for
    bit), relative L2 over the whole fields, the trace and the spectrum."""
    import statistics

    import numpy as np

    from torchfdtd_amd import workloads as W
    from torchfdtd_amd.solver import Simulation
    from torchfdtd_amd.monitors import point_spectrum
    out = {}
    for name, make in (('waveguide_2d_160x120', W.waveguide_2d), ('waveguide_3d_160x120x40', W.waveguide_3d)):
        project = make()
        r = project.region
        print(f'bench: {name}', file=sys.stderr, flush=True)
        Simulation(project).run(cuda_graph_steps=width)
        runs = [Simulation(project).run(cuda_graph_steps=width) for _ in range(repeats)]
        seconds = statistics.median(run.summary['seconds'] for run in runs)
        cells = math.prod(r.shape)
        exact = project.model_copy(deep=True)
        exact.region.precision = 'float64'
        ref = Simulation(exact).run(cuda_graph_steps=width)
        got = runs[0]
        rel = lambda a, b: float(np.linalg.norm((np.asarray(a)-np.asarray(b)).ravel())/np.linalg.norm(np.asarray(b).ravel()))
        m = project.resolved_monitor(project.monitors[0])
        settings = m.spectrum.model_copy(update=dict(sampling='wavelength', frequency_points=32, wavelength_start=1.5,
                                                     wavelength_stop=1.6, apodization='none'))
        spectrum = point_spectrum(got.times, got.signals[:, 0].astype(np.float64), settings)['value']
        exact_spectrum = point_spectrum(ref.times, ref.signals[:, 0], settings)['value']
        step_bytes = algorithmic_bytes(r)[0] if r.dimension == '3d' else 4*(6*cells+cells)+8*psi_tm(r)
        cpu_steps, cpu_seconds, cpu_rate = observed_cpu_baseline(project)
        out[name] = dict(value=round(cells*r.steps/seconds/1e9, 3), unit='Gcell-updates/s', steps=r.steps, cells=cells,
                         seconds=round(seconds, 4), samples_seconds=[round(run.summary['seconds'], 4) for run in runs],
                         steps_per_pass=got.summary['steps_per_pass'],
                         cpu_baseline=dict(value=round(cpu_rate, 5), unit='Gcell-updates/s', cores=1, kind='port',
                                           sample=f'NumPy oracle, one thread, first {cpu_steps} of {r.steps} steps, '
                                                  f'{cpu_seconds:.1f} s'),
                         gpu_over_cpu=round(cells*r.steps/seconds/1e9/cpu_rate, 1),
                         effective_gbs=round(step_bytes*r.steps/seconds/1e9, 1),
                         fp32_vs_fp64_rel_l2=dict(E=rel(got.electric, ref.electric), H=rel(got.magnetic, ref.magnetic),
                                                  trace=rel(got.signals, ref.signals),
                                                  spectrum=rel(spectrum, exact_spectrum)))
    return out


def psi_tm(region):
    """CPML memory elements of the TM components of a 2D grid (Ez from Hx/Hy, Hx/Hy from Ez)."""
    from torchfdtd_amd.planning import cpml_slabs
    total = sum(math.prod(s.psi_shape(region.shape)) for (_, axis, _), slabs in cpml_slabs(region).items() if axis != 2
                for s in slabs)
    return total/2      # of each axis's two terms per family, one feeds a TM component


def surrogate_line(batch=128, reps=5):
    """One training step (forward, data and weight gradients, AdamW) of the 1.2M-parameter 3D conv
    surrogate on a batch of 128 densities 64x64x8, on the BF16-MFMA implicit-GEMM kernels: dense
    conv FLOPs (forward + dgrad + wgrad) per second against the 2.5 PFLOP/s BF16 MFMA peak."""
    import torch

    from torchfdtd_amd.surrogate import SurrogateNet
    net = SurrogateNet().cuda()
    opt = torch.optim.AdamW(net.parameters(), lr=3e-4)
    rho = torch.rand((batch, 64, 64, 8), device='cuda')
    target = torch.rand((batch, 32), device='cuda')
    flops, dims, ci = 0, (64, 64, 8), 1
    for w in net.weights:
        co = w.shape[0]
        dims = tuple((n+2-3)//2+1 for n in dims)
        flops += 3*2*batch*math.prod(dims)*co*27*ci      # forward, dgrad and wgrad
        ci = co
    flops += 3*2*batch*(256*256+256*32)

    def step():
        opt.zero_grad(set_to_none=True)
        torch.nn.functional.mse_loss(net(rho), target).backward()
        opt.step()
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize()
    seconds = e0.elapsed_time(e1)/1e3/reps
    # The kernels on one large convolution (64 x 32x32x16 x 64 -> 128 channels, 3x3x3, stride 1): the
    # implicit GEMM's own rate, forward and forward + backward (dgrad + wgrad), against the peak.
    from torchfdtd_amd.surrogate import conv3d
    gen = torch.Generator(device='cuda').manual_seed(5)
    x = torch.randn((64, 32, 32, 16, 64), device='cuda', generator=gen).requires_grad_(True)
    w = (torch.randn((128, 3, 3, 3, 64), device='cuda', generator=gen)*.02).requires_grad_(True)
    b = torch.zeros(128, device='cuda', requires_grad=True)
    gemm = 2*64*32*32*16*128*27*64
    rates = {}
    for label, backward in (('forward', False), ('forward_dgrad_wgrad', True)):
        for _ in range(2):
            y = conv3d(x, w, b, 1, 1, True)
            if backward:
                y.sum().backward()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            y = conv3d(x, w, b, 1, 1, True)
            if backward:
                y.sum().backward()
        e1.record()
        torch.cuda.synchronize()
        rates[label] = round(gemm*(3 if backward else 1)*reps/(e0.elapsed_time(e1)/1e3)/1e12, 1)
    del x, w, b, y
    torch.cuda.empty_cache()
    return dict(workload='SurrogateNet training step: 4 stride-2 3x3x3 convs 1->32->64->128->256 on 128x64x64x8, MLP head',
                parameters=net.parameter_count(), ms_per_step=round(seconds*1e3, 3),
                conv_tflops=round(flops/seconds/1e12, 2), peak_tflops=2500.0,
                frac=round(flops/seconds/1e12/2500.0, 5), dtype='bf16 MFMA, fp32 accumulate',
                large_conv=dict(shape='64x32x32x16x64 -> 128, 3x3x3, stride 1', gemm_tflop=round(gemm/1e12, 3),
                                tflops=rates, frac={k: round(v/2500.0, 4) for k, v in rates.items()}),
                study=surrogate_study())


def surrogate_study():
    """The design study's summary (training set size, held-out error, and per campaign the unscreened
    solves needed to match the screened loop after 32, 96 and 192 solves), read from the committed
    record of its last run: the study takes about 25 GPU-minutes and is not rerun here."""
    path = Path(__file__).resolve().parent/'torchfdtd_amd'/'surrogate_study.json'
    if not path.exists():
        return None
    record = json.loads(path.read_text())
    return dict(record, source=f'read from {path.name} (python -m torchfdtd_amd.surrogate study, '
                               f'{record.get("measured", "MI355X")}); not measured in this run')


def adjoint_line(steps=2000):
    """Discrete-adjoint gradient on a 200x160x60 grid, 50x50x6 density design, |DFT Ez|^2 on a plane at 8 wavelengths."""
    import torch
    from torchfdtd_amd.adjoint import AdjointOptions, DifferentiablePlaneSimulation
    from torchfdtd_amd.density import cell_density_map
    from torchfdtd_amd.workloads import adjoint_case
    project, box, rho, frequency = adjoint_case(steps)
    from torchfdtd_amd.solver import Simulation
    model = DifferentiablePlaneSimulation(project, AdjointOptions(design_box=box))
    base = torch.full(tuple(project.region.shape), 2.085, device='cuda')
    with torch.no_grad():
        plain_eps = cell_density_map(torch.as_tensor(rho, device='cuda'), base, box, background_epsilon=2.085,
                                     design_epsilon=12.11)

    def plain_forward():
        # The plain forward run of the same design (frequency plane, graphs of 50 steps, default diagnostics).
        torch.cuda.synchronize()
        started = time.perf_counter()
        Simulation(project, epsilon=plain_eps).run(cuda_graph_steps=50)
        torch.cuda.synchronize()
        return time.perf_counter()-started

    def gradient():
        density = torch.as_tensor(rho, device='cuda').requires_grad_(True)
        torch.cuda.synchronize()
        started = time.perf_counter()
        eps = cell_density_map(density, base, box, background_epsilon=2.085, design_epsilon=12.11)
        planes = model(eps, frequency)
        objective = (planes['out'].fields[..., 2].abs()**2).sum()
        objective.backward()
        torch.cuda.synchronize()
        return time.perf_counter()-started, planes, objective, density

    # One untimed run of each first: library, kernel and BLAS initialisation are not part of a gradient.
    plain_forward()
    gradient()
    plain_seconds = plain_forward()
    total, planes, objective, recorded = gradient()
    report = planes['out'].report
    # The configured schedule: checkpoints every 45 steps, replayed in the reverse sweep.
    model = DifferentiablePlaneSimulation(project, AdjointOptions(design_box=box, checkpoint_interval=45))
    gradient()
    configured, configured_planes, _, replayed = gradient()
    configured_report = configured_planes['out'].report
    return dict(workload=f'200x160x60 cells at 0.04 um, {steps} steps, 12-cell CPML, 50x50x6 density design, '
                         'objective sum |DFT Ez|^2 over a plane at 8 wavelengths 1.50-1.60 um',
                seconds_per_gradient=round(total, 4), plain_forward_seconds=round(plain_seconds, 4),
                gradient_over_plain_forward=round(total/plain_seconds, 3), forward_seconds=round(report['forward_seconds'], 4),
                backward_seconds=round(report['backward_seconds'], 4),
                backward_over_forward=round(report['backward_seconds']/report['forward_seconds'], 3),
                checkpoint_interval=report['checkpoint_interval'], record_bytes=report['record_bytes'],
                observers=report['observers'], objective=float(objective.detach()),
                gradient_norm=float(recorded.grad.double().norm()),
                checkpoint_45=dict(seconds_per_gradient=round(configured, 4),
                                   gradient_over_plain_forward=round(configured/plain_seconds, 3),
                                   bound=2.5, checkpoint_interval=configured_report['checkpoint_interval'],
                                   checkpoints=configured_report['checkpoints'],
                                   replayed_steps=configured_report['replayed_steps'],
                                   record_bytes=configured_report['record_bytes'],
                                   gradient_rel_diff_vs_record_all=float((replayed.grad.double()-recorded.grad.double()).norm()
                                                                         / recorded.grad.double().norm())))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--n', type=int, default=512)
    ap.add_argument('--graph-width', type=int, default=50)
    ap.add_argument('--cpu-seconds', type=float, default=15.)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-batched', action='store_true')
    ap.add_argument('--no-adjoint', action='store_true')
    ap.add_argument('--adjoint-steps', type=int, default=2000, help='steps of the adjoint configuration (2000; fewer for profiling)')
    ap.add_argument('--no-observed', action='store_true', help='skip the observed lines (waveguides with their FP64 error, '
                    'the 512^3 cube with a plane monitor and traces)')
    ap.add_argument('--no-waveguides', action='store_true', help='of the observed lines, run only the 512^3 cube')
    ap.add_argument('--no-surrogate', action='store_true')
    ap.add_argument('--repeats', type=int, default=3, help='further timed samples of --steps steps (median reported)')
    ap.add_argument('--no-tune', action='store_true', help='keep the default launch shape of the cube step')
    ap.add_argument('--batch-grids', type=int, default=64)
    ap.add_argument('--batch-steps', type=int, default=2000)
    ap.add_argument('--batch-launch', choices=('tune', 'split', 'march'), default='tune')
    ap.add_argument('--no-cube', action='store_true', help='skip the 512^3 cube (profiling the other parts alone)')
    ap.add_argument('--one-step', action='store_true', help='time the one-step kernel (one step per pass over HBM)')
    ap.add_argument('--loop-grids', type=int, default=64)
    ap.add_argument('--copy-gib', type=float, default=0,
                    help='only copy a buffer of this many GiB three times (the calibration of HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: the fused Yee step runs the whole grid on one MI355X '
              f'(--gpus {args.gpus} is not run; see the sharding paragraph of the design notes).', file=sys.stderr)
        return 2
    import torch
    if args.copy_gib:
        count = int(args.copy_gib*2**30)//4
        src = torch.ones(count, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_read_per_launch': count*4, 'copy_bytes_written_per_launch': count*4, 'launches': 3}))
        return 0
    if args.no_cube:
        line = {}
        extra_lines(args, line)
        print(json.dumps(line), flush=True)
        return 0
    from torchfdtd_amd import _native as N
    from torchfdtd_amd.graphs import StepGraphs
    from torchfdtd_amd.solver import Stepper
    from torchfdtd_amd.workloads import cube, cube_inverse_permittivity, cube_noise_mask

    project = cube(n=args.n, steps=args.steps+args.warmup+8)
    region = project.region
    print(f'bench: building {args.n}^3 cube, {region.pml_cells}-cell CPML', file=sys.stderr, flush=True)
    mask = cube_noise_mask(args.n)
    run = Stepper(project, two_step=None if not args.one_step else False)
    g = run.grid
    g.inverse_permittivity.copy_(cube_inverse_permittivity(mask)[..., None])
    torch.cuda.synchronize()
    tuning = None
    if not args.no_tune:
        # The grid's launch shapes measured on this box (bitwise-identical candidates only).
        from torchfdtd_amd.tuning import ONE_PASS, tune_launch
        print('bench: tuning the launch shape', file=sys.stderr, flush=True)
        # --one-step keeps the one-step kernel: only its shapes are candidates.
        only = [('one_pass', dict(tile_y=t, chunk_x=c)) for t, c in ONE_PASS] if args.one_step else None
        tuning = tune_launch(run, candidates=only, sweeps=3, units=2)
    # Graphs of whole fours keep every step on the two-step pass (a graph of 50 steps ends in a pair
    # of one-step launches).
    width = args.graph_width if not run.two_step else args.graph_width // 4 * 4
    graph = StepGraphs(run.step, run.states(), region.steps, width, warmup=1,
                       pair=run.step_pair if run.one_pass else None,
                       prepare=run.refresh_materials if run.one_pass else None,
                       quad=run.quad if run.two_step else None)
    print(f'bench: graphs captured (one-pass kernel: {run.one_pass}, two steps per pass: {run.two_step})',
          file=sys.stderr, flush=True)

    def advance(steps):
        while steps >= width:
            graph.replay(width)
            steps -= width
        while run.two_step and steps >= 4:
            graph.replay(4)
            steps -= 4
        for _ in range(steps):
            graph.replay(1)

    advance(args.warmup)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall = time.perf_counter()
    start.record()
    advance(args.steps)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter()-wall
    seconds = start.elapsed_time(stop)/1e3
    # Further samples of the same number of steps (the state keeps evolving; every step costs the same).
    samples = [seconds]
    for _ in range(args.repeats-1):
        start.record()
        advance(args.steps)
        stop.record()
        torch.cuda.synchronize()
        samples.append(start.elapsed_time(stop)/1e3)
    cells = math.prod(region.shape)
    value = cells*args.steps/seconds/1e9
    assert torch.isfinite(g.E).all() and torch.isfinite(g.H).all()

    # Per-kernel durations with HIP events on the launch stream (the kernels are launched on torch's current stream).
    lib = N.library()
    stream = N.stream_handle()
    kernel_ms = {}
    reps = 20
    launches = [('yee_half_step_E', lambda: lib.tfdtd_yee_half_step(ctypes.byref(run.e_step), stream)),
                ('yee_half_step_H', lambda: lib.tfdtd_yee_half_step(ctypes.byref(run.h_step), stream))]
    if run.one_pass:
        launches.append(('march', lambda k: run.march_step(k % 2, stream)))
    if run.two_step:
        launches.append(('march2_per_step', lambda k: run.march2_step(k % 2, stream)))
    for name, launch in launches:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(reps):
            if name.startswith('march'):
                launch(k)
            else:
                N.check(launch(), name)
        e1.record()
        torch.cuda.synchronize()
        kernel_ms[name] = e0.elapsed_time(e1)/reps/(2 if name == 'march2_per_step' else 1)
    copy_gbs = stream_copy_rate(torch)
    step_bytes, psi_per_cell = algorithmic_bytes(region)
    if run.one_pass:
        # The step is three concurrent tile launches (fork/join on the launch stream): its duration is
        # the per-step span of the timed graph replays, from HIP events on the stream they run on.
        per_step_ms = seconds*1e3/args.steps
        kernel = ('one-pass Yee step: interior, z-face and boundary tile launches running concurrently; '
                  'duration = per-step span of the timed region')
        if run.two_step:
            kernel = ('two-step Yee pass (two steps per pass over HBM): interior, z-face and boundary tile launches '
                      'running concurrently; duration = per-step span of the timed region')
    else:
        per_step_ms, kernel = kernel_ms['yee_half_step_E']+kernel_ms['yee_half_step_H'], 'yee_half_step E+H'
    achieved = step_bytes/(per_step_ms/1e3)/1e9
    traffic = counted_traffic(args.n, run.two_step) if run.one_pass else None
    roofline = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved/HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=('copied from profiles/march_512cube_hbm_counters.md (rocprofv3 --pmc FETCH_SIZE and '
                                    'WRITE_SIZE per dispatch, scaled by a plain copy), not measured in this run')
                    if traffic else None,
                    kernel=kernel, kernel_ms=kernel_ms, stream_copy=round(copy_gbs, 1),
                    frac_of_copy=round(achieved/copy_gbs, 4),
                    bytes_per_cell_update=round(step_bytes/cells, 2), psi_per_cell=round(psi_per_cell, 4))
    baseline = None
    if not args.no_cpu_baseline:
        print('bench: CPU baseline (NumPy oracle)', file=sys.stderr, flush=True)
        baseline = cpu_baseline(args.cpu_seconds, mask)
    line = {'metric': 'Gcell-updates/s (fused 3D Yee+CPML step, 512^3 FP32, %HBM in roofline)',
            'value': round(value, 3), 'unit': 'Gcell-updates/s', 'n_gpus': 1, 'steps': args.steps,
            'warmup': args.warmup, 'ms_per_step': round(seconds*1e3/args.steps, 4), 'higher_is_better': True,
            'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded smoothed-threshold eps noise)',
            'config': {'workload': f'{args.n}^3 cube, 16-cell CPML all faces, eps=1+11s per cell, Ez dipole, '
                                   f'{region.steps - 8} steps of which {args.steps} timed',
                       'cells': cells, 'graph_steps': width, 'wall_s': round(wall, 3),
                       'launch': run.launch_shape(), 'launch_tuning': tuning},
            'samples_ms_per_step': [round(t*1e3/args.steps, 4) for t in samples],
            'median_ms_per_step': round(sorted(samples)[len(samples)//2]*1e3/args.steps, 4),
            'roofline': roofline, 'cpu_baseline': baseline}
    del graph, run, g, mask
    torch.cuda.empty_cache()
    extra_lines(args, line)
    print(json.dumps(line), flush=True)
    return 0


def extra_lines(args, line):
    """The other configurations' lines, each under its own key of the one JSON line."""
    if not args.no_batched:
        line['batched'] = batched_line(args.batch_steps, args.batch_grids, args.loop_grids, args.batch_launch)
    if not args.no_observed:
        line['observed'] = {} if args.no_waveguides else observed_lines()
        line['observed']['cube_512_plane_and_traces'] = observed_cube_line()
    if not args.no_adjoint:
        print('bench: adjoint gradient', file=sys.stderr, flush=True)
        line['adjoint'] = adjoint_line(args.adjoint_steps)
    if not args.no_surrogate:
        print('bench: surrogate training step', file=sys.stderr, flush=True)
        line['surrogate'] = surrogate_line()


if __name__ == '__main__':
    sys.exit(main())
