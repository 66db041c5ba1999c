"""The named workloads of the engine's benchmarks and tests, with their sizes.

* ``waveguide_2d``: 2D TM waveguide, 8x6 um at 0.05 um (160x120 cells), 1000 steps,
  10-cell CPML, eps 2.085 background with a 0.65 um eps 12.11 core, broadband Ez point
  source at (-2.5, 0) for 1.50-1.60 um, point monitor at (2, 0) with a 32-wavelength DFT.
* ``waveguide_3d``: 8x6x2 um at 0.05 um (160x120x40), 0.65x0.4 um core, 12-cell CPML,
  Ey dipole, 1000 steps.
* ``cube_512``: 512^3 cells, 16-cell CPML on all faces, 500 steps, per-cell
  eps = 1 + 11 s with s = [Gaussian(sigma = 3 cells) * U(0,1) noise > 0.5] (periodic
  filter, fixed seed), Ez dipole at the centre.
* ``uneven_batch``: 64 3D grids with sizes drawn uniformly from 64x64x32 to 192x160x64
  cells at 0.04 um, 10-cell CPML, 2000 steps; each has an Ez dipole (1.45-1.65 um
  broadband pulse around 1.55 um) 12 cells inside its low-x edge, a 16-wavelength
  float64 DFT plane 12 cells inside its high-x edge over the CPML-free cross-section,
  and a 32x32x8-cell binary design block at its centre (eps 2.085 or 12.11 from
  seeded U(0,1) noise, conic filter of radius 0.1 um, threshold 0.5).
* ``adjoint_case``: 200x160x60 cells at 0.04 um, 2000 steps, 12-cell CPML, an Ez dipole
  2.4 um left of centre, a 50x50x6-cell design box at the centre with seeded uniform density
  rho (eps = 2.085 + 10.025 rho) and an x-normal float64 DFT plane 2.4 um right of centre over
  the CPML-free cross-section, sampled at 8 wavelengths 1.50-1.60 um.
* ``design_slab``: the surrogate's data and arbiter.  A 64x64x8-cell density slab
  (eps = 2.085 + 10.025 rho) in a 96x96x52-cell grid at 0.04 um with 10-cell CPML, lit by an
  Ex plane source below it (1.45-1.65 um broadband) and measured by a 32-wavelength float64
  DFT plane above it; the transmission is the plane's flux over that of the empty slab.
"""
from __future__ import annotations

import math

import numpy as np

from .models import FieldMonitor, Material, Monitor, Project, Region, Source, SpectrumSettings, Structure

CUBE_SEED = 20260512
BATCH_SEED = 20260513
EPS_LOW, EPS_HIGH = 2.085, 12.11


def waveguide_2d(steps=1000, backend='cuda'):
    return Project(name='2D TM waveguide 160x120',
                   region=Region(dimension='2d', size=(8, 6, .05), mesh=.05, steps=steps, pml_cells=10,
                                 background_index=math.sqrt(2.085), backend=backend, snapshot_interval=50),
                   materials=[Material(name='core', index=math.sqrt(12.11))],
                   structures=[Structure(id='core', size=(8, .65, 1), material='core')],
                   sources=[Source(id='src', center=(-2.5, 0, 0), component='Ez', pulse='broadband', time_definition='wavelength',
                                   wavelength_start=1.5, wavelength_stop=1.6)],
                   monitors=[Monitor(id='probe', center=(2, 0, 0), spectrum=SpectrumSettings(
                       sampling='wavelength', frequency_points=32, wavelength_start=1.5, wavelength_stop=1.6,
                       apodization='none'))])


def waveguide_3d(steps=1000, backend='cuda'):
    return Project(name='3D waveguide 160x120x40',
                   region=Region(dimension='3d', size=(8, 6, 2), mesh=.05, steps=steps, pml_cells=12,
                                 background_index=math.sqrt(2.085), backend=backend, snapshot_interval=50),
                   materials=[Material(name='core', index=math.sqrt(12.11))],
                   structures=[Structure(id='core', size=(8, .65, .4), material='core')],
                   sources=[Source(id='src', center=(-2.5, 0, 0), component='Ey', pulse='broadband', time_definition='wavelength',
                                   wavelength_start=1.5, wavelength_stop=1.6)],
                   monitors=[Monitor(id='probe', center=(2, 0, 0), component='Ey')])


def cube(n=512, steps=500, pml=16, mesh=.04, backend='cuda', precision='float32'):
    return Project(name=f'{n}^3 roofline cube',
                   region=Region(dimension='3d', size=(n*mesh,)*3, mesh=mesh, steps=steps, pml_cells=pml,
                                 backend=backend, precision=precision, snapshot_interval=max(20, steps)),
                   sources=[Source(id='src', center=(0, 0, 0), component='Ez')])


def cube_noise_mask(n=512, sigma=3., seed=CUBE_SEED, device='cuda'):
    """s in {0,1}: U(0,1) noise (NumPy PCG64, `seed`) smoothed by a periodic Gaussian of `sigma` cells, > 0.5."""
    import torch
    noise = np.random.default_rng(seed).random((n, n, n), dtype=np.float32)
    u = torch.as_tensor(noise, device=device)
    spectrum = torch.fft.rfftn(u)
    k = [torch.fft.fftfreq(n, device=device), torch.fft.fftfreq(n, device=device), torch.fft.rfftfreq(n, device=device)]
    gain = [torch.exp(-2*(math.pi*sigma*f)**2) for f in k]
    spectrum *= gain[0][:, None, None]*gain[1][None, :, None]*gain[2][None, None, :]
    smooth = torch.fft.irfftn(spectrum, s=(n, n, n))
    return (smooth > .5)


def cube_inverse_permittivity(mask):
    """1/eps for eps = 1 + 11 s, in FP32 with the correctly rounded reciprocal 1/12."""
    import torch
    inv12 = float(np.float32(1)/np.float32(12))
    return torch.where(mask, torch.tensor(inv12, device=mask.device), torch.tensor(1., device=mask.device)).float()


def conic_filter(density, radius_cells):
    """Normalised conic filter (1 - d/r)+ over the x-y plane of a 2D density, edge-padded."""
    reach = int(math.floor(radius_cells))
    offsets = np.arange(-reach, reach+1)
    weight = np.maximum(0., 1.-np.hypot(offsets[:, None], offsets[None, :])/radius_cells)
    weight /= weight.sum()
    padded = np.pad(density, reach, mode='edge')
    out = np.zeros_like(density, dtype=np.float64)
    for i, dx in enumerate(offsets):
        for j, dy in enumerate(offsets):
            if weight[i, j]:
                out += weight[i, j]*padded[reach+dx:reach+dx+density.shape[0], reach+dy:reach+dy+density.shape[1]]
    return out


def binary_design(rng, shape=(32, 32), radius_um=.1, mesh=.04, fill=None):
    """0/1 design: U(0,1) noise, conic filter of `radius_um`, threshold 0.5 (or, with `fill`, the
    quantile that sets that fraction of the cells to 1)."""
    smooth = conic_filter(rng.random(shape), radius_um/mesh)
    return smooth > (.5 if fill is None else np.quantile(smooth, 1-fill))


SLAB_FILL = (.2, .8)


def uneven_grid(shape, steps=2000, mesh=.04, pml=10, frequencies=16, precision='float32', backend='cuda', name='grid',
                inset_cells=12):
    """One grid of the uneven batch: Ez dipole near low x, float64 DFT plane near high x."""
    size = tuple(n*mesh for n in shape)
    inset = inset_cells*mesh
    return Project(name=name,
                   region=Region(dimension='3d', size=size, mesh=mesh, steps=steps, pml_cells=pml,
                                 background_index=math.sqrt(EPS_LOW), backend=backend, precision=precision,
                                 snapshot_interval=steps),
                   sources=[Source(id='src', center=(-size[0]/2+inset+mesh/2, 0, 0), component='Ez', pulse='broadband',
                                   time_definition='wavelength', wavelength_start=1.45, wavelength_stop=1.65)],
                   monitors=[FieldMonitor(id='plane', center=(size[0]/2-inset-mesh/2, 0, 0), normal='x',
                                          size=(0, size[1]-2*pml*mesh, size[2]-2*pml*mesh), dft_precision='float64',
                                          spectrum=SpectrumSettings(sampling='wavelength', frequency_points=frequencies,
                                                                    wavelength_start=1.45, wavelength_stop=1.65,
                                                                    apodization='none'))])


def design_epsilon(project, design, depth=8):
    """Per-cell eps of a grid with the binary x-y `design` extruded over `depth` z cells at its centre."""
    from .planning import voxelize
    eps = voxelize(project)
    nx, ny, nz = eps.shape
    dx, dy = design.shape
    x0, y0, z0 = (nx-dx)//2, (ny-dy)//2, (nz-depth)//2
    block = np.where(design, EPS_HIGH, EPS_LOW).astype(eps.dtype)
    eps[x0:x0+dx, y0:y0+dy, z0:z0+depth] = block[:, :, None]
    return eps


def uneven_shapes(count=64, seed=BATCH_SEED, low=(64, 64, 32), high=(192, 160, 64)):
    rng = np.random.default_rng(seed)
    return [tuple(int(rng.integers(lo, hi+1)) for lo, hi in zip(low, high)) for _ in range(count)]


def uneven_batch(count=64, steps=2000, seed=BATCH_SEED, frequencies=16, shapes=None, precision='float32'):
    """BatchCases of the uneven batch (sizes, designs and ids fixed by `seed`)."""
    from .batch import BatchCase
    shapes = uneven_shapes(count, seed) if shapes is None else shapes
    rng = np.random.default_rng(seed+1)
    cases = []
    for i, shape in enumerate(shapes):
        p = uneven_grid(shape, steps, frequencies=frequencies, precision=precision, name=f'uneven {i} {shape}')
        design = binary_design(rng)
        cases.append(BatchCase(f'grid-{i:02d}', p, dict(shape=list(shape)), epsilon=design_epsilon(p, design)))
    return cases


SLAB_SHAPE = (96, 96, 52)
SLAB_DESIGN = (64, 64, 8)


def design_slab(steps=2000, frequencies=32, mesh=.04, precision='float32', name='design slab'):
    """The design-slab project (no design: background eps 2.085 everywhere)."""
    nx, ny, nz = SLAB_SHAPE
    size = (nx*mesh, ny*mesh, nz*mesh)
    pml = 10
    inner = ((nx-2*pml)*mesh, (ny-2*pml)*mesh)
    spectrum = SpectrumSettings(sampling='wavelength', frequency_points=frequencies, wavelength_start=1.45,
                                wavelength_stop=1.65, apodization='none')
    return Project(name=name,
                   region=Region(dimension='3d', size=size, mesh=mesh, steps=steps, pml_cells=pml,
                                 background_index=math.sqrt(EPS_LOW), backend='cuda', precision=precision,
                                 snapshot_interval=steps),
                   sources=[Source(id='plane', kind='plane', center=(0, 0, -size[2]/2+(pml+4.5)*mesh), size=(*inner, 0),
                                   normal='z', component='Ex', pulse='broadband', time_definition='wavelength',
                                   wavelength_start=1.45, wavelength_stop=1.65)],
                   monitors=[FieldMonitor(id='out', center=(0, 0, size[2]/2-(pml+4.5)*mesh), normal='z',
                                          size=(*inner, 0), dft_precision='float64', spectrum=spectrum)])


def slab_epsilon(project, rho):
    """Per-cell eps of the slab grid with density rho (64, 64, 8) at its centre: 2.085 + 10.025 rho."""
    from .planning import voxelize
    eps = voxelize(project).astype(np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    nx, ny, nz = eps.shape
    dx, dy, dz = rho.shape
    x0, y0, z0 = (nx-dx)//2, (ny-dy)//2, (nz-dz)//2
    eps[x0:x0+dx, y0:y0+dy, z0:z0+dz] = EPS_LOW+(EPS_HIGH-EPS_LOW)*rho
    return eps.astype(np.float32 if project.region.precision == 'float32' else np.float64)


def slab_densities(count, seed, radius_um=.1, mesh=.04):
    """`count` binary slab densities: smoothed-threshold noise patterns (64x64) extruded over 8 cells,
    each with a fill fraction drawn uniformly from SLAB_FILL."""
    rng = np.random.default_rng(seed)
    out = np.empty((count, *SLAB_DESIGN), dtype=np.float32)
    for i in range(count):
        fill = rng.uniform(*SLAB_FILL)
        out[i] = binary_design(rng, SLAB_DESIGN[:2], radius_um, mesh, fill)[:, :, None]
    return out


ADJOINT_SEED = 20260514


def adjoint_case(steps=2000, seed=ADJOINT_SEED, precision='float32'):
    """(project, design box cells, rho (50, 50, 6), frequencies in Hz) of the adjoint benchmark."""
    mesh, shape, pml = .04, (200, 160, 60), 12
    size = tuple(n*mesh for n in shape)
    spectrum = SpectrumSettings(sampling='wavelength', frequency_points=8, wavelength_start=1.5, wavelength_stop=1.6,
                                apodization='none')
    project = Project(name='adjoint 200x160x60',
                      region=Region(dimension='3d', size=size, mesh=mesh, steps=steps, pml_cells=pml,
                                    background_index=math.sqrt(EPS_LOW), backend='cuda', precision=precision,
                                    snapshot_interval=steps),
                      sources=[Source(id='src', center=(-2.4, 0, 0), component='Ez', pulse='broadband',
                                      time_definition='wavelength', wavelength_start=1.5, wavelength_stop=1.6)],
                      monitors=[FieldMonitor(id='out', center=(2.4+mesh/2, 0, 0), normal='x',
                                             size=(0, size[1]-2*pml*mesh, size[2]-2*pml*mesh), dft_precision='float64',
                                             spectrum=spectrum)])
    design = (50, 50, 6)
    box = tuple(((n-d)//2, (n-d)//2+d) for n, d in zip(shape, design))
    rho = np.random.default_rng(seed).random(design).astype(np.float32 if precision == 'float32' else np.float64)
    from .monitors import frequency_samples
    return project, box, rho, frequency_samples(project.monitors[0].spectrum)
