"""The named workloads of the engine's benchmarks and tests, with their sizes.

* ``waveguide_2d``: 2D TM waveguide, 8x6 um at 0.05 um (160x120 cells), 1000 steps,
  10-cell CPML, eps 2.085 background with a 0.65 um eps 12.11 core, broadband Ez point
  source at (-2.5, 0) for 1.50-1.60 um, point monitor at (2, 0) with a 32-wavelength DFT.
* ``waveguide_3d``: 8x6x2 um at 0.05 um (160x120x40), 0.65x0.4 um core, 12-cell CPML,
  Ey dipole, 1000 steps.
* ``cube_512``: 512^3 cells, 16-cell CPML on all faces, 500 steps, per-cell
  eps = 1 + 11 s with s = [Gaussian(sigma = 3 cells) * U(0,1) noise > 0.5] (periodic
  filter, fixed seed), Ez dipole at the centre.
"""
from __future__ import annotations

import math

import numpy as np

from .models import Material, Monitor, Project, Region, Source, SpectrumSettings, Structure

CUBE_SEED = 20260512


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
