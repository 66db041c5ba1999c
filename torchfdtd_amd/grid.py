"""Resident Yee state in torch-owned GPU memory and the HIP kernel object behind it.

``YeeGrid`` mirrors the reference grid's attributes (torchfdtd/boundaries.py:225-256):
``E``, ``H`` of shape (Nx, Ny, Nz, 3), ``inverse_permittivity``, ``inverse_permeability``,
``cpml[(forward, axis, comp)]`` slab lists and ``memory_states``.  The inverse
permittivity is stored per cell, (Nx, Ny, Nz, 1), when the material is sampled at
cell centres and per component, (Nx, Ny, Nz, 3), when it is sampled on the Yee
lattice; torch broadcasting makes both read like the reference array.

``HipYee`` is the drop-in for ``FusedYeeCUDA`` (torchfdtd/cuda_kernels.py:88-365):
``update_E()`` / ``update_H()`` launch the gfx950 kernels on torch's current stream,
mutate the grid in place, allocate nothing and never synchronise, so a HIP graph can
capture them.  ``configure_cuda_kernel`` patches them into the grid the way the
reference does (cuda_kernels.py:368-373).
"""
from __future__ import annotations

import ctypes
import math
import weakref

import numpy as np
import torch

from . import _native as N
from .planning import CURL_TERMS, cpml_slabs, pec_upper, periodic_axes, real_dtype


class YeeGrid:
    def __init__(self, region, device='cuda'):
        refusal = region.unsupported()
        if refusal:
            raise ValueError(refusal)
        if region.complex_fields:
            raise ValueError('Bloch (complex) fields are not run by this engine.')
        self.region = region
        self.Nx, self.Ny, self.Nz = region.shape
        self.dtype = torch.float64 if region.precision == 'float64' else torch.float32
        self.device = torch.device(device)
        shape = region.shape
        self.E = torch.zeros((*shape, 3), device=self.device, dtype=self.dtype)
        self.H = torch.zeros((*shape, 3), device=self.device, dtype=self.dtype)
        samples = 3 if region.material_sampling == 'yee' else 1
        self.inverse_permittivity = torch.ones((*shape, samples), device=self.device, dtype=self.dtype)
        self.inverse_permeability = torch.ones((1,), device=self.device, dtype=self.dtype)
        self.courant_number = region.rectangular_courant
        self.time_step = region.time_step
        self.is_torch = True
        self.wrap = {a: 1.0 for a in periodic_axes(region)}
        self.pec_upper = pec_upper(region)
        self.cpml = {}
        self.memory_states = []
        for key, slabs in cpml_slabs(region).items():
            entries = []
            for s in slabs:
                psi = torch.zeros(s.psi_shape(shape), device=self.device, dtype=self.dtype)
                self.memory_states.append(psi)
                coefficient = lambda a: torch.as_tensor(a, device=self.device)
                entries.append({'slab': s, 'psi': psi, 'b': coefficient(s.b), 'c': coefficient(s.c),
                                'inv_k': coefficient(s.inv_k), 'side': s.side, 'lo': s.lo, 'hi': s.hi,
                                'kappa': s.kappa, 'sigma': s.sigma, 'alpha': s.alpha})
            self.cpml[key] = entries

    def set_permittivity(self, eps):
        """inverse_permittivity = 1/eps, with the reciprocal taken in the field precision (solver.py:742-745)."""
        eps = np.asarray(eps)
        real = real_dtype(self.region)
        inverse = (1/(eps.astype(real) if eps.ndim == 4 else eps.astype(real)[..., None]))
        target = self.inverse_permittivity
        if inverse.shape[-1] != target.shape[-1]:
            if inverse.shape[-1] == 3 and target.shape[-1] == 1:
                self.inverse_permittivity = target = torch.empty((*self.region.shape, 3), device=self.device, dtype=self.dtype)
            else:
                inverse = np.broadcast_to(inverse, target.shape)
        target.copy_(torch.as_tensor(np.ascontiguousarray(inverse), device=self.device))

    def state(self):
        return [self.E, self.H, *self.memory_states]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def half_step_descriptor(grid, forward, sources=(), counter=None):
    """tfdtd_half_step for one family; `sources` are planning.SourceTerm of that family with device values."""
    d = N.HalfStep()
    d.dtype = N.F64 if grid.dtype == torch.float64 else N.F32
    d.family = N.FAMILY_H if forward else N.FAMILY_E
    d.nx, d.ny, d.nz = grid.Nx, grid.Ny, grid.Nz
    src, dst = (grid.E, grid.H) if forward else (grid.H, grid.E)
    inverse = grid.inverse_permeability if forward else grid.inverse_permittivity
    for t in (src, dst, inverse):
        if not t.is_cuda or not t.is_contiguous() or t.dtype != grid.dtype:
            raise ValueError('HIP kernels need contiguous GPU tensors of the field dtype.')
    d.src, d.dst, d.inverse = _ptr(src), _ptr(dst), _ptr(inverse)
    d.inverse_stride = 0 if inverse.numel() == 1 else inverse.shape[-1]
    d.courant = grid.courant_number
    d.periodic_mask = sum(1 << a for a in grid.wrap)
    d.pec_upper_mask = sum(1 << a for a in grid.pec_upper) if forward else 0
    for a, f in grid.pec_upper.items():
        d.pec_upper_factor[a] = f
    k = 0
    for axis, comp, _, _ in CURL_TERMS:
        for entry in grid.cpml.get((forward, axis, comp), []):
            seg = d.seg[k]
            seg.axis, seg.comp, seg.lo, seg.hi = axis, comp, entry['lo'], entry['hi']
            seg.psi, seg.b, seg.c, seg.inv_k = (_ptr(entry[n]) for n in ('psi', 'b', 'c', 'inv_k'))
            k += 1
    d.nseg = k
    if len(sources) > N.MAX_SOURCES:
        raise ValueError(f'at most {N.MAX_SOURCES} soft sources per field family are folded into the HIP step')
    for k, (comp, box, values) in enumerate(sources):
        s = d.srcs[k]
        s.comp = comp
        for a in range(3):
            s.lo[a], s.hi[a] = box[a].start, box[a].stop
        s.values = _ptr(values)
        if k == 0:
            d.source_samples = values.numel()
        elif values.numel() != d.source_samples:
            raise ValueError('every source waveform of one family must have the same length')
    d.nsrc = len(sources)
    d.counter = _ptr(counter) if counter is not None else None
    return d


class HipYee:
    """Drop-in for the reference's FusedYeeCUDA: E and H half steps on gfx950."""

    complex_fields = False

    def __init__(self, grid, **_ignored):
        if not getattr(grid, 'is_torch', False) or not grid.E.is_cuda:
            raise ValueError('The HIP kernel requires backend="cuda" and a ROCm GPU.')
        if grid.E.is_complex():
            raise ValueError('The HIP kernel supports real fields only.')
        if 3*math.prod(grid.E.shape[:3]) >= 2**62:
            raise ValueError('Field indexing exceeds the 64-bit range.')
        self.lib = N.library()
        self._grid = weakref.ref(grid)
        self.device = grid.E.device
        self.launches = {}
        self.tuned_blocks = {False: 256, True: 256}
        self.refresh()

    @property
    def grid(self):
        g = self._grid()
        if g is None:
            raise RuntimeError('The HIP grid is no longer alive.')
        return g

    def refresh(self):
        """Rebuild the descriptors (after a tensor of the grid was replaced)."""
        g = self.grid
        self.launches = {forward: half_step_descriptor(g, forward) for forward in (False, True)}

    @staticmethod
    def launch_count(grid, forward):
        return math.prod(grid.E.shape[:3])

    def update(self, forward):
        N.check(self.lib.tfdtd_yee_half_step(ctypes.byref(self.launches[forward]), N.stream_handle(self.device)),
                'tfdtd_yee_half_step')

    def update_E(self):
        self.update(False)

    def update_H(self):
        self.update(True)


def resolve_cuda_kernel(region, field='cuda_kernel'):
    """'auto' and 'fused' select the HIP kernels; 'torch' keeps the reference's unfused torch update."""
    choice = getattr(region, field)
    return 'torch' if choice == 'torch' else 'hip'


def configure_cuda_kernel(grid, choice):
    """Patch the grid's update methods, as torchfdtd/cuda_kernels.py:368-373 does."""
    if choice == 'torch':
        from .torch_path import configure_torch_kernel
        configure_torch_kernel(grid)
        return None
    implementation = HipYee(grid)
    grid.update_E = implementation.update_E
    grid.update_H = implementation.update_H
    grid.kernel = implementation
    return implementation
