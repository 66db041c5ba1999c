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
        # All psi arrays of a family share one buffer (each slab a contiguous view), and the
        # two curl terms along an axis share one coefficient block per face: b, c, inv_k.
        table = cpml_slabs(region)
        self.psi_buffers, self.coef_buffers = {}, {}
        for forward in (False, True):
            keys = [(forward, axis, comp) for axis, comp, _, _ in CURL_TERMS if (forward, axis, comp) in table]
            psi_total = sum(math.prod(s.psi_shape(shape)) for k in keys for s in table[k])
            blocks, coef = {}, []
            for key in keys:
                for s in table[key]:
                    face = (s.axis, s.side)
                    if face in blocks:
                        first = blocks[face][1]
                        if (first.lo, first.hi) != (s.lo, s.hi) or not all(
                                np.array_equal(getattr(first, n), getattr(s, n)) for n in ('b', 'c', 'inv_k')):
                            raise ValueError('CPML slabs of one axis and face must share rows and coefficients')
                    else:
                        blocks[face] = (sum(len(c) for c in coef), s)
                        coef.extend([s.b, s.c, s.inv_k])
            psi_buffer = torch.zeros(max(psi_total, 1), device=self.device, dtype=self.dtype)
            coef_host = np.concatenate(coef) if coef else np.zeros(1, real_dtype(region))
            coef_buffer = torch.as_tensor(coef_host, device=self.device)
            self.psi_buffers[forward], self.coef_buffers[forward] = psi_buffer, coef_buffer
            offset = 0
            for key in keys:
                entries = []
                for s in table[key]:
                    size = math.prod(s.psi_shape(shape))
                    psi = psi_buffer[offset:offset+size].view(s.psi_shape(shape))
                    self.memory_states.append(psi)
                    w = s.hi-s.lo
                    c0 = blocks[(s.axis, s.side)][0]
                    entries.append({'slab': s, 'psi': psi, 'psi_offset': offset, 'coef_offset': c0,
                                    'b': coef_buffer[c0:c0+w], 'c': coef_buffer[c0+w:c0+2*w], 'inv_k': coef_buffer[c0+2*w:c0+3*w],
                                    'side': s.side, 'lo': s.lo, 'hi': s.hi, 'kappa': s.kappa, 'sigma': s.sigma, 'alpha': s.alpha})
                    offset += size
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


MARCH_SOURCES = 8


def march_supported(grid, sources=None):
    """Whether the one-pass kernel runs this grid: 3D, no periodic axis, 32-bit offsets, <= 8 sources per family."""
    r = grid.region
    if r.dimension != '3d' or min(r.shape) < 2 or grid.wrap:
        return False
    if 3*math.prod(r.shape) >= 2**32 or max(b.numel() for b in grid.psi_buffers.values()) >= 2**32:
        return False
    if sources is not None and any(len(v) > MARCH_SOURCES for v in sources.values()):
        return False
    return True


def march_family(grid, forward, psi_in, psi_out):
    """tfdtd_march_family: offsets of every slab in the family's psi and coefficient buffers."""
    f = N.MarchFamily()
    f.psi_in, f.psi_out, f.coef = _ptr(psi_in), _ptr(psi_out), _ptr(grid.coef_buffers[forward])
    for axis in range(3):
        terms = [(forward, a, comp) for a, comp, _, _ in CURL_TERMS if a == axis]
        for side in (0, 1):
            entries = [next((e for e in grid.cpml.get(t, []) if e['side'] == side), None) for t in terms]
            if entries[0] is None:
                continue
            face = f.face[axis][side]
            face.lo, face.hi = entries[0]['lo'], entries[0]['hi']
            face.psi[0], face.psi[1] = entries[0]['psi_offset'], entries[1]['psi_offset']
            face.coef = entries[0]['coef_offset']
    return f


class MaterialIndex:
    """The grid's per-cell 1/eps as a one-byte material index and a table of at most 256 values, for
    the one-pass step (which then reads a quarter of the bytes).  Only grids with one value per
    cell and at most 256 distinct values are indexed; the table holds the exact values, so the
    step's arithmetic is unchanged.  `refresh` rebuilds index and table in place (their device
    addresses never change) when inverse_permittivity was modified or replaced since the last
    build, as torch's version counter of the tensor tells."""

    MAX = 256

    def __init__(self, grid):
        self.grid = grid
        cells = grid.Nx*grid.Ny*grid.Nz
        self.index = torch.zeros(cells, dtype=torch.uint8, device=grid.device)
        self.table = torch.zeros(self.MAX, dtype=grid.dtype, device=grid.device)
        self.count = 0
        self.seen = None
        self.refresh()

    def _key(self):
        inv = self.grid.inverse_permittivity
        return id(inv), inv._version

    def stale(self):
        return self._key() != self.seen

    def refresh(self):
        """Rebuild from inverse_permittivity if it changed; returns True when it did."""
        if not self.stale():
            return False
        inv = self.grid.inverse_permittivity
        self.count = 0
        if inv.shape[-1] == 1:
            values, inverse = torch.unique(inv.reshape(-1), return_inverse=True)
            if values.numel() <= self.MAX:
                self.table[:values.numel()] = values
                self.index.copy_(inverse.to(torch.uint8))
                self.count = values.numel()
        self.seen = self._key()
        return True

    def describe(self, d):
        """Point tfdtd_march `d` at the index (or at the plain 1/eps when the grid is not indexed)."""
        d.material = _ptr(self.index) if self.count else None
        d.material_table = _ptr(self.table) if self.count else None
        d.materials = self.count


def march_descriptor(grid, e_in, h_in, e_out, h_out, psi_e_in, psi_e_out, sources, counter, tile_y=8, chunk_x=8,
                     materials=None, psi_h=None, snapshot=None):
    """tfdtd_march: one whole step reading (e_in, h_in, psi_e_in) and writing (e_out, h_out, psi_e_out);
    `materials` (a MaterialIndex) lets the step read 1/eps through the one-byte index.  The H-family
    CPML memory updates in place unless `psi_h` = (input, output) is given (the two-step march).
    `snapshot` = (E buffer, H buffer, [(lo, hi), ...]) with lo, hi cell triples: the two-step pass also
    writes step n+1 in those boxes, one after another in the buffers."""
    d = N.March()
    d.dtype = N.F64 if grid.dtype == torch.float64 else N.F32
    d.nx, d.ny, d.nz = grid.Nx, grid.Ny, grid.Nz
    d.e_in, d.h_in, d.e_out, d.h_out = (_ptr(t) for t in (e_in, h_in, e_out, h_out))
    d.inverse_permittivity = _ptr(grid.inverse_permittivity)
    d.eps_stride = grid.inverse_permittivity.shape[-1]
    d.inverse_permeability = _ptr(grid.inverse_permeability)
    d.courant = grid.courant_number
    d.pec_upper_mask = sum(1 << a for a in grid.pec_upper)
    for a, f in grid.pec_upper.items():
        d.pec_upper_factor[a] = f
    d.e = march_family(grid, False, psi_e_in, psi_e_out)
    psi_h_in, psi_h_out = psi_h if psi_h is not None else (grid.psi_buffers[True],)*2
    d.h = march_family(grid, True, psi_h_in, psi_h_out)
    samples = 0
    for family, count_name, table_name in (('E', 'nsrc_e', 'srcs_e'), ('H', 'nsrc_h', 'srcs_h')):
        entries = sources.get(family, ())
        if len(entries) > MARCH_SOURCES:
            raise ValueError(f'the one-pass kernel folds at most {MARCH_SOURCES} sources per family')
        table = getattr(d, table_name)
        for k, (comp, box, values) in enumerate(entries):
            s = table[k]
            s.comp = comp
            for a in range(3):
                s.lo[a], s.hi[a] = box[a].start, box[a].stop
            s.values = _ptr(values)
            if samples and values.numel() != samples:
                raise ValueError('every source waveform must have the same length')
            samples = values.numel()
        setattr(d, count_name, len(entries))
    d.source_samples = samples
    d.counter = _ptr(counter)
    d.tile_y, d.chunk_x = tile_y, chunk_x
    if materials is not None:
        materials.describe(d)
    if snapshot is not None:
        snap_e, snap_h, ranges = snapshot
        d.snap_e, d.snap_h, d.snap_count = _ptr(snap_e), _ptr(snap_h), len(ranges)
        for k, (lo, hi) in enumerate(ranges):
            for a in range(3):
                d.snap_lo[k][a], d.snap_hi[k][a] = lo[a], hi[a]
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
