"""Oracle: oracle/yee.py's half steps through the C restatement -- TEST INFRASTRUCTURE ONLY.

``NativeGrid`` is ``yee.Grid`` with ``update_E`` / ``update_H`` done by
``oracle/yee_native.c`` (same arrays, same rounding, OpenMP over x planes), so that the
oracle reaches configuration sizes in seconds.  The tests pin it to the
NumPy grid bit for bit in every boundary layout and on the reference-written golden runs
(tests/test_oracle_native.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from .yee import Grid

HERE = Path(__file__).resolve().parent
SOURCE = HERE/'yee_native.c'
SOURCES = (SOURCE, HERE/'yee_native_body.h')
LIBRARY = HERE/'liboracle_yee.so'
CFLAGS = ['-O3', '-fPIC', '-shared', '-fopenmp', '-std=c11',
          # every product and sum rounds on its own, as the NumPy ufuncs do
          '-ffp-contract=off', '-fno-fast-math']


class Segment(ctypes.Structure):
    _fields_ = [('axis', ctypes.c_int32), ('comp', ctypes.c_int32), ('lo', ctypes.c_int32), ('hi', ctypes.c_int32),
                ('psi', ctypes.c_void_p), ('b', ctypes.c_void_p), ('c', ctypes.c_void_p), ('inv_k', ctypes.c_void_p)]


def build(verbose=False):
    cc = os.environ.get('CC', 'gcc')
    tmp = LIBRARY.with_name(LIBRARY.name+'.tmp')
    cmd = [cc, *CFLAGS, str(SOURCE), '-o', str(tmp)]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


_lib = None


def library():
    global _lib
    if _lib is None:
        if not LIBRARY.exists() or LIBRARY.stat().st_mtime < max(s.stat().st_mtime for s in SOURCES):
            build()
        lib = ctypes.CDLL(str(LIBRARY))
        lib.oracle_half_step.restype = ctypes.c_int
        lib.oracle_half_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _lib = lib
    return _lib


def _ptr(a):
    assert a.flags.c_contiguous
    return a.ctypes.data


class NativeGrid(Grid):
    """yee.Grid whose half steps run in C (bitwise equal to the NumPy ones)."""

    def __init__(self, scene, inverse_permittivity=None):
        super().__init__(scene, inverse_permittivity)
        self._n = np.array(scene.shape, np.int32)
        self._segments = {}
        for forward in (False, True):
            segs = [(axis, comp, s) for (f, axis, comp), group in self.cpml.items() if f == forward for s in group]
            table = (Segment*max(1, len(segs)))()
            for k, (axis, comp, s) in enumerate(segs):
                for name in ('psi', 'b', 'c', 'inv_k'):
                    assert s[name].flags.c_contiguous and s[name].dtype == scene.real
                table[k] = Segment(axis, comp, s['lo'], s['hi'], _ptr(s['psi']), _ptr(s['b']), _ptr(s['c']),
                                   _ptr(s['inv_k']))
            self._segments[forward] = (table, len(segs))
        self._wrap = sum(1 << a for a in self.wrap)
        self._pec = sum(1 << a for a in self.pec_upper)
        self._pec_factor = np.array([self.pec_upper.get(a, 0.) for a in range(3)], np.float64)

    def _half(self, forward):
        src, dst = (self.E, self.H) if forward else (self.H, self.E)
        inverse = self.inverse_permeability if forward else self.inverse_permittivity
        real = self.scene.real
        for a in (src, dst, inverse):
            assert a.dtype == real and a.flags.c_contiguous
        assert src.shape == dst.shape == (*self.scene.shape, 3)
        stride = 0 if inverse.size == 1 else inverse.shape[-1]
        if stride:
            assert inverse.shape == (*self.scene.shape, stride) and stride in (1, 3)
        table, count = self._segments[forward]
        status = library().oracle_half_step(0 if real == np.float32 else 1, int(forward), _ptr(self._n), _ptr(src),
                                            _ptr(dst), _ptr(inverse), stride, float(self.courant), self._wrap,
                                            self._pec if forward else 0, _ptr(self._pec_factor), count, table)
        if status:
            raise RuntimeError('oracle_half_step rejected its arguments')

    def update_E(self):
        self._half(False)

    def update_H(self):
        self._half(True)
