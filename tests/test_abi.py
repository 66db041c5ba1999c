"""The C ABI (CPU only): the library loads, exports every function include/tfdtd.h
declares, and the ctypes mirror has the C layout of every descriptor."""
import ctypes
import re
import subprocess

import pytest

from tests.conftest import ROOT
from torchfdtd_amd import _native as N

HEADER = ROOT/'include'/'tfdtd.h'


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r'^\w[\w\s\*]*?\b(tfdtd_\w+)\s*\(', text, flags=re.M)))


def test_library_exports_every_declared_function():
    if not N.LIBRARY.exists():
        pytest.fail('libtfdtd.so is not built: run __graft_entry__.build()')
    lib = ctypes.CDLL(str(N.LIBRARY))
    names = declared_functions()
    assert 'tfdtd_yee_half_step' in names and 'tfdtd_last_error' in names
    for name in names:
        assert hasattr(lib, name), name
    assert lib.tfdtd_abi_version() == 1


def test_library_is_built_from_the_sources_beside_it():
    """The build stamps the sources' digest into the library; the loader refuses a stale one."""
    lib = N.load_library()
    assert lib.tfdtd_source_digest().decode() == N.source_digest()
    stale = ctypes.CDLL(str(N.LIBRARY))
    stale.tfdtd_source_digest.restype = ctypes.c_char_p
    assert len(stale.tfdtd_source_digest()) == 64


def test_errors_are_reported_without_a_gpu():
    lib = N.load_library()
    d = N.HalfStep()
    d.dtype, d.family, d.nx, d.ny, d.nz = 7, 0, 4, 4, 4
    d.src = d.dst = d.inverse = 16
    assert lib.tfdtd_yee_half_step(ctypes.byref(d), None) == 1
    assert 'dtype' in N.last_error()
    d.dtype, d.nseg = 0, 1
    d.seg[0].axis, d.seg[0].comp, d.seg[0].lo, d.seg[0].hi = 0, 0, 0, 2   # no term differentiates Ex along x
    assert lib.tfdtd_yee_half_step(ctypes.byref(d), None) == 1
    assert 'no curl term' in N.last_error()


LAYOUT_C = r'''
#include <stdio.h>
#include <stddef.h>
#include "tfdtd.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
int main(void) {
  printf("tfdtd_half_step %zu\n", sizeof(tfdtd_half_step));
  printf("tfdtd_plane %zu\n", sizeof(tfdtd_plane));
  printf("tfdtd_observe %zu\n", sizeof(tfdtd_observe));
  printf("tfdtd_cpml_segment %zu\n", sizeof(tfdtd_cpml_segment));
  printf("tfdtd_source %zu\n", sizeof(tfdtd_source));
  F(tfdtd_half_step, courant); F(tfdtd_half_step, seg); F(tfdtd_half_step, srcs); F(tfdtd_half_step, source_samples); F(tfdtd_half_step, counter);
  F(tfdtd_plane, points); F(tfdtd_plane, indices); F(tfdtd_plane, samples); F(tfdtd_plane, counter);
  F(tfdtd_plane, block); F(tfdtd_plane, stage); F(tfdtd_plane, flushed);
  F(tfdtd_observe, steps); F(tfdtd_observe, counter);
  printf("tfdtd_march %zu\n", sizeof(tfdtd_march));
  printf("tfdtd_march_face %zu\n", sizeof(tfdtd_march_face));
  printf("tfdtd_march_family %zu\n", sizeof(tfdtd_march_family));
  F(tfdtd_march, courant); F(tfdtd_march, e); F(tfdtd_march, h); F(tfdtd_march, nsrc_e); F(tfdtd_march, srcs_h);
  F(tfdtd_march, source_samples); F(tfdtd_march, counter); F(tfdtd_march, chunk_x);
  F(tfdtd_march, material); F(tfdtd_march, material_table); F(tfdtd_march, materials);
  F(tfdtd_march_family, face); F(tfdtd_march_face, psi); F(tfdtd_march_face, coef);
  printf("tfdtd_norm_grid %zu\n", sizeof(tfdtd_norm_grid));
  F(tfdtd_norm_grid, eps_stride); F(tfdtd_norm_grid, widths); F(tfdtd_norm_grid, psi); F(tfdtd_norm_grid, psi_count);
  printf("tfdtd_curl_record %zu\n", sizeof(tfdtd_curl_record));
  F(tfdtd_curl_record, box_lo); F(tfdtd_curl_record, record); F(tfdtd_curl_record, modulus); F(tfdtd_curl_record, counter);
  F(tfdtd_curl_record, e); F(tfdtd_adjoint, record_cpml);
  printf("tfdtd_decay_rule %zu\n", sizeof(tfdtd_decay_rule));
  printf("tfdtd_decay_state %zu\n", sizeof(tfdtd_decay_state));
  F(tfdtd_decay_rule, field_limit); F(tfdtd_decay_rule, auto_shutoff);
  F(tfdtd_decay_state, verdict_step); F(tfdtd_decay_state, checks);
  printf("tfdtd_ade %zu\n", sizeof(tfdtd_ade));
  F(tfdtd_ade, samples); F(tfdtd_ade, index); F(tfdtd_ade, a); F(tfdtd_ade, k); F(tfdtd_ade, eps_inf);
  printf("tfdtd_adjoint %zu\n", sizeof(tfdtd_adjoint));
  F(tfdtd_adjoint, courant); F(tfdtd_adjoint, e); F(tfdtd_adjoint, h); F(tfdtd_adjoint, nseed); F(tfdtd_adjoint, seeds); F(tfdtd_adjoint, seed_rows);
  printf("tfdtd_conv %zu\n", sizeof(tfdtd_conv));
  F(tfdtd_conv, co); F(tfdtd_conv, kp);
  F(tfdtd_adjoint, box_lo); F(tfdtd_adjoint, record_modulus); F(tfdtd_adjoint, gradient_stride); F(tfdtd_adjoint, counter);
  return 0;
}
'''


def test_ctypes_mirror_matches_the_c_layout(tmp_path):
    source = tmp_path/'layout.c'
    source.write_text(LAYOUT_C)
    exe = tmp_path/'layout'
    subprocess.run(['gcc', '-std=c11', '-I', str(HEADER.parent), str(source), '-o', str(exe)], check=True)
    lines = dict(line.rsplit(' ', 1) for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split('\n') if line)
    mirror = {'tfdtd_half_step': N.HalfStep, 'tfdtd_plane': N.Plane, 'tfdtd_observe': N.Observe,
              'tfdtd_cpml_segment': N.CpmlSegment, 'tfdtd_source': N.SourceBox, 'tfdtd_march': N.March,
              'tfdtd_march_face': N.MarchFace, 'tfdtd_march_family': N.MarchFamily, 'tfdtd_norm_grid': N.NormGrid,
              'tfdtd_curl_record': N.CurlRecord, 'tfdtd_adjoint': N.Adjoint,
              'tfdtd_decay_rule': N.DecayRule, 'tfdtd_decay_state': N.DecayState, 'tfdtd_ade': N.Ade,
              'tfdtd_conv': __import__('torchfdtd_amd.surrogate', fromlist=['ConvShape']).ConvShape}
    for key, value in lines.items():
        if '.' in key:
            struct, member = key.split('.')
            assert getattr(mirror[struct], member).offset == int(value), key
        else:
            assert ctypes.sizeof(mirror[key]) == int(value), key


def test_batch_and_norm_packing_validate_on_the_host():
    lib = N.load_library()
    steps = (N.HalfStep*2)()
    for k, family in enumerate((0, 1)):
        d = steps[k]
        d.dtype, d.family, d.nx, d.ny, d.nz = 0, family, 4, 4, 4
        d.src = d.dst = d.inverse = 16
    nbytes = lib.tfdtd_batch_table_bytes(2, 0)
    host = (ctypes.c_ubyte*nbytes)()
    prefix = (ctypes.c_int32*3)()
    assert lib.tfdtd_batch_pack(steps, 2, host, nbytes, prefix) == 1
    assert 'differs in family' in N.last_error()
    steps[1].family = 0
    assert lib.tfdtd_batch_pack(steps, 2, host, nbytes-1, prefix) == 1
    assert lib.tfdtd_batch_pack(steps, 2, host, nbytes, prefix) == 0
    assert list(prefix) == [0, 1, 2]    # 64 cells: one 256-thread block per grid
    grids = (N.NormGrid*1)()
    grids[0].dtype, grids[0].nx, grids[0].ny, grids[0].nz, grids[0].eps_stride = 0, 0, 4, 4, 1
    assert lib.tfdtd_norm_pack(grids, 1, host, nbytes) == 1
    grids[0].nx, grids[0].E, grids[0].H, grids[0].inverse_permittivity = 4, 16, 16, 16
    grids[0].widths[0] = 16
    nbytes = lib.tfdtd_norm_table_bytes(grids, 1)
    host = (ctypes.c_ubyte*nbytes)()
    assert lib.tfdtd_norm_pack(grids, 1, host, nbytes) == 1
    assert 'width' in N.last_error()
    grids[0].widths[0] = None
    assert lib.tfdtd_norm_pack(grids, 1, host, nbytes) == 0
    assert lib.tfdtd_norm_tasks(grids, 1) == 2    # E and H, 192 values each
