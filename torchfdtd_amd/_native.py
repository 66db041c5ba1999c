"""Build and load libtfdtd.so, the gfx950 HIP library behind include/tfdtd.h.

The library is compiled in-tree with hipcc for gfx950 only.  Loading it requires
a ROCm GPU; there is no fallback: every engine entry point raises when the
library or the GPU is missing.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
LIBRARY = PACKAGE/'libtfdtd.so'
SOURCES = [PACKAGE/'csrc'/name for name in ('yee.hip', 'march.hip', 'norm.hip', 'adjoint.hip', 'conv.hip', 'ade.hip')]
HEADER = ROOT/'include'/'tfdtd.h'

HIPCC_FLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-shared', '-std=c++17',
               # Every product and sum rounds separately, as in the reference update.
               '-ffp-contract=off', '-fno-gpu-flush-denormals-to-zero',
               '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']


def source_digest():
    """SHA-256 over the library's sources (every file under csrc/ and the C header), in name order;
    None where the sources are not beside the package."""
    files = sorted(p for p in (PACKAGE/'csrc').glob('*') if p.suffix in ('.hip', '.h')) + [HEADER]
    if not all(f.exists() for f in files):
        return None
    digest = hashlib.sha256()
    for f in files:
        digest.update(f.name.encode()+b'\0'+f.read_bytes()+b'\0')
    return digest.hexdigest()


def build(verbose=False):
    """Compile every HIP source into LIBRARY (always recompiles), stamped with source_digest()."""
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    sources = [str(s) for s in SOURCES if s.exists()]
    tmp = LIBRARY.with_name(LIBRARY.name+'.tmp')
    cmd = [hipcc, *HIPCC_FLAGS, f'-DTFDTD_SOURCE_DIGEST="{source_digest()}"', '-I', str(ROOT/'include'), *sources,
           '-o', str(tmp)]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=str(PACKAGE))
    os.replace(tmp, LIBRARY)
    return LIBRARY


# ------------------------------------------------------------------ ctypes mirror of include/tfdtd.h

MAX_SEGMENTS = 12
MAX_SOURCES = 16
F32, F64 = 0, 1
FAMILY_E, FAMILY_H = 0, 1

vp = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64


class CpmlSegment(ctypes.Structure):
    _fields_ = [('axis', i32), ('comp', i32), ('lo', i32), ('hi', i32),
                ('psi', vp), ('b', vp), ('c', vp), ('inv_k', vp)]


class SourceBox(ctypes.Structure):
    _fields_ = [('comp', i32), ('lo', i32*3), ('hi', i32*3), ('values', vp)]


class HalfStep(ctypes.Structure):
    _fields_ = [('dtype', i32), ('family', i32), ('nx', i32), ('ny', i32), ('nz', i32),
                ('src', vp), ('dst', vp), ('inverse', vp), ('inverse_stride', i32),
                ('courant', ctypes.c_double), ('periodic_mask', i32), ('pec_upper_mask', i32),
                ('pec_upper_factor', ctypes.c_double*3),
                ('nseg', i32), ('seg', CpmlSegment*MAX_SEGMENTS),
                ('nsrc', i32), ('srcs', SourceBox*MAX_SOURCES),
                ('source_samples', i64), ('counter', vp)]


class Observe(ctypes.Structure):
    _fields_ = [('dtype', i32), ('count', i32), ('table', vp), ('traces', vp), ('steps', i64), ('counter', vp)]


class Plane(ctypes.Structure):
    _fields_ = [('dtype', i32), ('spectral', i32), ('points', i64), ('frequencies', i32),
                ('components', i32), ('electric', i32), ('corners', i32), ('time_downsample', i32),
                ('E', vp), ('H', vp), ('indices', vp*6), ('weights', vp*6),
                ('window_e', vp), ('window_h', vp), ('omega', vp), ('half', vp), ('value', vp),
                ('samples', i64), ('counter', vp), ('block', i32), ('stage', vp), ('flushed', vp)]


class MarchFace(ctypes.Structure):
    _fields_ = [('lo', i32), ('hi', i32), ('psi', i64*2), ('coef', i64)]


class MarchFamily(ctypes.Structure):
    _fields_ = [('psi_in', vp), ('psi_out', vp), ('coef', vp), ('face', (MarchFace*2)*3)]


class March(ctypes.Structure):
    _fields_ = [('dtype', i32), ('nx', i32), ('ny', i32), ('nz', i32),
                ('e_in', vp), ('h_in', vp), ('e_out', vp), ('h_out', vp),
                ('inverse_permittivity', vp), ('eps_stride', i32), ('inverse_permeability', vp),
                ('courant', ctypes.c_double), ('pec_upper_mask', i32), ('pec_upper_factor', ctypes.c_double*3),
                ('e', MarchFamily), ('h', MarchFamily),
                ('nsrc_e', i32), ('srcs_e', SourceBox*MAX_SOURCES),
                ('nsrc_h', i32), ('srcs_h', SourceBox*MAX_SOURCES),
                ('source_samples', i64), ('counter', vp), ('tile_y', i32), ('chunk_x', i32),
                ('material', vp), ('material_table', vp), ('materials', i32),
                ('snap_e', vp), ('snap_h', vp), ('snap_count', i32), ('snap_lo', (i32*3)*4), ('snap_hi', (i32*3)*4)]


class NormGrid(ctypes.Structure):
    _fields_ = [('dtype', i32), ('nx', i32), ('ny', i32), ('nz', i32), ('E', vp), ('H', vp),
                ('inverse_permittivity', vp), ('eps_stride', i32), ('widths', vp*3), ('psi', vp*2), ('psi_count', i64*2)]


ADE_MAX_POLES = 16


class Ade(ctypes.Structure):
    _fields_ = [('dtype', i32), ('poles', i32), ('samples', i64), ('per_cell', i32), ('index', vp), ('field', vp),
                ('old', vp), ('P', vp), ('Q', vp), ('a', ctypes.c_double*ADE_MAX_POLES),
                ('d', ctypes.c_double*ADE_MAX_POLES), ('k', ctypes.c_double*ADE_MAX_POLES), ('eps_inf', ctypes.c_double)]


class DecayRule(ctypes.Structure):
    _fields_ = [('eligible_time', ctypes.c_double), ('time_step', ctypes.c_double), ('decay_threshold', ctypes.c_double),
                ('growth_limit', ctypes.c_double), ('field_limit', ctypes.c_double), ('min_steps', i32),
                ('consecutive_checks', i32), ('divergence_check', i32), ('auto_shutoff', i32)]


class DecayState(ctypes.Structure):
    _fields_ = [('peak_energy', ctypes.c_double), ('baseline', ctypes.c_double), ('verdict_value', ctypes.c_double),
                ('verdict_step', i64), ('frozen', i32), ('below', i32), ('verdict', i32), ('checks', i32)]


RUN_GOING, RUN_DECAYED, RUN_NONFINITE, RUN_FIELD_LIMIT, RUN_GROWTH = range(5)


class CurlRecord(ctypes.Structure):
    _fields_ = [('dtype', i32), ('nx', i32), ('ny', i32), ('nz', i32), ('H', vp), ('periodic_mask', i32),
                ('box_lo', i32*3), ('box_hi', i32*3), ('record', vp), ('modulus', i64), ('counter', vp),
                ('e', MarchFamily)]


class Adjoint(ctypes.Structure):
    _fields_ = [('dtype', i32), ('nx', i32), ('ny', i32), ('nz', i32), ('e_bar', vp), ('h_bar', vp),
                ('inverse_permittivity', vp), ('eps_stride', i32), ('inverse_permeability', vp), ('courant', ctypes.c_double),
                ('periodic_mask', i32), ('pec_upper_mask', i32), ('pec_upper_factor', ctypes.c_double*3),
                ('e', MarchFamily), ('h', MarchFamily), ('nseed', i32), ('seed_targets', vp), ('seed_offsets', vp),
                ('seed_columns', vp), ('seeds', vp), ('seed_width', i32), ('seed_rows', i64), ('box_lo', i32*3), ('box_hi', i32*3),
                ('record', vp), ('record_modulus', i64), ('gradient', vp), ('gradient_stride', i32), ('counter', vp),
                ('record_cpml', i32)]


_lib = None

PROTOTYPES = {
    'tfdtd_yee_step_march': (i32, [ctypes.POINTER(March), vp]),
    'tfdtd_yee_step_march2': (i32, [ctypes.POINTER(March), vp]),
    'tfdtd_batch_march2_pack': (i32, [ctypes.POINTER(March), i32, vp, ctypes.c_size_t, ctypes.POINTER(i32)]),
    'tfdtd_batch_march2': (i32, [vp, vp, i32, i32, i32, i32, ctypes.POINTER(i32), vp]),
    'tfdtd_batch_table_bytes': (ctypes.c_size_t, [i32, i32]),
    'tfdtd_batch_pack': (i32, [ctypes.POINTER(HalfStep), i32, vp, ctypes.c_size_t, ctypes.POINTER(i32)]),
    'tfdtd_batch_half_step': (i32, [vp, vp, i32, i32, i32, i32, vp]),
    'tfdtd_batch_march_bytes': (ctypes.c_size_t, [i32, i32]),
    'tfdtd_batch_march_pack': (i32, [ctypes.POINTER(March), i32, vp, ctypes.c_size_t, ctypes.POINTER(i32)]),
    'tfdtd_batch_march': (i32, [vp, vp, i32, i32, i32, i32, ctypes.POINTER(i32), vp]),
    'tfdtd_batch_planes_bytes': (ctypes.c_size_t, [i32, i32, i32]),
    'tfdtd_batch_planes_pack': (i32, [ctypes.POINTER(Plane), i32, vp, ctypes.c_size_t, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
    'tfdtd_batch_dft_update': (i32, [vp, i32, i32, i32, i32, i32, vp]),
    'tfdtd_norm_tasks': (i64, [ctypes.POINTER(NormGrid), i32]),
    'tfdtd_norm_table_bytes': (ctypes.c_size_t, [ctypes.POINTER(NormGrid), i32]),
    'tfdtd_norm_pack': (i32, [ctypes.POINTER(NormGrid), i32, vp, ctypes.c_size_t]),
    'tfdtd_state_norm': (i32, [vp, i32, i64, i32, vp, vp, vp]),
    'tfdtd_state_norm_noalloc': (i32, [vp, i32, i64, i32, vp, vp]),
    'tfdtd_decay_update': (i32, [vp, vp, vp, i32, i64, vp, i32, vp]),
    'tfdtd_inject': (i32, [ctypes.POINTER(HalfStep), vp]),
    'tfdtd_ade_save': (i32, [ctypes.POINTER(Ade), vp]),
    'tfdtd_ade_correct': (i32, [ctypes.POINTER(Ade), vp]),
    'tfdtd_record_curl': (i32, [ctypes.POINTER(CurlRecord), vp]),
    'tfdtd_adjoint_step': (i32, [ctypes.POINTER(Adjoint), vp]),
    'tfdtd_conv_output': (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]),
    'tfdtd_conv_forward': (i32, [vp, vp, vp, vp, i32, vp, vp, vp]),
    'tfdtd_conv_dgrad': (i32, [vp, vp, vp, vp, vp]),
    'tfdtd_conv_wgrad': (i32, [vp, vp, vp, i32, vp, vp, vp]),
    'tfdtd_gelu_backward': (i32, [vp, vp, i64, i32, vp, vp, vp]),
    'tfdtd_abi_version': (i32, []),
    'tfdtd_source_digest': (ctypes.c_char_p, []),
    'tfdtd_last_error': (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    'tfdtd_yee_half_step': (i32, [ctypes.POINTER(HalfStep), vp]),
    'tfdtd_dft_update': (i32, [ctypes.POINTER(Plane), i32, vp]),
    'tfdtd_dft_flush': (i32, [ctypes.POINTER(Plane), i32, vp]),
    'tfdtd_observe_step': (i32, [ctypes.POINTER(Observe), vp]),
    'tfdtd_step_fused': (i32, [ctypes.POINTER(HalfStep), ctypes.POINTER(HalfStep), ctypes.POINTER(Plane), i32,
                               ctypes.POINTER(Observe), i32, vp]),
}


def load_library(path=None):
    """dlopen the library and bind the prototypes (no GPU needed to load)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    import torch  # noqa: F401  (torch's HIP runtime is the one the library binds to)
    target = Path(path) if path else LIBRARY
    if not target.exists():
        raise RuntimeError(f'{target} is missing: run torchfdtd_amd._native.build() (hipcc for gfx950) first')
    lib = ctypes.CDLL(str(target))
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    # A library compiled from other sources may lay its descriptors out differently from the
    # ctypes mirror above: refuse it rather than run it.
    expected = source_digest()
    built = lib.tfdtd_source_digest().decode()
    if path is None and expected is not None and built != expected:
        raise RuntimeError(f'{target} was built from other sources than those beside it (digest {built[:12]}, '
                           f'sources {expected[:12]}): rebuild with torchfdtd_amd._native.build()')
    if path is None:
        _lib = lib
    return lib


def library():
    """The loaded library, with a GPU present; raises otherwise (no CPU fallback)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('torchfdtd_amd runs on an AMD GPU (gfx950); no GPU is visible')
    return load_library()


def last_error():
    lib = load_library()
    buf = ctypes.create_string_buffer(1024)
    lib.tfdtd_last_error(buf, len(buf))
    return buf.value.decode(errors='replace')


def check(status, what):
    if status != 0:
        raise RuntimeError(f'{what} failed ({status}): {last_error()}')


def stream_handle(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
