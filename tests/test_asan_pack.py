"""The C ABI's host-side validation and table packing under AddressSanitizer (CPU).

tests/asan/pack_driver.cpp packs one-step, two-step, half-step and state-norm tables for grids of
several shapes into host buffers of exactly the size the library reports, and feeds the packers
malformed descriptors (slabs past an axis, overlapping faces, too many sources or segments, short
tables, missing counters) that must be refused.  The library sources are compiled with the host
half instrumented (the device half is built as usual and never launched), so any out-of-bounds
host access in the validation or packing code stops the driver.  Takes about 1.5 minutes: the
sources compile in parallel."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
SANITIZE = ['-Xarch_host', '-fsanitize=address', '-Xarch_host', '-fno-omit-frame-pointer']


@pytest.mark.skipif(not Path(HIPCC).exists(), reason='hipcc is not installed')
def test_packing_and_validation_are_clean_under_asan(tmp_path):
    from torchfdtd_amd import _native as N
    flags = ['--offload-arch=gfx950', '-O1', '-g', '-std=c++17', *SANITIZE, '-I', str(ROOT/'include')]

    def compile_one(source):
        out = tmp_path/(source.stem+'.o')
        subprocess.run([HIPCC, *flags, '-c', str(source), '-o', str(out)], check=True, capture_output=True)
        return out

    sources = [s for s in N.SOURCES if s.exists()]+[ROOT/'tests'/'asan'/'pack_driver.cpp']
    with ThreadPoolExecutor(max_workers=min(len(sources), max(1, (os.cpu_count() or 2)//2))) as pool:
        objects = list(pool.map(compile_one, sources))
    driver = tmp_path/'pack_driver'
    subprocess.run([HIPCC, '--offload-arch=gfx950', *SANITIZE[:2], *map(str, objects), '-o', str(driver)],
                   check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:halt_on_error=1')
    run = subprocess.run([str(driver)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout+run.stderr
    assert 'failures' in run.stdout and ' 0 failures' in run.stdout and 'AddressSanitizer' not in run.stderr
    shutil.rmtree(tmp_path, ignore_errors=True)
