"""Copy the reference's own test fixtures used as golden vectors into tests/golden/.

These files are data the reference's tests hold (a CPU FP32 run written by the
reference and its project JSON): tests/fixtures/npz/tiny_2d_complete.npz,
tests/fixtures/npz/tiny_2d_cancelled.npz and tests/fixtures/projects/tiny_result.json.
They are copied byte for byte; nothing of the reference is imported or run.
Run from the repository root: python tests/golden/copy_fixtures.py [reference root].
"""
import hashlib
import shutil
import sys
from pathlib import Path

FILES = {'tests/fixtures/npz/tiny_2d_complete.npz': 'tiny_2d_complete.npz',
         'tests/fixtures/npz/tiny_2d_cancelled.npz': 'tiny_2d_cancelled.npz',
         'tests/fixtures/projects/tiny_result.json': 'tiny_result.json'}


def main(reference='/root/reference'):
    here = Path(__file__).resolve().parent
    for source, name in FILES.items():
        src = Path(reference)/source
        shutil.copyfile(src, here/name)
        print(name, hashlib.sha256((here/name).read_bytes()).hexdigest())


if __name__ == '__main__':
    main(*sys.argv[1:])
