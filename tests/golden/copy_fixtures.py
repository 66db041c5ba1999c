"""Copy the reference's own test fixtures used as golden vectors into tests/golden/.

These files are data the reference holds (a CPU FP32 run written by the reference and its
project JSON): tests/fixtures/npz/tiny_2d_complete.npz, tests/fixtures/npz/tiny_2d_cancelled.npz
and tests/fixtures/projects/tiny_result.json, copied byte for byte; and the inverse-design
example's recorded run: examples/adjoint_inverse_design/splitter.json (the thresholded splitter
as 283 rectangles, copied byte for byte) and four arrays of
examples/adjoint_inverse_design/reference.npz (binary_epsilon, field_ez, field_x_um, field_y_um:
the FP64 permittivity on the Ez nodes and the 1200-step spectral Ez of 173 x-normal lines at
1.31 and 1.55 um) re-saved as splitter_fields.npz.  Nothing of the reference is imported or run;
the npz is read with allow_pickle=False.
Run from the repository root: python tests/golden/copy_fixtures.py [reference root].
"""
import hashlib
import shutil
import sys
from pathlib import Path

import numpy as np

FILES = {'tests/fixtures/npz/tiny_2d_complete.npz': 'tiny_2d_complete.npz',
         'tests/fixtures/npz/tiny_2d_cancelled.npz': 'tiny_2d_cancelled.npz',
         'tests/fixtures/projects/tiny_result.json': 'tiny_result.json',
         'examples/adjoint_inverse_design/splitter.json': 'splitter.json'}
SPLITTER_ARRAYS = ('binary_epsilon', 'field_ez', 'field_x_um', 'field_y_um')


def main(reference='/root/reference'):
    here = Path(__file__).resolve().parent
    for source, name in FILES.items():
        src = Path(reference)/source
        shutil.copyfile(src, here/name)
        print(name, hashlib.sha256((here/name).read_bytes()).hexdigest())
    with np.load(Path(reference)/'examples/adjoint_inverse_design/reference.npz', allow_pickle=False) as z:
        np.savez_compressed(here/'splitter_fields.npz', **{k: z[k] for k in SPLITTER_ARRAYS})
    print('splitter_fields.npz', ', '.join(SPLITTER_ARRAYS))


if __name__ == '__main__':
    main(*sys.argv[1:])
