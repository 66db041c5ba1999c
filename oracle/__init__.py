"""CPU oracle for the Yee/CPML time-stepping path -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy restatement of the reference's CPU path
(hyoseokp/TorchFDTD, whose ``backend='cpu'`` runs ``YeeGrid`` on NumPy arrays,
torchfdtd/solver.py:708-721).  It keeps the reference's operation order so that
FP32 results are reproducible bit for bit where the reference's own arithmetic is.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product package
``torchfdtd_amd`` never imports, links or runs anything here.

Pinning (what checks each part, besides restating the reference's source):

* ``scene`` geometry, CPML coefficients, source waveform, point index, and the
  ``yee`` update: the reference-written golden run
  ``tests/golden/tiny_2d_complete.npz`` (CPU FP32, 20x20 cells, 3-cell CPML,
  40 steps) -- final E/H, the 40-sample trace, the mesh nodes and epsilon.
* ``yee`` curl signs/indices in 2D and 3D: the closed-form discrete plane-wave
  eigenmode of the Yee scheme (omega = 2 asin(S |sin(q/2)|)), FP64, periodic.
* CPML: pulse absorption (residual of a plane pulse after the PML round trip).
* ``monitors.point_spectrum``: the same golden file's ``monitor_0_spectrum``.
* ``monitors`` plane DFT: a direct time-domain sum of the sampled trace
  (definition) and trilinear interpolation of linear fields (exact for affine
  fields).

Functions whose header says "parity unpinned" have none of the above.
"""
