"""torchfdtd_amd: an MI355X-native (gfx950 HIP) engine for TorchFDTD's Yee/CPML time stepping.

The public names follow the reference package (``Project``, ``Region``,
``Simulation``, ...): a project that selects ``Region(cuda_kernel='hip')`` (or
``'auto'``) runs its fused Yee + CPML + source + DFT step on hand-written HIP
kernels through the C ABI of ``include/tfdtd.h``.
"""
from .models import (Boundaries, BoundaryFace, FieldMonitor, Material, MeshRefinement, Monitor, Project, Region,
                     RunControl, Source, SourceTimeSettings, SpectrumSettings, Structure)

__all__ = ['Boundaries', 'BoundaryFace', 'FieldMonitor', 'Material', 'MeshRefinement', 'Monitor', 'Project', 'Region',
           'RunControl', 'Result', 'Simulation', 'Source', 'SourceTimeSettings', 'SpectrumSettings', 'Structure']


def __getattr__(name):
    if name in ('Simulation', 'Result'):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
