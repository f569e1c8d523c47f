"""openmeeg_amd: MI355X-native symmetric-BEM EEG forward path.

Drop-in names of the reference's Python module ``openmeeg``
(wrapping/python/openmeeg/__init__.py:15-39) for the EEG forward path:
HeadMat, DipSourceMat, Head2EEGMat, GainEEG, GainEEGadjoint, Integrator,
Matrix, SymMatrix, Sensors, make_geometry, make_nested_geometry,
read_geometry; plus EEGForward, the EEG gain differentiable in the dipoles
and the conductivities, and BatchedGain, the gains of many heads in one
graph-captured pass (no reference counterparts).  All computation runs in the HIP library
(libopenmeeg_amd.so) through its C ABI (include/openmeeg_amd.h).
"""
from .api import (  # noqa: F401
    DipSource2InternalPotMat,
    DipSource2MEGMat,
    DipSourceMat,
    EITSourceMat,
    Factor,
    Forward,
    GainEEG,
    GainEEGadjoint,
    GainEEGMEGadjoint,
    GainEITInternalPot,
    GainInternalPot,
    GainMEG,
    GainMEGadjoint,
    Geometry,
    Head2EEGMat,
    Head2MEGMat,
    HeadMat,
    Integrator,
    Matrix,
    Mesh,
    Sensors,
    SparseMatrix,
    Surf2VolMat,
    SurfSource2MEGMat,
    SurfSourceMat,
    SymMatrix,
    factor,
    gain_eeg_meg_pipeline,
    gain_eeg_pipeline,
    gain_eeg_surf_pipeline,
    gain_internal_pot_pipeline,
    mesh_intersects,
    gain_to_numpy,
    get_log_level,
    make_geometry,
    make_nested_geometry,
    read_geometry,
    set_log_level,
    use_log_level,
)

__version__ = "0.1.0"
from ._lib import OverlappingSourceMesh  # noqa: F401,E402
from .differentiable import EEGForward  # noqa: F401,E402
from .batch import BatchedGain  # noqa: F401,E402
