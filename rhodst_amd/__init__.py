"""rhodst_amd -- MI355X-native block-coupled implicit step of rhoDST.

The path (face loop, Roe Jacobian assembly into block-CSR, block-ILU(0)
GMRES/BiCGStab solve, primitive/BC update, local time stepping) runs as HIP
kernels for gfx950 behind the C-ABI in include/rhodst.h; this package holds
the ctypes binding (_capi), the C descriptors (_abi), mesh generation and
polyMesh I/O (meshgen), case descriptions (cases) and the torch harness
(adjoint).
"""
from . import meshgen, cases  # noqa: F401
from ._capi import DeviceCase, DeviceMesh, RdstError, init, lib  # noqa: F401
