"""vfd_amd - MI355X-native DFSPH step (hand-written HIP for gfx950) behind VFD's simulation-component interface.

    from vfd_amd import DFSPHSimulation, DFSPHSimulationDescription, FluidObject, RigidBodyDescription

The compute path is the in-tree HIP library vfd_amd/libvfd_hip.so (C ABI: include/vfd.h). Importing the package
does not load it; the first DFSPHSimulation does, and raises if the library is missing.
"""
from ._abi import (BoundaryDesc, DebugInfo, Desc, Info, VfdError, load_library, make_boundary, make_desc,
                   EXPORTED, FIELDS, PHASES)
from .mesh import TriangleMesh
from .simulation import (DFSPHBatch, DFSPHParticleFrame, DFSPHSimulation, DFSPHSimulationDescription, FluidObject,
                         FluidObjectDescription, RigidBodyDescription, read_descriptions, sample_mesh_volume,
                         write_description)

__all__ = [
    "BoundaryDesc", "DebugInfo", "Desc", "Info", "VfdError", "load_library", "make_boundary", "make_desc",
    "EXPORTED", "FIELDS", "PHASES", "DFSPHBatch", "DFSPHParticleFrame", "DFSPHSimulation",
    "DFSPHSimulationDescription", "FluidObject", "FluidObjectDescription", "RigidBodyDescription", "TriangleMesh",
    "read_descriptions", "sample_mesh_volume", "write_description",
]
