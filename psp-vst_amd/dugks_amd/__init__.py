"""MI355X-native DUGKS time step (dugksFoam fvDVM::evolution) -- host-side Python layer.

Submodules: ``quadrature`` (velocity sets), ``mesh`` (polyMesh + geometry), ``cases``
(case descriptions and flattened C-ABI arrays), ``timecontrol`` (adaptive dt),
``hip`` (ctypes binding of the HIP C-ABI library ``libdugks_hip.so``).
"""
__all__ = ["quadrature", "mesh", "cases", "timecontrol"]
