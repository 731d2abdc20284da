"""MI355X (gfx950) engine for paraCLAWs' 2D wave-propagation update.

The compute path is ``libpcw_hip.so`` (hand-written HIP kernels behind the
C ABI of ``include/pcw.h``); ``abi`` binds it for Python callers and the
Fortran host under ``paraclaws_amd/fortran`` binds it with iso_c_binding.
"""
__all__ = ["abi", "control", "timectl"]
