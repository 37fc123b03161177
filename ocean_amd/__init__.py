"""MI355X-native OCEAN2 core-level BSE action and Krylov solvers.

``ocean_amd.core`` binds the HIP C-ABI library (``libocean_gpu.so``); it raises
on import of any GPU entry point when the library is not built.  ``synth`` and
``setup_host`` are host-side numpy and need no GPU.
"""
__all__ = ["core", "synth", "setup_host"]
