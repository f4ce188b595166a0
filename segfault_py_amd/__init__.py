"""segfault_py_amd: MI355X-native ERI and J/K (Fock-build) engine.

A drop-in for the two-electron-integral and J/K call of the reference's
Hartree-Fock code (src/hf/hf.py:26-30, :264, :300).  The compute path is
hand-written HIP for gfx950 in ``csrc/``, reached through the C-ABI library
``libsfeng.so`` via ctypes (``segfault_py_amd.engine``).
"""
__all__ = ["molecule", "workloads"]
