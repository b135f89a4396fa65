"""MI355X-native SFCI ground-state engine.

- ``sfci_amd.gpu``        ctypes binding of libsfci_hip.so (H.v and the sparse-vector
                          operations of the Lanczos driver on the GPU, C-ABI in
                          include/sfci_gpu.h)
- ``sfci_amd.integrals``  synthetic integrals, FCIDUMP writer, FCIDUMP -> _IN converter
"""
__all__ = ["gpu", "integrals"]
