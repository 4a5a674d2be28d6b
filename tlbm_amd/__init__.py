"""tlbm_amd: MI355X-native lattice-Boltzmann time step for tLBM.

The product is the HIP library built from tlbm_amd/csrc (libtlbm_gpu.so,
C-ABI in include/tlbm_gpu.h) and the C++ drop-in host (tlbm_amd/host).
This package carries the Python side used by tests and bench.py: the
ctypes binding of the C-ABI (tlbm_amd.gpu), input generation
(tlbm_amd.geometry) and the tLBM file formats (tlbm_amd.io_formats).
"""
