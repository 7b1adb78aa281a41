"""MI355X-native symmetry-blocked DMRG contraction path (H*psi and boundary growth) behind a C ABI.

Modules:
  abi        ctypes mirror of include/qcmaquis_amd.h and the library loader
  model      host-side description of indices, MPO tensors, boundaries, MPS layouts
  engine     Context / DeviceBoundary / DeviceMPS / HpsiPlan and the growth steps
  synthetic  random charge-consistent chains (tests, benchmarks)
  build      in-tree hipcc build of libqcmaquis_amd.so for gfx950
"""
__version__ = "0.1.0"
