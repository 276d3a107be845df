"""ddevol_amd: MI355X-native Pauli-propagation branch engine.

Native libraries (built in-tree by ``python -m ddevol_amd.build``):
  lib/libtpgp_hip.so     HIP/gfx950 kernels and the C ABI of include/tpgp_hip.h
  lib/libtpgp_engine.so  C17 host engine with the reference core's API (include/tpgp_engine.h)
"""
__all__ = ["native", "engine", "hamiltonians"]
