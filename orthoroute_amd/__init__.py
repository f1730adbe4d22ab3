"""orthoroute_amd -- MI355X-native PathFinder / SSSP hot path for OrthoRoute.

Device kernels live in liborx_hip.so (csrc/, C ABI in include/orthoroute_hip.h);
this package holds the host side: lattice and workload generators, the net
order and disjoint-ROI wave planner, the device-resident PathFinder driver and
the drop-in solver seam.
"""
__version__ = "0.1.0"
