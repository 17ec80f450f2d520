"""pyto_amd: MI355X (gfx950) implementation of PyTO's voxel-hex structural FEA
path -- the matrix-free SIMP operator, the deflated PCG (FP64 and FP32 inside
FP64 refinement), the cone filter and the compliance / p-norm stress
sensitivities -- behind the reference's Python call surface
(hex_structural_fea, linear_solvers, deflation, topopt_obj_cons_sensitivities).

The compute runs in libpytohip.so (hand-written HIP, C ABI in
include/pyto_hip.h); build it with ``python -m pyto_amd.csrc.build``.
"""
__version__ = "0.1.0"
