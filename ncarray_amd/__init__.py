"""ncarray_amd: MI355X (gfx950) device backend for ncarray's hot path.

Fused elementwise expressions (hipRTC), stencils, axis and full reductions and binning,
behind a C ABI (include/nca_hip.h, libnca_hip.so).
"""
__version__ = "0.1.0"
