"""fiber3d_amd -- the fiber3d GMRES-matvec hot path on one AMD MI355X (gfx950).

Host code in Python on PyTorch-ROCm (device memory and streams); every
numerical kernel is hand-written HIP in csrc/, reached through the C ABI of
libf3d.so (include/f3d.h).  There is no CPU fallback: without the built
library or a GPU the entry points raise.
"""
__version__ = "0.1.0"
