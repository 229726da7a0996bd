"""FP64 CPU oracle of the fiber3d GMRES-matvec hot path.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain NumPy (and, for the
CPU baseline, C), the reference algorithm of bioflume/fiber3d for the path that
fiber3d_amd implements on the GPU.  It is imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
checker -- never as the thing measured or shipped.  fiber3d_amd never imports
it.

Each function cites the reference file:line it follows (paths relative to the
reference's python/ directory).  How each function is pinned (closed forms,
invariants, golden vectors written from the reference) is listed in the
module docstrings and in tests/test_oracle_*.py.
"""
