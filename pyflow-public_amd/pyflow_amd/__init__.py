"""pyflow_amd: PyFlow's single-box time-step path on MI355X (gfx950).

Host classes keep the reference's names, signatures and in-place semantics
(metric_uniform, rhs_NavierStokes, RK4, solver_bicgstab, Domain, run ...);
every numerical operation runs in hand-written HIP kernels of
libpyflow_hip.so, reached through its C-ABI (include/pyflow_hip.h).
"""

__version__ = '0.1.0'
