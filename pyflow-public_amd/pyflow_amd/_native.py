"""ctypes binding of libpyflow_hip.so (declared in include/pyflow_hip.h).

The library is required: importing the GPU classes without it, or without a
ROCm device, raises instead of silently computing something else.
"""

import ctypes
import os

import torch

from . import _build

PF_F64 = 0
PF_F32 = 1
PF_POISSON_BICGSTAB = 0
PF_POISSON_JACOBI = 1
PF_POISSON_BICGSTAB_MPIR = 2
SOLVE_STATUS = {0: 'converged', 1: 'maxit', 2: 'breakdown <v,r_hat>=0', 3: 'breakdown <t,t>=0',
                4: 'breakdown rho2*omega=0', 5: 'stalled'}

_lib = None

# (name, restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
SIGNATURES = [
    ('pf_abi_version', _I, []),
    ('pf_last_error', ctypes.c_char_p, []),
    ('pf_box_create', _I, [ctypes.POINTER(_P), _I, _I, _I, _D, _D, _D, _I]),
    ('pf_box_destroy', _I, [_P]),
    ('pf_box_device_bytes', ctypes.c_int64, [_P]),
    ('pf_box_set_spacing', _I, [_P, _D, _D, _D]),
    ('pf_box_set_persistent', _I, [_P, _I]),
    ('pf_solve_last_persistent', _I, [_P]),
    ('pf_fill_halo', _I, [_P, _P, _P]),
    ('pf_rk4_step', _I, [_P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _PP, _P]),
    ('pf_euler_step', _I, [_P, _P, _P, _P, _D, _D, _D, _PP, _P]),
    ('pf_rhs_ns', _I, [_P, _P, _P, _P, _D, _D, _PP, _P, _P, _P, _P]),
    ('pf_div_vel', _I, [_P, _P, _P, _P, _P, _P]),
    ('pf_bicgstab_solve', _I, [_P, _P, _P, _P, _D, _D, _D, _I, _PD, _PI, _P]),
    ('pf_jacobi_solve', _I, [_P, _P, _P, _P, _D, _D, _I, _PD, _P]),
    ('pf_vel_corr', _I, [_P, _P, _P, _P, _P, _P, _P, _I, _D, _D, _D, _I, _PD, _PI, _P]),
    ('pf_apply_laplacian', _I, [_P, _P, _P, _P]),
    ('pf_state_op', _I, [_P, _I, _D, _P, _P, _I, _P, _I, _I, _P]),
    ('pf_stats', _I, [_P, _P, _P, _P, _P, _PD, _P]),
    ('pf_sfs_gradient', _I, [_P, _P, _P, _P, _PP, _P]),
    ('pf_metric_pair', _I, [_P, _I, _I, _I, _P, _P, _D, _P]),
    ('pf_metric_edge4', _I, [_P, _I, _P, _P, _P]),
    ('pf_div_visc', _I, [_P, _P, _P, _P, _P, _P]),
    ('pf_div_vel_over', _I, [_P, _P, _P, _P, _P, _P]),
    ('pf_vel_conv', _I, [_P, _I, _I, _P, _P, _P, _D, _P]),
    ('pf_sfs_mlp', _I, [_P, _P, _P, _P, _P, _PP, _P]),
    ('pf_sfs_mlp_param_count', _I, []),
    ('pf_sfs_mlp_f32', _I, [_P, _P, _P, _P, _P, _PP, _P]),
    ('pf_rhs_adj', _I, [_P, _P, _P, _P, _P, _P, _P, _D, _D, _PP, _P, _P, _P, _P]),
    ('pf_adj_rk4_step', _I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _I, _P, _P, _PP,
                             _PD, _P]),
    ('pf_adj_euler_step', _I, [_P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _I, _P, _P, _PP, _PD, _P]),
    ('pf_rhs_adj_visc', _I, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _PP, _P, _P, _P, _P]),
    ('pf_adj_rk4_step_visc', _I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _D, _I, _P, _P,
                                  _PP, _PD, _P]),
    ('pf_adj_euler_step_visc', _I, [_P, _P, _P, _P, _P, _P, _P, _D, _P, _D, _I, _P, _P, _PP, _PD,
                                    _P]),
    ('pf_adj_work_bytes', ctypes.c_size_t, [_P]),
    ('pf_adj_closure_term', _I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _PP, _PD, _P]),
    ('pf_adj_conv', _I, [_P, _I, _I, _I, _P, _P, _P, _D, _P]),
    ('pf_rhs_vjp', _I, [_P, _P, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P]),
    ('pf_sfs_vjp', _I, [_P, _P, _P, _P, _PP, _I, _P]),
    ('pf_fold_halo', _I, [_P, _P, _P]),
    ('pf_closure_vjp_waves', _I, [_P]),
    ('pf_closure_vjp', _I, [_P, _I, _P, _P, _P, _P, _PP, _P, _P, _P, _P, _PD, _PD, _P]),
    ('pf_batch_create', _I, [ctypes.POINTER(_P), _I, _I]),
    ('pf_batch_destroy', _I, [_P]),
    ('pf_batch_set', _I, [_P, _I, _PP, _PP, _PP, _PD]),
    ('pf_batch_advance', _I, [_P, _PI, _PI, _PD, _PI, _P]),
    ('pf_batch_chunk_iterations', _I, []),
    ('pf_batch_set_group_bytes', _I, [_P, ctypes.c_int64]),
    ('pf_batch_set_persistent', _I, [_P, _I]),
    ('pf_profile_enable', _I, [_P, _I]),
    ('pf_profile_read', _I, [_P, _I, _PD, ctypes.POINTER(ctypes.c_int64)]),
    ('pf_profile_read_persistent', _I, [_P, _PD, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64)]),
    ('pf_residual_history_enable', _I, [_P, _I]),
    ('pf_rk4_step_visc', _I, [_P, _P, _P, _P, _P, _P, _P, _D, _P, _D, _P, _P]),
    ('pf_euler_step_visc', _I, [_P, _P, _P, _P, _D, _P, _D, _P, _P]),
    ('pf_rhs_ns_visc', _I, [_P, _P, _P, _P, _P, _D, _P, _P, _P, _P, _P]),
    ('pf_sfs_smagorinsky', _I, [_P, _P, _P, _P, _D, _D, _D, _D, _P, _P]),
    ('pf_bicgstab_mpir_solve', _I, [_P, _P, _P, _P, _D, _D, _D, _I, _PD, ctypes.POINTER(_I), _P]),
    ('pf_mpir_configure', _I, [_P, _D, _I, _I]),
    ('pf_mpir_solution', _I, [_P, _P, _P, _P]),
    ('pf_residual_history', _I, [_P, _PD, _I, ctypes.POINTER(ctypes.c_int), _P]),
    ('pf_source_inputs', _I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    ('pf_rhs_ns_src', _I, [_P, _P, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P, _P]),
    ('pf_euler_step_src', _I, [_P, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P]),
    ('pf_rk4_source_inputs', _I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    ('pf_rk4_stage_src', _I, [_P, _I, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P]),
    ('pf_sfs_ml_source', _I, [_P, _P, ctypes.POINTER(ctypes.c_float), _P, _P, _P, _P, _P, _P, _P]),
    ('pf_sfs_ml_source_params', ctypes.c_int64, []),
    ('pf_sfs_ml_source_reserve', _I, [_P]),
    ('pf_sfs_ml_source_train_reserve', _I, [_P, ctypes.POINTER(ctypes.c_float)]),
    ('pf_sfs_ml_source_cache_budget', _I, [_P, ctypes.c_int64]),
    ('pf_sfs_ml_source_vjp', _I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _PD, _P]),
    ('pf_rk4_step_noalloc', _I, [_I, _I, _I, _D, _D, _D, _I, _P, _P, _P, _P, _P, _P, _D, _D, _D,
                                 _PP, _P]),
    ('pf_rk4_step_noalloc_max_bytes', ctypes.c_int64, []),
]


def lib_path():
    return os.environ.get('PYFLOW_HIP_LIB', _build.LIB_PATH)


def lib():
    """Load the library once (after torch, so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError('libpyflow_hip.so is not built (expected at %s); run '
                           '`python -c "import __graft_entry__ as g; g.build()"`' % path)
    L = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().pf_last_error().decode(errors='replace')


def check(rc, what):
    if rc < 0:
        raise RuntimeError('%s failed (%d): %s' % (what, rc, last_error()))
    return rc


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('pyflow_amd runs its kernels on a ROCm GPU; no GPU is visible')


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dtype_code(dtype):
    if dtype == torch.float64:
        return PF_F64
    if dtype == torch.float32:
        return PF_F32
    raise ValueError('unsupported dtype %s (float64 or float32)' % dtype)


class Box:
    """Owner of one pf_box handle: grid sizes, spacing, dtype and all device
    workspace for one periodic box."""

    def __init__(self, nx, ny, nz, dx=1.0, dy=1.0, dz=1.0, dtype=torch.float64, device=None):
        require_gpu()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dtype = dtype
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.padded_shape = (self.nx + 4, self.ny + 4, self.nz + 4)
        self.interior_shape = (self.nx, self.ny, self.nz)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().pf_box_create(ctypes.byref(h), self.nx, self.ny, self.nz,
                                      float(dx), float(dy), float(dz), dtype_code(dtype)),
                  'pf_box_create')
        self.h = h
        self.spacing = (float(dx), float(dy), float(dz))

    def set_spacing(self, dx, dy, dz):
        check(lib().pf_box_set_spacing(self.h, float(dx), float(dy), float(dz)),
              'pf_box_set_spacing')
        self.spacing = (float(dx), float(dy), float(dz))

    def set_persistent(self, on):
        """Allow (or forbid) the persistent on-chip BiCGStab kernel for this
        box's solves; returns whether the box has that kernel's buffers."""
        rc = lib().pf_box_set_persistent(self.h, 1 if on else 0)
        check(rc, 'pf_box_set_persistent')
        return rc == 1

    @property
    def last_solve_persistent(self):
        """Whether the last BiCGStab solve ran as the persistent kernel."""
        return check(lib().pf_solve_last_persistent(self.h), 'pf_solve_last_persistent') == 1

    @property
    def device_bytes(self):
        return int(lib().pf_box_device_bytes(self.h))

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            lib().pf_box_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- argument checking -------------------------------------------------
    def _ptr(self, t, shape, name):
        if not isinstance(t, torch.Tensor):
            raise TypeError('%s must be a torch.Tensor' % name)
        if t.device.type != 'cuda':
            raise ValueError('%s must live on the GPU (got %s)' % (name, t.device))
        if t.dtype != self.dtype:
            raise ValueError('%s has dtype %s, box uses %s' % (name, t.dtype, self.dtype))
        if tuple(t.shape) != tuple(shape):
            raise ValueError('%s has shape %s, expected %s' % (name, tuple(t.shape), tuple(shape)))
        if not t.is_contiguous():
            raise ValueError('%s must be contiguous' % name)
        return ctypes.c_void_p(t.data_ptr())

    def padded(self, t, name='field'):
        return self._ptr(t, self.padded_shape, name)

    def interior(self, t, name='field'):
        return self._ptr(t, self.interior_shape, name)

    def sfs_array(self, sfs6):
        if sfs6 is None:
            return None
        arr = (ctypes.c_void_p * 6)(*[self.padded(f, 'sfs[%d]' % m).value for m, f in enumerate(sfs6)])
        return ctypes.cast(arr, _PP), arr

    def stream(self):
        return stream_ptr(self.device)

    # -- entry points --------------------------------------------------------
    def fill_halo(self, a):
        check(lib().pf_fill_halo(self.h, self.padded(a), self.stream()), 'pf_fill_halo')

    def rk4_step(self, u, v, w, ut, vt, wt, dt, mu, rho, sfs6=None):
        sa = self.sfs_array(sfs6)
        check(lib().pf_rk4_step(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                                self.padded(ut, 'ut'), self.padded(vt, 'vt'), self.padded(wt, 'wt'),
                                float(dt), float(mu), float(rho), sa[0] if sa else None,
                                self.stream()), 'pf_rk4_step')

    def euler_step(self, u, v, w, dt, mu, rho, sfs6=None):
        sa = self.sfs_array(sfs6)
        check(lib().pf_euler_step(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                                  float(dt), float(mu), float(rho), sa[0] if sa else None,
                                  self.stream()), 'pf_euler_step')

    def rhs_ns(self, u, v, w, mu, rho, ru, rv, rw, sfs6=None):
        sa = self.sfs_array(sfs6)
        check(lib().pf_rhs_ns(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                              float(mu), float(rho), sa[0] if sa else None,
                              self.interior(ru, 'rhs_u'), self.interior(rv, 'rhs_v'),
                              self.interior(rw, 'rhs_w'), self.stream()), 'pf_rhs_ns')

    def rk4_step_visc(self, u, v, w, ut, vt, wt, dt, visc, rho, sfs6=None):
        sa = self.sfs_array(sfs6)
        check(lib().pf_rk4_step_visc(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                     self.padded(w, 'w'), self.padded(ut, 'ut'),
                                     self.padded(vt, 'vt'), self.padded(wt, 'wt'), float(dt),
                                     self.padded(visc, 'VISC'), float(rho),
                                     sa[0] if sa else None, self.stream()), 'pf_rk4_step_visc')

    def euler_step_visc(self, u, v, w, dt, visc, rho, sfs6=None):
        sa = self.sfs_array(sfs6)
        check(lib().pf_euler_step_visc(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                       self.padded(w, 'w'), float(dt), self.padded(visc, 'VISC'),
                                       float(rho), sa[0] if sa else None, self.stream()),
              'pf_euler_step_visc')

    def rhs_ns_visc(self, u, v, w, visc, rho, ru, rv, rw, sfs6=None):
        sa = self.sfs_array(sfs6)
        check(lib().pf_rhs_ns_visc(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                   self.padded(w, 'w'), self.padded(visc, 'VISC'), float(rho),
                                   sa[0] if sa else None, self.interior(ru, 'rhs_u'),
                                   self.interior(rv, 'rhs_v'), self.interior(rw, 'rhs_w'),
                                   self.stream()), 'pf_rhs_ns_visc')

    # -- 'source' closures (velocity.py:360-383) ------------------------------
    def source_fields(self, G):
        """GX, GY, GZ as padded pointers; (nxo, nyo, nzo, 1) arrays, the
        reference's shape (sfsmodel_ML.py:146-151), are taken as padded."""
        out = []
        for name, g in zip(('GX', 'GY', 'GZ'), G):
            if isinstance(g, torch.Tensor) and g.dim() == 4 and g.shape[3] == 1:
                g = g[..., 0]
            out.append(self.padded(g, name))
        return out

    def source_inputs(self, u, v, w, ui, vi, wi):
        check(lib().pf_source_inputs(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                     self.padded(w, 'w'), self.padded(ui, 'u.var_i'),
                                     self.padded(vi, 'v.var_i'), self.padded(wi, 'w.var_i'),
                                     self.stream()), 'pf_source_inputs')

    def rhs_ns_src(self, u, v, w, mu, rho, G, ru, rv, rw):
        g = self.source_fields(G)
        check(lib().pf_rhs_ns_src(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                  self.padded(w, 'w'), float(mu), float(rho), *g,
                                  self.interior(ru, 'rhs_u'), self.interior(rv, 'rhs_v'),
                                  self.interior(rw, 'rhs_w'), self.stream()), 'pf_rhs_ns_src')

    def euler_step_src(self, u, v, w, dt, mu, rho, G):
        g = self.source_fields(G)
        check(lib().pf_euler_step_src(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                      self.padded(w, 'w'), float(dt), float(mu), float(rho), *g,
                                      self.stream()), 'pf_euler_step_src')

    def rk4_source_inputs(self, stage, u, v, w, ut, vt, wt, ui, vi, wi):
        check(lib().pf_rk4_source_inputs(self.h, int(stage), self.padded(u, 'u'),
                                         self.padded(v, 'v'), self.padded(w, 'w'),
                                         self.padded(ut, 'ut'), self.padded(vt, 'vt'),
                                         self.padded(wt, 'wt'), self.padded(ui, 'u.var_i'),
                                         self.padded(vi, 'v.var_i'), self.padded(wi, 'w.var_i'),
                                         self.stream()), 'pf_rk4_source_inputs')

    def rk4_stage_src(self, stage, u, v, w, ut, vt, wt, dt, mu, rho, G):
        g = self.source_fields(G)
        check(lib().pf_rk4_stage_src(self.h, int(stage), self.padded(u, 'u'), self.padded(v, 'v'),
                                     self.padded(w, 'w'), self.padded(ut, 'ut'),
                                     self.padded(vt, 'vt'), self.padded(wt, 'wt'), float(dt),
                                     float(mu), float(rho), *g, self.stream()),
              'pf_rk4_stage_src')

    def sfs_ml_source(self, params, norm21, ui, vi, wi, G):
        """The reference's 'source' ML closure: var_i arrays -> GX, GY, GZ."""
        n = int(lib().pf_sfs_ml_source_params())
        if not isinstance(params, torch.Tensor) or params.dtype != torch.float32 or \
                params.device.type != 'cuda' or params.numel() != n or not params.is_contiguous():
            raise ValueError('params must be a contiguous float32 GPU tensor of %d values' % n)
        nrm = (ctypes.c_float * 21)(*[float(x) for x in norm21])
        g = []
        for name, a in zip(('GX', 'GY', 'GZ'), G):
            if a.dim() == 4 and a.shape[3] == 1:
                a = a[..., 0]
            g.append(self.padded(a, name))
        check(lib().pf_sfs_ml_source(self.h, ctypes.c_void_p(params.data_ptr()), nrm,
                                     self.padded(ui, 'u'), self.padded(vi, 'v'),
                                     self.padded(wi, 'w'), *g, self.stream()),
              'pf_sfs_ml_source')

    def sfs_ml_source_reserve(self):
        """Reserve the ML closure's workspace in the box (allocates once)."""
        check(lib().pf_sfs_ml_source_reserve(self.h), 'pf_sfs_ml_source_reserve')

    def sfs_ml_source_train_reserve(self, norm21):
        """Reserve the closure reverse's workspace (allocates once) and keep
        the normalisation table for the adjoint closure term."""
        nrm = (ctypes.c_float * 21)(*[float(x) for x in norm21])
        check(lib().pf_sfs_ml_source_train_reserve(self.h, nrm), 'pf_sfs_ml_source_train_reserve')

    def sfs_ml_source_cache_budget(self, nbytes):
        """Bytes the closure's all-cell training cache may take (before
        sfs_ml_source_train_reserve; negative = the default, 48 GiB)."""
        check(lib().pf_sfs_ml_source_cache_budget(self.h, int(nbytes)),
              'pf_sfs_ml_source_cache_budget')

    def sfs_ml_source_vjp(self, params, ui, vi, wi, G, dI=None, dparams=None):
        """dg/d(ui, vi, wi) into dI (three FP32 padded arrays, optional) and
        dg/dparams added to dparams (FP64, optional), for g = sum(G_closure * G)."""
        n = int(lib().pf_sfs_ml_source_params())
        if not isinstance(params, torch.Tensor) or params.dtype != torch.float32 or \
                params.device.type != 'cuda' or params.numel() != n or not params.is_contiguous():
            raise ValueError('params must be a contiguous float32 GPU tensor of %d values' % n)
        g = [self.padded(a, name) for name, a in zip(('gx', 'gy', 'gz'), G)]
        d = [None, None, None]
        if dI is not None:
            for q, a in enumerate(dI):
                if a.dtype != torch.float32 or tuple(a.shape) != self.padded_shape or \
                        not a.is_contiguous() or a.device.type != 'cuda':
                    raise ValueError('dI must be contiguous float32 padded GPU arrays')
                d[q] = ctypes.c_void_p(a.data_ptr())
        dp = None
        if dparams is not None:
            if dparams.dtype != torch.float64 or dparams.numel() != n or not dparams.is_contiguous():
                raise ValueError('dparams must be a contiguous float64 tensor of %d values' % n)
            dp = ctypes.cast(ctypes.c_void_p(dparams.data_ptr()), _PD)
        check(lib().pf_sfs_ml_source_vjp(self.h, ctypes.c_void_p(params.data_ptr()),
                                         self.padded(ui, 'u'), self.padded(vi, 'v'),
                                         self.padded(wi, 'w'), *g, *d, dp, self.stream()),
              'pf_sfs_ml_source_vjp')

    def sfs_smagorinsky(self, u, v, w, rho, Cs, Delta, mu, visc):
        check(lib().pf_sfs_smagorinsky(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                       self.padded(w, 'w'), float(rho), float(Cs), float(Delta),
                                       float(mu), self.padded(visc, 'VISC'), self.stream()),
              'pf_sfs_smagorinsky')

    def div_vel(self, u, v, w, divg):
        check(lib().pf_div_vel(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                               self.interior(divg, 'divg'), self.stream()), 'pf_div_vel')

    def bicgstab(self, DP, p, source, rho, dt, tol, maxit):
        res = ctypes.c_double(0.0)
        its = ctypes.c_int(0)
        st = check(lib().pf_bicgstab_solve(self.h, self.padded(DP, 'DP'), self.padded(p, 'p'),
                                           self.interior(source, 'source'), float(rho), float(dt),
                                           float(tol), int(maxit), ctypes.byref(res),
                                           ctypes.byref(its), self.stream()), 'pf_bicgstab_solve')
        return res.value, its.value, st

    def bicgstab_mpir(self, DP, p, source, rho, dt, tol, maxit):
        res = ctypes.c_double(0.0)
        its = ctypes.c_int(0)
        st = check(lib().pf_bicgstab_mpir_solve(self.h, self.padded(DP, 'DP'), self.padded(p, 'p'),
                                                self.interior(source, 'source'), float(rho),
                                                float(dt), float(tol), int(maxit),
                                                ctypes.byref(res), ctypes.byref(its),
                                                self.stream()), 'pf_bicgstab_mpir_solve')
        return res.value, its.value, st

    def mpir_configure(self, inner_tol=1e-1, inner_maxit=500, outer_maxit=100):
        """Refinement settings; reserves the refinement workspace (allocates),
        which bicgstab_mpir and vel_corr with the refinement solver need."""
        check(lib().pf_mpir_configure(self.h, float(inner_tol), int(inner_maxit), int(outer_maxit)),
              'pf_mpir_configure')

    def mpir_solution(self):
        """(x, b) of the last refinement solve: FP64 interior tensors."""
        x = torch.empty(self.interior_shape, dtype=torch.float64, device=self.device)
        b = torch.empty_like(x)
        check(lib().pf_mpir_solution(self.h, ctypes.c_void_p(x.data_ptr()),
                                     ctypes.c_void_p(b.data_ptr()), self.stream()),
              'pf_mpir_solution')
        return x, b

    def jacobi(self, pOld, p, source, rho, dt, nit):
        res = ctypes.c_double(0.0)
        check(lib().pf_jacobi_solve(self.h, self.padded(pOld, 'pOld'), self.padded(p, 'p'),
                                    self.interior(source, 'source'), float(rho), float(dt), int(nit),
                                    ctypes.byref(res), self.stream()), 'pf_jacobi_solve')
        return res.value

    def vel_corr(self, u, v, w, DP, p, source, solver, src_scale, corr_scale, tol, iters):
        res = ctypes.c_double(0.0)
        its = ctypes.c_int(0)
        st = check(lib().pf_vel_corr(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                     self.padded(w, 'w'), self.padded(DP, 'DP'), self.padded(p, 'p'),
                                     self.interior(source, 'source'), int(solver), float(src_scale),
                                     float(corr_scale), float(tol), int(iters), ctypes.byref(res),
                                     ctypes.byref(its), self.stream()), 'pf_vel_corr')
        return res.value, its.value, st

    def apply_laplacian(self, A, B):
        check(lib().pf_apply_laplacian(self.h, self.padded(A, 'A'), self.interior(B, 'B'),
                                       self.stream()), 'pf_apply_laplacian')

    def stats(self, u, v, w, divg=None):
        out = (ctypes.c_double * 5)()
        check(lib().pf_stats(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                             self.interior(divg, 'divg') if divg is not None else None,
                             out, self.stream()), 'pf_stats')
        return tuple(out)

    def metric_pair(self, kind, axis, upper, dst, src, coef):
        check(lib().pf_metric_pair(self.h, kind, axis, 1 if upper else 0, self.padded(dst, 'dst'),
                                   self.padded(src, 'src'), float(coef), self.stream()),
              'pf_metric_pair')

    def metric_edge4(self, plane, dst, src):
        check(lib().pf_metric_edge4(self.h, plane, self.padded(dst, 'dst'), self.padded(src, 'src'),
                                    self.stream()), 'pf_metric_edge4')

    def div_visc(self, FX, FY, FZ, rhs):
        check(lib().pf_div_visc(self.h, self.padded(FX, 'FX'), self.padded(FY, 'FY'),
                                self.padded(FZ, 'FZ'), self.interior(rhs, 'rhs'), self.stream()),
              'pf_div_visc')

    def div_vel_over(self, u, v, w, divg_o):
        check(lib().pf_div_vel_over(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                    self.padded(w, 'w'), self.padded(divg_o, 'divg_o'),
                                    self.stream()), 'pf_div_vel_over')

    def vel_conv(self, axis, diagonal, A, B, rhs, sign):
        check(lib().pf_vel_conv(self.h, axis, 1 if diagonal else 0, self.padded(A, 'A'),
                                self.padded(B, 'B'), self.interior(rhs, 'rhs'), float(sign),
                                self.stream()), 'pf_vel_conv')

    def profile_enable(self, period=1):
        """Time every `period`-th launch of each Krylov kernel (0/False = off)."""
        check(lib().pf_profile_enable(self.h, int(period)), 'pf_profile_enable')

    def profile_read(self):
        """{'KA': (ms_total, launches), 'KB': ..., 'KC': ...,
        'KP': (ms_total, launches, iterations) of persistent solves}"""
        out = {}
        for k, name in enumerate(('KA', 'KB', 'KC')):
            ms = ctypes.c_double(0.0)
            n = ctypes.c_int64(0)
            check(lib().pf_profile_read(self.h, k, ctypes.byref(ms), ctypes.byref(n)), 'pf_profile_read')
            out[name] = (ms.value, n.value)
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        it = ctypes.c_int64(0)
        check(lib().pf_profile_read_persistent(self.h, ctypes.byref(ms), ctypes.byref(n),
                                               ctypes.byref(it)), 'pf_profile_read_persistent')
        out['KP'] = (ms.value, n.value, it.value)
        return out

    def residual_history_enable(self, capacity):
        """Keep max|r_k|/max|r0| of up to `capacity` BiCGStab iterations per solve (0 = off)."""
        check(lib().pf_residual_history_enable(self.h, int(capacity)), 'pf_residual_history_enable')
        self._hist_cap = int(capacity)

    def residual_history(self):
        """The last solve's per-iteration relative residuals (list of floats)."""
        cap = getattr(self, '_hist_cap', 0)
        buf = (ctypes.c_double * max(cap, 1))()
        n = ctypes.c_int(0)
        check(lib().pf_residual_history(self.h, buf, cap, ctypes.byref(n), self.stream()),
              'pf_residual_history')
        return [buf[i] for i in range(n.value)]

    def sfs_mlp(self, u, v, w, params, sfs6):
        if params.dtype != torch.float32 or params.device.type != 'cuda' or not params.is_contiguous():
            raise ValueError('MLP parameters must be a contiguous float32 GPU tensor')
        if params.numel() != lib().pf_sfs_mlp_param_count():
            raise ValueError('wrong number of MLP parameters')
        arr = (ctypes.c_void_p * 6)(*[self.padded(f, 'sfs[%d]' % m).value for m, f in enumerate(sfs6)])
        check(lib().pf_sfs_mlp(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                               ctypes.c_void_p(params.data_ptr()), ctypes.cast(arr, _PP),
                               self.stream()), 'pf_sfs_mlp')

    def sfs_gradient(self, u, v, w, sfs6):
        arr = (ctypes.c_void_p * 6)(*[self.padded(f, 'sfs[%d]' % m).value for m, f in enumerate(sfs6)])
        check(lib().pf_sfs_gradient(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                    self.padded(w, 'w'), ctypes.cast(arr, _PP), self.stream()),
              'pf_sfs_gradient')

    # -- adjoint step and reverse-mode kernels ---------------------------------
    @staticmethod
    def _param_count(kind):
        """Packed parameters of closure kind 1 (MLP) or 2 ('source' ML closure)."""
        if kind == 2:
            return int(lib().pf_sfs_ml_source_params())
        return int(lib().pf_sfs_mlp_param_count())

    def _params(self, params, kind=1):
        if params is None:
            return None
        if params.dtype != torch.float32 or params.device.type != 'cuda' or not params.is_contiguous():
            raise ValueError('closure parameters must be a contiguous float32 GPU tensor')
        if params.numel() != self._param_count(kind):
            raise ValueError('closure kind %d takes %d parameters, got %d'
                             % (kind, self._param_count(kind), params.numel()))
        return ctypes.c_void_p(params.data_ptr())

    def _ptrs(self, fields, name):
        arr = (ctypes.c_void_p * len(fields))(
            *[self.padded(f, '%s[%d]' % (name, m)).value for m, f in enumerate(fields)])
        return ctypes.cast(arr, _PP), arr

    @staticmethod
    def _dparams(d, kind=1):
        if d is None:
            return None
        if d.dtype != torch.float64 or d.device.type != 'cuda' or not d.is_contiguous() or \
                d.numel() != Box._param_count(kind):
            raise ValueError('dparams must be a contiguous float64 GPU tensor of the closure '
                             "kind's parameter count")
        return ctypes.cast(ctypes.c_void_p(d.data_ptr()), _PD)

    def sfs_mlp_f32(self, u, v, w, params, sfs6):
        pa = self._ptrs(sfs6, 'sfs')
        check(lib().pf_sfs_mlp_f32(self.h, self.padded(u, 'u'), self.padded(v, 'v'),
                                   self.padded(w, 'w'), self._params(params), pa[0], self.stream()),
              'pf_sfs_mlp_f32')

    def adj_work_bytes(self):
        return int(lib().pf_adj_work_bytes(self.h))

    def rhs_adj(self, u, v, w, ua, va, wa, mu, rho, ru, rv, rw, delta3=None):
        """mu: the constant viscosity, or the padded VISC field (tensor)."""
        da = self._ptrs(delta3, 'delta') if delta3 is not None else None
        args = (self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                self.padded(ua, 'ua'), self.padded(va, 'va'), self.padded(wa, 'wa'))
        tail = (float(rho), da[0] if da else None, self.interior(ru, 'rhs_u'),
                self.interior(rv, 'rhs_v'), self.interior(rw, 'rhs_w'), self.stream())
        if isinstance(mu, torch.Tensor):
            check(lib().pf_rhs_adj_visc(self.h, *args, self.padded(mu, 'VISC'), *tail),
                  'pf_rhs_adj_visc')
        else:
            check(lib().pf_rhs_adj(self.h, *args, float(mu), *tail), 'pf_rhs_adj')

    def adj_closure_term(self, kind, u, v, w, params, ua, va, wa, work, delta3, dparams=None):
        da = self._ptrs(delta3, 'delta')
        check(lib().pf_adj_closure_term(self.h, int(kind), self.padded(u, 'u'), self.padded(v, 'v'),
                                        self.padded(w, 'w'), self._params(params, kind),
                                        self.padded(ua, 'ua'), self.padded(va, 'va'),
                                        self.padded(wa, 'wa'), self._work(work), da[0],
                                        self._dparams(dparams, kind), self.stream()),
              'pf_adj_closure_term')

    def _work(self, work):
        if work is None:
            return None
        if work.device.type != 'cuda' or work.numel() * work.element_size() < self.adj_work_bytes():
            raise ValueError('adjoint work buffer must be a GPU tensor of >= %d bytes'
                             % self.adj_work_bytes())
        return ctypes.c_void_p(work.data_ptr())

    def adj_rk4_step(self, u, v, w, ua, va, wa, uta, vta, wta, dt, mu, rho, closure=-1,
                     params=None, work=None, delta3=None, dparams=None):
        """mu: the constant viscosity, or the padded VISC field (tensor)."""
        da = self._ptrs(delta3, 'delta') if delta3 is not None else None
        head = (self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                self.padded(ua, 'ua'), self.padded(va, 'va'), self.padded(wa, 'wa'),
                self.padded(uta, 'uta'), self.padded(vta, 'vta'), self.padded(wta, 'wta'),
                float(dt))
        tail = (float(rho), int(closure), self._params(params, closure), self._work(work),
                da[0] if da else None, self._dparams(dparams, closure), self.stream())
        if isinstance(mu, torch.Tensor):
            check(lib().pf_adj_rk4_step_visc(self.h, *head, self.padded(mu, 'VISC'), *tail),
                  'pf_adj_rk4_step_visc')
        else:
            check(lib().pf_adj_rk4_step(self.h, *head, float(mu), *tail), 'pf_adj_rk4_step')

    def adj_euler_step(self, u, v, w, ua, va, wa, dt, mu, rho, closure=-1, params=None, work=None,
                       delta3=None, dparams=None):
        """mu: the constant viscosity, or the padded VISC field (tensor)."""
        da = self._ptrs(delta3, 'delta') if delta3 is not None else None
        head = (self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                self.padded(ua, 'ua'), self.padded(va, 'va'), self.padded(wa, 'wa'), float(dt))
        tail = (float(rho), int(closure), self._params(params, closure), self._work(work),
                da[0] if da else None, self._dparams(dparams, closure), self.stream())
        if isinstance(mu, torch.Tensor):
            check(lib().pf_adj_euler_step_visc(self.h, *head, self.padded(mu, 'VISC'), *tail),
                  'pf_adj_euler_step_visc')
        else:
            check(lib().pf_adj_euler_step(self.h, *head, float(mu), *tail), 'pf_adj_euler_step')

    def adj_conv(self, d_axis, s2_axis, centre, a, vi, rhs, coef):
        check(lib().pf_adj_conv(self.h, int(d_axis), int(s2_axis), 1 if centre else 0,
                                self.padded(a, 'a'), self.padded(vi, 'vi'), self.interior(rhs, 'rhs'),
                                float(coef), self.stream()), 'pf_adj_conv')

    def rhs_vjp(self, u, v, w, mu, rho, lam3, work, out3, kprev3=None, hcoef=0.0):
        if work.device.type != 'cuda' or work.dtype != self.dtype or \
                work.numel() < 33 * self.nx * self.ny * self.nz:
            raise ValueError('rhs_vjp work must hold 33*nx*ny*nz elements of the box dtype')
        k = [self.padded(t, 'kprev') for t in kprev3] if kprev3 is not None else [None] * 3
        check(lib().pf_rhs_vjp(self.h, self.padded(u, 'u'), self.padded(v, 'v'), self.padded(w, 'w'),
                               float(mu), float(rho), self.padded(lam3[0], 'lu'),
                               self.padded(lam3[1], 'lv'), self.padded(lam3[2], 'lw'),
                               ctypes.c_void_p(work.data_ptr()), self.padded(out3[0], 'ou'),
                               self.padded(out3[1], 'ov'), self.padded(out3[2], 'ow'),
                               k[0], k[1], k[2], float(hcoef), self.stream()), 'pf_rhs_vjp')

    def sfs_vjp(self, lam3, taubar6, accumulate=True):
        ta = self._ptrs(taubar6, 'taubar')
        check(lib().pf_sfs_vjp(self.h, self.padded(lam3[0], 'lu'), self.padded(lam3[1], 'lv'),
                               self.padded(lam3[2], 'lw'), ta[0], 1 if accumulate else 0,
                               self.stream()), 'pf_sfs_vjp')

    def fold_halo(self, a):
        check(lib().pf_fold_halo(self.h, self.padded(a, 'field'), self.stream()), 'pf_fold_halo')

    def closure_vjp_waves(self):
        return int(lib().pf_closure_vjp_waves(self.h))

    def closure_vjp(self, kind, u, v, w, params, taubar6, gbar9, f3, dparams=None, wpart=None):
        ta = self._ptrs(taubar6, 'taubar')
        if gbar9.dtype != self.dtype or gbar9.numel() != 9 * self.padded_shape[0] * \
                self.padded_shape[1] * self.padded_shape[2] or not gbar9.is_contiguous():
            raise ValueError('gbar9 must hold 9 padded arrays of the box dtype')
        wp = None
        if wpart is not None:
            if wpart.dtype != torch.float64 or wpart.numel() < self.closure_vjp_waves() * \
                    lib().pf_sfs_mlp_param_count():
                raise ValueError('wpart must hold closure_vjp_waves()*746 float64 values')
            wp = ctypes.cast(ctypes.c_void_p(wpart.data_ptr()), _PD)
        check(lib().pf_closure_vjp(self.h, int(kind), self.padded(u, 'u'), self.padded(v, 'v'),
                                   self.padded(w, 'w'), self._params(params), ta[0],
                                   ctypes.c_void_p(gbar9.data_ptr()), self.padded(f3[0], 'fu'),
                                   self.padded(f3[1], 'fv'), self.padded(f3[2], 'fw'),
                                   self._dparams(dparams), wp, self.stream()), 'pf_closure_vjp')


class Batch:
    """Owner of one pf_batch: batched, graph-captured time steps of many
    resident boxes (all of one dtype).  set() takes, per box, the Box and its
    fields; step() advances all of them by one time step."""

    FIELDS = ('u', 'v', 'w', 'ut', 'vt', 'wt', 'DP', 'p')

    def __init__(self, capacity, dtype=torch.float32):
        require_gpu()
        self.capacity = int(capacity)
        self.dtype = dtype
        h = ctypes.c_void_p()
        check(lib().pf_batch_create(ctypes.byref(h), self.capacity, dtype_code(dtype)),
              'pf_batch_create')
        self.h = h
        self.n = 0
        self._keep = None

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            lib().pf_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, members):
        """members: list of (Box, dict of the 8 padded fields + 'src', params)
        with params = (dt, mu, rho, tol, maxit)."""
        n = len(members)
        if n > self.capacity:
            raise ValueError('batch holds at most %d boxes' % self.capacity)
        boxes = (ctypes.c_void_p * max(n, 1))()
        fields = (ctypes.c_void_p * max(8 * n, 1))()
        srcs = (ctypes.c_void_p * max(n, 1))()
        prm = (ctypes.c_double * max(5 * n, 1))()
        for i, (box, f, params) in enumerate(members):
            if box.dtype != self.dtype:
                raise ValueError('box dtype %s differs from the batch dtype %s' % (box.dtype, self.dtype))
            boxes[i] = box.h.value
            for q, name in enumerate(self.FIELDS):
                fields[8 * i + q] = box.padded(f[name], name).value
            srcs[i] = box.interior(f['src'], 'src').value
            for q in range(5):
                prm[5 * i + q] = float(params[q])
        check(lib().pf_batch_set(self.h, n, ctypes.cast(boxes, _PP), ctypes.cast(fields, _PP),
                                 ctypes.cast(srcs, _PP), ctypes.cast(prm, _PD)), 'pf_batch_set')
        self.n = n
        self._keep = members

    def set_group_bytes(self, nbytes):
        """Cache-residency budget of the chunk's box groups (0: one group)."""
        check(lib().pf_batch_set_group_bytes(self.h, int(nbytes)), 'pf_batch_set_group_bytes')

    def set_persistent(self, on):
        """Big boxes (a persistent solve that fills the GPU) solve each step as
        one cooperative launch or on the chunk kernels (on = False or 0).  An
        int caps the persistent solves per advance call (True: 1, the
        default: the other boxes get a chunk between persistent solves).
        True if a resident box now takes the persistent solve."""
        cap = int(on)
        if cap < 0:
            raise ValueError('set_persistent: on must be a bool or an int >= 0')
        rc = lib().pf_batch_set_persistent(self.h, cap)
        check(rc if rc < 0 else 0, 'pf_batch_set_persistent')
        return rc == 1

    def advance(self, start, device=None):
        """Start a step for the boxes flagged in `start` (idle ones), run until
        at least one running box finishes its step; returns (indices of the
        boxes that finished, their monitor rows {i: [8 values]}, chunks)."""
        n = self.n
        st = (ctypes.c_int * max(n, 1))(*[1 if x else 0 for x in start])
        fin = (ctypes.c_int * max(n, 1))()
        out = (ctypes.c_double * max(8 * n, 1))()
        chunks = ctypes.c_int(0)
        check(lib().pf_batch_advance(self.h, ctypes.cast(st, _PI), ctypes.cast(fin, _PI),
                                     ctypes.cast(out, _PD), ctypes.byref(chunks),
                                     stream_ptr(device)), 'pf_batch_advance')
        done = [i for i in range(n) if fin[i]]
        return done, {i: list(out[8 * i:8 * i + 8]) for i in done}, chunks.value
