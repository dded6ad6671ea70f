"""ctypes binding of the p2d C ABI (include/p2d.h) -> libp2d.so.

The HIP library is the only compute path: when libp2d.so is missing this
module raises at import, and when no GPU is visible every compute call raises
P2DError (status P2D_NO_DEVICE).  Arrays may be numpy (host memory) or torch
tensors on the GPU (device memory, used in place); shapes follow the C ABI:
cells (ny, nx), x-faces (ny, nx+1), y-faces (ny+1, nx), row-major.

PyTorch-ROCm ships its own HIP runtime (libamdhip64); the first one loaded in
a process serves both it and libp2d.so.  In a process that uses torch tensors
with this binding, import torch before this module (the package does so when
torch is installed) so that both share torch's runtime, its streams included.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libp2d.so")

P2D_OK, P2D_NOT_CONVERGED, P2D_NAN, P2D_BAD_ARG, P2D_HIP_ERROR, P2D_NO_DEVICE = range(6)
MEM_HOST, MEM_DEVICE = 0, 1
FP64, MIXED = 0, 1
PREC_F64, PREC_F32 = 0, 1

# Kernels p2d_time_kernel can time (finest level).
KERNEL_DIR_F32 = 0
KERNEL_DIR_F64 = 1
KERNEL_PRE0_F32 = 2
KERNEL_POST0_F32 = 3
KERNEL_CORRECT_RESID = 4
KERNEL_CORRECT_RESID_F64 = 5
KERNEL_PRE0_F64 = 6
KERNEL_POST0_F64 = 7
# p2d_time_kernel only: one level-set re-initialisation step
TIME_REINIT_STEP = 5

EXPORTS = [
    "p2d_config_default", "p2d_create", "p2d_destroy", "p2d_set_stream", "p2d_set_fluid",
    "p2d_set_coefficients", "p2d_project", "p2d_get_pressure", "p2d_get_coefficients", "p2d_adjoint",
    "p2d_apply_stencil", "p2d_vcycle", "p2d_mg_levels", "p2d_time_kernel", "p2d_kernel_timing",
    "p2d_kernel_time", "p2d_status_string", "p2d_reinitialize", "p2d_levelset_normals", "p2d_advect_levelset",
    "p2d_advect_velocity", "p2d_apply_viscosity", "p2d_step_params_default", "p2d_advance",
    "p2d_cfl_dt", "p2d_visc_create", "p2d_visc_destroy", "p2d_visc_set_stream", "p2d_visc_solve",
    "p2d_oil_water_params_default", "p2d_visc_oil_water", "p2d_divergence", "p2d_solve_pressure",
    "p2d_cfl_dt_noalloc",
]


class P2DConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("h", ctypes.c_double),
        ("precision", ctypes.c_int), ("mg_max_levels", ctypes.c_int),
        ("mg_pre_sweeps", ctypes.c_int), ("mg_post_sweeps", ctypes.c_int),
        ("rtol", ctypes.c_double), ("div_tol", ctypes.c_double),
        ("max_iterations", ctypes.c_int), ("max_outer", ctypes.c_int),
        ("inner_rtol", ctypes.c_double), ("abs_tol", ctypes.c_double), ("mg_omega", ctypes.c_double),
        ("mg_omega_coarse", ctypes.c_double),
    ]


class P2DStats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int), ("iterations", ctypes.c_int), ("outer", ctypes.c_int),
        ("rel_residual", ctypes.c_double), ("max_abs_residual", ctypes.c_double),
        ("max_div", ctypes.c_double), ("max_u", ctypes.c_double), ("max_v", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class P2DStepParams(ctypes.Structure):
    _fields_ = [
        ("gravity_x", ctypes.c_double), ("gravity_y", ctypes.c_double),
        ("rho_minus", ctypes.c_double), ("rho_plus", ctypes.c_double), ("sigma", ctypes.c_double),
        ("mu_minus", ctypes.c_double), ("mu_plus", ctypes.c_double), ("mollification", ctypes.c_double),
        ("weno_epsilon", ctypes.c_double), ("reinit_iterations", ctypes.c_int), ("reinit_scaling", ctypes.c_double),
        ("rk_order", ctypes.c_int),
    ]


class P2DOilWaterParams(ctypes.Structure):
    _fields_ = [
        ("x_min", ctypes.c_double), ("Re", ctypes.c_double), ("rho_oil", ctypes.c_double),
        ("rho_water", ctypes.c_double), ("mu_oil", ctypes.c_double), ("mu_water", ctypes.c_double),
        ("mollification", ctypes.c_double), ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int),
        ("boundary", ctypes.c_int),
    ]


class P2DError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    cfgp, stp = ctypes.POINTER(P2DConfig), ctypes.POINTER(P2DStats)
    sig = {
        "p2d_config_default": (None, [cfgp, i, i, d]),
        "p2d_create": (i, [cfgp, ctypes.POINTER(vp)]),
        "p2d_destroy": (i, [vp]),
        "p2d_set_stream": (i, [vp, vp]),
        "p2d_set_fluid": (i, [vp, vp, d, d, d, i]),
        "p2d_set_coefficients": (i, [vp, vp, vp, vp, vp, i]),
        "p2d_project": (i, [vp, vp, vp, d, stp, i]),
        "p2d_get_pressure": (i, [vp, vp, i, i]),
        "p2d_divergence": (i, [vp, vp, vp, d, vp, i]),
        "p2d_solve_pressure": (i, [vp, vp, vp, stp, i]),
        "p2d_get_coefficients": (i, [vp, vp, vp, vp, i]),
        "p2d_adjoint": (i, [vp, vp, vp, vp, vp, vp, vp, stp, i]),
        "p2d_apply_stencil": (i, [vp, vp, vp, i, i]),
        "p2d_vcycle": (i, [vp, vp, vp, i, i]),
        "p2d_mg_levels": (i, [vp]),
        "p2d_time_kernel": (i, [vp, i, i, ctypes.POINTER(d)]),
        "p2d_kernel_timing": (i, [vp, i, i]),
        "p2d_kernel_time": (i, [vp, i, ctypes.POINTER(d), ctypes.POINTER(i)]),
        "p2d_status_string": (ctypes.c_char_p, [i]),
        "p2d_reinitialize": (i, [vp, vp, i, d, d, i]),
        "p2d_levelset_normals": (i, [vp, vp, vp, vp, i]),
        "p2d_advect_levelset": (i, [vp, vp, vp, vp, d, d, i]),
        "p2d_advect_velocity": (i, [vp, vp, vp, d, i]),
        "p2d_apply_viscosity": (i, [vp, vp, vp, d, d, d, d, i]),
        "p2d_step_params_default": (None, [ctypes.POINTER(P2DStepParams)]),
        "p2d_advance": (i, [vp, vp, vp, vp, d, ctypes.POINTER(P2DStepParams), stp, i]),
        "p2d_cfl_dt": (i, [vp, vp, vp, vp, ctypes.POINTER(P2DStepParams), d, ctypes.POINTER(d), i]),
        "p2d_cfl_dt_noalloc": (i, [i, i, d, vp, vp, vp, ctypes.POINTER(P2DStepParams), d, vp, vp]),
        "p2d_visc_create": (i, [i, i, d, d, ctypes.POINTER(vp)]),
        "p2d_visc_destroy": (i, [vp]),
        "p2d_visc_set_stream": (i, [vp, vp]),
        "p2d_visc_solve": (i, [vp, i, vp, vp, vp, vp, vp, vp, vp, i, d, i, stp, i]),
        "p2d_oil_water_params_default": (None, [ctypes.POINTER(P2DOilWaterParams)]),
        "p2d_visc_oil_water": (i, [vp, vp, vp, vp, d, ctypes.POINTER(P2DOilWaterParams), stp, stp, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def status_string(status):
    return lib.p2d_status_string(status).decode()


def default_config(nx, ny, h):
    c = P2DConfig()
    lib.p2d_config_default(ctypes.byref(c), nx, ny, h)
    return c


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class _Buf:
    """Pointer + memory kind for one argument; host numpy arrays are made contiguous copies when needed."""

    def __init__(self, a, shape, dtype=np.float64, writable=False):
        if a is None:
            self.ptr, self.mem, self.arr = None, None, None
            return
        if _is_torch(a):
            import torch
            want = torch.float64 if dtype == np.float64 else torch.float32
            if tuple(a.shape) != tuple(shape) or a.dtype != want or not a.is_contiguous():
                raise ValueError(f"tensor must be contiguous {want} of shape {shape}, got {a.dtype} {tuple(a.shape)}")
            self.mem = MEM_DEVICE if a.is_cuda else MEM_HOST
            self.ptr, self.arr = a.data_ptr(), a
            return
        arr = np.asarray(a)
        if arr.shape != tuple(shape):
            raise ValueError(f"array must have shape {shape}, got {arr.shape}")
        if writable:
            if arr.dtype != dtype or not arr.flags["C_CONTIGUOUS"]:
                raise ValueError(f"output array must be C-contiguous {np.dtype(dtype)}")
        else:
            arr = np.ascontiguousarray(arr, dtype=dtype)
        self.mem, self.ptr, self.arr = MEM_HOST, arr.ctypes.data, arr


def _mem(*bufs):
    kinds = {b.mem for b in bufs if b.mem is not None}
    if len(kinds) > 1:
        raise ValueError("mixing host and device arrays in one call")
    return kinds.pop() if kinds else MEM_HOST


def _check(rc, what, allow=(P2D_OK,)):
    if rc not in allow:
        raise P2DError(rc, what)
    return rc


class Context:
    """One projection context (p2d_ctx) for an nx x ny grid of spacing h."""

    def __init__(self, nx, ny, h, precision=FP64, **cfg):
        c = default_config(nx, ny, h)
        c.precision = precision
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown config field {k}")
            setattr(c, k, v)
        self.cfg = c
        self.nx, self.ny, self.h = nx, ny, h
        self._h = ctypes.c_void_p()
        _check(lib.p2d_create(ctypes.byref(c), ctypes.byref(self._h)), "p2d_create")

    # shapes
    @property
    def cell_shape(self):
        return (self.ny, self.nx)

    @property
    def u_shape(self):
        return (self.ny, self.nx + 1)

    @property
    def v_shape(self):
        return (self.ny + 1, self.nx)

    def close(self):
        if self._h:
            lib.p2d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        _check(lib.p2d_set_stream(self._h, stream_handle), "p2d_set_stream")

    def set_fluid(self, phi, rho_minus, rho_plus, sigma):
        b = _Buf(phi, self.cell_shape)
        _check(lib.p2d_set_fluid(self._h, b.ptr, rho_minus, rho_plus, sigma, b.mem), "p2d_set_fluid")

    def set_coefficients(self, beta_x, beta_y, phi, jc=None):
        a, b, p, j = (_Buf(beta_x, self.u_shape), _Buf(beta_y, self.v_shape), _Buf(phi, self.cell_shape),
                      _Buf(jc, self.cell_shape))
        _check(lib.p2d_set_coefficients(self._h, a.ptr, b.ptr, p.ptr, j.ptr, _mem(a, b, p, j)),
               "p2d_set_coefficients")

    def project(self, u, v, dt=1.0, allow_not_converged=False):
        """Project u, v in place (device tensors) or return projected copies (numpy). Returns (u, v, stats)."""
        if not _is_torch(u):
            u = np.array(u, dtype=np.float64, order="C", copy=True)
            v = np.array(v, dtype=np.float64, order="C", copy=True)
        bu, bv = _Buf(u, self.u_shape, writable=True), _Buf(v, self.v_shape, writable=True)
        st = P2DStats()
        allow = (P2D_OK, P2D_NOT_CONVERGED) if allow_not_converged else (P2D_OK,)
        _check(lib.p2d_project(self._h, bu.ptr, bv.ptr, dt, ctypes.byref(st), _mem(bu, bv)), "p2d_project", allow)
        return u, v, st.as_dict()

    def pressure(self, mean_zero=True, out=None):
        if out is None:
            out = np.empty(self.cell_shape)
        b = _Buf(out, self.cell_shape, writable=True)
        _check(lib.p2d_get_pressure(self._h, b.ptr, 1 if mean_zero else 0, b.mem), "p2d_get_pressure")
        return out

    def divergence(self, u, v, dt=1.0):
        """DetermineDivergence: (D u) / dt per cell, numpy or a torch tensor on the device."""
        bu, bv = _Buf(u, self.u_shape), _Buf(v, self.v_shape)
        out = np.empty(self.cell_shape) if bu.mem == MEM_HOST else u.new_empty(self.cell_shape)
        bo = _Buf(out, self.cell_shape, writable=True)
        _check(lib.p2d_divergence(self._h, bu.ptr, bv.ptr, dt, bo.ptr, _mem(bu, bv, bo)), "p2d_divergence")
        return out

    def solve_pressure(self, div, allow_not_converged=False):
        """DeterminePressure for a given divergence field: returns (p in the reference's gauge
        p(0,0) = 0, stats); the pressure is also kept for pressure()."""
        bd = _Buf(div, self.cell_shape)
        out = np.empty(self.cell_shape) if bd.mem == MEM_HOST else div.new_empty(self.cell_shape)
        bo = _Buf(out, self.cell_shape, writable=True)
        st = P2DStats()
        allow = (P2D_OK, P2D_NOT_CONVERGED) if allow_not_converged else (P2D_OK,)
        _check(lib.p2d_solve_pressure(self._h, bd.ptr, bo.ptr, ctypes.byref(st), _mem(bd, bo)),
               "p2d_solve_pressure", allow)
        return out, st.as_dict()

    def coefficients(self):
        bx, by, jc = np.empty(self.u_shape), np.empty(self.v_shape), np.empty(self.cell_shape)
        _check(lib.p2d_get_coefficients(self._h, bx.ctypes.data, by.ctypes.data, jc.ctypes.data, MEM_HOST),
               "p2d_get_coefficients")
        return bx, by, jc

    def adjoint(self, gu, gv, outs=None):
        """Returns (du, dv, dbeta_x, dbeta_y, stats)."""
        a, b = _Buf(gu, self.u_shape), _Buf(gv, self.v_shape)
        if outs is None:
            if a.mem == MEM_DEVICE:
                import torch
                dev = gu.device
                outs = (torch.empty(self.u_shape, dtype=torch.float64, device=dev),
                        torch.empty(self.v_shape, dtype=torch.float64, device=dev),
                        torch.empty(self.u_shape, dtype=torch.float64, device=dev),
                        torch.empty(self.v_shape, dtype=torch.float64, device=dev))
            else:
                outs = (np.empty(self.u_shape), np.empty(self.v_shape), np.empty(self.u_shape), np.empty(self.v_shape))
        o = [_Buf(x, s, writable=True) for x, s in zip(outs, (self.u_shape, self.v_shape, self.u_shape, self.v_shape))]
        st = P2DStats()
        _check(lib.p2d_adjoint(self._h, a.ptr, b.ptr, o[0].ptr, o[1].ptr, o[2].ptr, o[3].ptr, ctypes.byref(st),
                               _mem(a, b, *o)), "p2d_adjoint")
        return outs[0], outs[1], outs[2], outs[3], st.as_dict()

    def apply_stencil(self, x, prec=PREC_F64):
        dt = np.float64 if prec == PREC_F64 else np.float32
        bx = _Buf(x, self.cell_shape, dtype=dt)
        y = np.empty(self.cell_shape, dtype=dt) if bx.mem == MEM_HOST else x.new_empty(x.shape)
        by = _Buf(y, self.cell_shape, dtype=dt, writable=True)
        _check(lib.p2d_apply_stencil(self._h, bx.ptr, by.ptr, prec, _mem(bx, by)), "p2d_apply_stencil")
        return y

    def vcycle(self, r_scaled, prec=PREC_F64):
        dt = np.float64 if prec == PREC_F64 else np.float32
        br = _Buf(r_scaled, self.cell_shape, dtype=dt)
        z = np.empty(self.cell_shape, dtype=dt) if br.mem == MEM_HOST else r_scaled.new_empty(r_scaled.shape)
        bz = _Buf(z, self.cell_shape, dtype=dt, writable=True)
        _check(lib.p2d_vcycle(self._h, br.ptr, bz.ptr, prec, _mem(br, bz)), "p2d_vcycle")
        return z

    def reinitialize(self, phi, iterations=25, scaling=5.0, epsilon=1e-6):
        """Sussman WENO5 re-initialisation of phi (EULERIAN_FLUID_SOLVER_2D.h:2365-2400).

        A torch CUDA tensor is updated in place and returned; a numpy array gives a new array."""
        if _is_torch(phi):
            b = _Buf(phi, self.cell_shape, writable=True)
            out = phi
        else:
            out = np.array(phi, dtype=np.float64, order="C", copy=True)
            b = _Buf(out, self.cell_shape, writable=True)
        _check(lib.p2d_reinitialize(self._h, b.ptr, int(iterations), float(scaling), float(epsilon), b.mem),
               "p2d_reinitialize")
        return out

    def advect_levelset(self, phi, u, v, dt, epsilon=1e-6):
        """One forward-Euler WENO5 advection stage of phi by the MAC velocity (u, v)
        (ADVECTION_2D.cpp:165-200).  Torch CUDA tensors: phi in place; numpy: a new array."""
        bu, bv = _Buf(u, self.u_shape), _Buf(v, self.v_shape)
        if _is_torch(phi):
            b = _Buf(phi, self.cell_shape, writable=True)
            out = phi
        else:
            out = np.array(phi, dtype=np.float64, order="C", copy=True)
            b = _Buf(out, self.cell_shape, writable=True)
        _check(lib.p2d_advect_levelset(self._h, b.ptr, bu.ptr, bv.ptr, float(dt), float(epsilon), _mem(b, bu, bv)),
               "p2d_advect_levelset")
        return out

    def advect_velocity(self, u, v, dt):
        """One forward-Euler ENO3 advection stage of the MAC velocity by itself
        (ADVECTION_2D.cpp:95-160).  Torch CUDA tensors: in place; numpy: new arrays."""
        if _is_torch(u):
            bu, bv = _Buf(u, self.u_shape, writable=True), _Buf(v, self.v_shape, writable=True)
            uo, vo = u, v
        else:
            uo = np.array(u, dtype=np.float64, order="C", copy=True)
            vo = np.array(v, dtype=np.float64, order="C", copy=True)
            bu, bv = _Buf(uo, self.u_shape, writable=True), _Buf(vo, self.v_shape, writable=True)
        _check(lib.p2d_advect_velocity(self._h, bu.ptr, bv.ptr, float(dt), _mem(bu, bv)), "p2d_advect_velocity")
        return uo, vo

    def apply_viscosity(self, u, v, dt, mu_minus, mu_plus, epsilon=None):
        """Explicit air-water viscosity step of the MAC velocity with the fluid of the last
        set_fluid (VISCOSITY_2D.h:460-548); epsilon defaults to 1.5 h.  Torch CUDA tensors:
        in place; numpy: new arrays."""
        eps = 1.5 * self.h if epsilon is None else float(epsilon)
        if _is_torch(u):
            bu, bv = _Buf(u, self.u_shape, writable=True), _Buf(v, self.v_shape, writable=True)
            uo, vo = u, v
        else:
            uo = np.array(u, dtype=np.float64, order="C", copy=True)
            vo = np.array(v, dtype=np.float64, order="C", copy=True)
            bu, bv = _Buf(uo, self.u_shape, writable=True), _Buf(vo, self.v_shape, writable=True)
        _check(lib.p2d_apply_viscosity(self._h, bu.ptr, bv.ptr, float(dt), float(mu_minus), float(mu_plus), eps,
                                       _mem(bu, bv)), "p2d_apply_viscosity")
        return uo, vo

    @staticmethod
    def _step_params(params):
        prm = P2DStepParams()
        lib.p2d_step_params_default(ctypes.byref(prm))
        names = {k for k, _ in P2DStepParams._fields_}
        for k, val in params.items():
            if k not in names:
                raise TypeError(f"unknown step parameter {k}")
            setattr(prm, k, val)
        return prm

    def cfl_dt(self, phi, u, v, cfl_number=0.5, **params):
        """The CFL time step (EULERIAN_FLUID_SOLVER_2D.h:3517-3637); 0.0 when no term
        restricts it.  phi may be None when sigma is 0."""
        prm = self._step_params(params)
        bu, bv = _Buf(u, self.u_shape), _Buf(v, self.v_shape)
        bp = _Buf(phi, self.cell_shape) if phi is not None else None
        out = ctypes.c_double(0.0)
        bufs = (bu, bv) if bp is None else (bp, bu, bv)
        _check(lib.p2d_cfl_dt(self._h, bp.ptr if bp else None, bu.ptr, bv.ptr, ctypes.byref(prm), float(cfl_number),
                              ctypes.byref(out), _mem(*bufs)), "p2d_cfl_dt")
        return out.value

    def advance(self, phi, u, v, dt, allow_not_converged=False, **params):
        """One air-water time step (EULERIAN_FLUID_SOLVER_2D.h:1142-1797; TVD-RK3 unless
        rk_order says 1 or 2): level-set advection and re-initialisation, then the velocity
        stages (advection, gravity, viscosity, wall velocities, projection).  Keyword parameters
        are the fields of p2d_step_params (gravity_x, gravity_y, rho_minus, rho_plus, sigma,
        mu_minus, mu_plus, mollification, weno_epsilon, reinit_iterations, reinit_scaling,
        rk_order).  Torch CUDA tensors are
        advanced in place; numpy arrays are copied.  Returns (phi, u, v, stats of the last
        projection)."""
        prm = self._step_params(params)
        if not _is_torch(phi):
            phi = np.array(phi, dtype=np.float64, order="C", copy=True)
            u = np.array(u, dtype=np.float64, order="C", copy=True)
            v = np.array(v, dtype=np.float64, order="C", copy=True)
        bp = _Buf(phi, self.cell_shape, writable=True)
        bu, bv = _Buf(u, self.u_shape, writable=True), _Buf(v, self.v_shape, writable=True)
        st = P2DStats()
        allow = (P2D_OK, P2D_NOT_CONVERGED) if allow_not_converged else (P2D_OK,)
        _check(lib.p2d_advance(self._h, bp.ptr, bu.ptr, bv.ptr, float(dt), ctypes.byref(prm), ctypes.byref(st),
                               _mem(bp, bu, bv)), "p2d_advance", allow)
        return phi, u, v, st.as_dict()

    def normals(self, phi):
        """Unit normals (nx, ny) of the level set phi (LEVELSET_2D.h:593-620)."""
        bp = _Buf(phi, self.cell_shape)
        if bp.mem == MEM_HOST:
            a, c = np.empty(self.cell_shape), np.empty(self.cell_shape)
        else:
            a, c = phi.new_empty(phi.shape), phi.new_empty(phi.shape)
        ba, bc = _Buf(a, self.cell_shape, writable=True), _Buf(c, self.cell_shape, writable=True)
        _check(lib.p2d_levelset_normals(self._h, bp.ptr, ba.ptr, bc.ptr, _mem(bp, ba, bc)), "p2d_levelset_normals")
        return a, c

    @property
    def mg_levels(self):
        return lib.p2d_mg_levels(self._h)

    def kernel_timing(self, which, enable=True):
        """Time every launch of kernel `which` inside the solves (HIP events on the context stream)."""
        _check(lib.p2d_kernel_timing(self._h, which, 1 if enable else 0), "p2d_kernel_timing")

    def kernel_time(self, which):
        """(total ms, launches) of the launches timed since kernel_timing(which)."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(lib.p2d_kernel_time(self._h, which, ctypes.byref(ms), ctypes.byref(n)), "p2d_kernel_time")
        return ms.value, n.value

    def time_kernel(self, which, reps=50):
        ms = ctypes.c_double()
        _check(lib.p2d_time_kernel(self._h, which, reps, ctypes.byref(ms)), "p2d_time_kernel")
        return ms.value


VISC_REFERENCE, VISC_DIRICHLET = 0, 1


class ViscositySolver:
    """The semi-implicit viscosity solve (POISSON_SOLVER::SolveForViscosity,
    POISSON_SOLVER.cpp:175-184, 620-712) and the oil-water driver around it
    (VISCOSITY_2D.h:546-640) for an nx x ny cell grid with spacings dx, dy."""

    def __init__(self, nx, ny, dx, dy):
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        h = ctypes.c_void_p()
        _check(lib.p2d_visc_create(nx, ny, dx, dy, ctypes.byref(h)), "p2d_visc_create")
        self._h = h
        self.u_shape, self.v_shape, self.cell_shape = (ny, nx + 1), (ny + 1, nx), (ny, nx)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.p2d_visc_destroy(h)
            self._h = None

    def set_stream(self, stream):
        _check(lib.p2d_visc_set_stream(self._h, stream), "p2d_visc_set_stream")

    def solve(self, component, velocity, c1, c2, c3, c4, c5, explicit_term, boundary=VISC_REFERENCE,
              tolerance=1e-6, max_iterations=10000, allow_not_converged=False):
        """Returns (solution, stats); torch CUDA tensors are solved in place."""
        shape = self.u_shape if component == 0 else self.v_shape
        if not _is_torch(velocity):
            velocity = np.array(velocity, dtype=np.float64, order="C", copy=True)
        bx = _Buf(velocity, shape, writable=True)
        bc = [_Buf(a, shape) for a in (c1, c2, c3, c4, c5, explicit_term)]
        st = P2DStats()
        allow = (P2D_OK, P2D_NOT_CONVERGED) if allow_not_converged else (P2D_OK,)
        _check(lib.p2d_visc_solve(self._h, component, bx.ptr, *[b.ptr for b in bc], boundary, float(tolerance),
                                  int(max_iterations), ctypes.byref(st), _mem(bx, *bc)), "p2d_visc_solve", allow)
        return velocity, st.as_dict()

    def oil_water(self, phi, u, v, dt, allow_not_converged=False, **params):
        """VISCOSITY_2D::ApplyViscosity, oil-water, semi-implicit, dimensionless.  Keyword
        parameters: the fields of p2d_oil_water_params.  Returns (u, v, stats_x, stats_y)."""
        prm = P2DOilWaterParams()
        lib.p2d_oil_water_params_default(ctypes.byref(prm))
        names = {k for k, _ in P2DOilWaterParams._fields_}
        for k, val in params.items():
            if k not in names:
                raise TypeError(f"unknown oil-water parameter {k}")
            setattr(prm, k, val)
        if not _is_torch(u):
            u = np.array(u, dtype=np.float64, order="C", copy=True)
            v = np.array(v, dtype=np.float64, order="C", copy=True)
        bp = _Buf(phi, self.cell_shape)
        bu, bv = _Buf(u, self.u_shape, writable=True), _Buf(v, self.v_shape, writable=True)
        sx, sy = P2DStats(), P2DStats()
        allow = (P2D_OK, P2D_NOT_CONVERGED) if allow_not_converged else (P2D_OK,)
        _check(lib.p2d_visc_oil_water(self._h, bp.ptr, bu.ptr, bv.ptr, float(dt), ctypes.byref(prm), ctypes.byref(sx),
                                      ctypes.byref(sy), _mem(bp, bu, bv)), "p2d_visc_oil_water", allow)
        return u, v, sx.as_dict(), sy.as_dict()
