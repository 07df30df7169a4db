"""ctypes binding of libgm.so (include/gm_capi.h).

Mirrors Goma's entry points for this path: ``matrix_fill_full``
(src/mm_fill.c:158) and ``solve_nonlinear_problem`` (src/mm_sol_nonlinear.c:596),
on a context built from a :class:`goma_amd.problems.Problem`.

There is no fallback: if libgm.so is missing or cannot be loaded, importing
this module's ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgm.so")

GM_OK = 0
GM_ERR_ELEM = -1
GM_ERR_ARG = -2
GM_ERR_HIP = -3
GM_ERR_PIVOT = -4
GM_ERR_NAN = -5
GM_ERR_CONV = -6
GM_PREC_FP64 = 0
GM_PREC_MIXED = 1
GM_PC_NONE = 0
GM_PC_ILU0 = 1
GM_PC_MG = 2
KERNEL_NAMES = ["k_fill", "k_spmv_msr", "k_ilu_lower", "k_ilu_upper", "k_ilu_factor", "k_mdot",
                "k_orth_update", "k_mg_resid", "k_vanka_update", "k_csr_apply", "k_vanka_invert",
                "k_fill(residual only)", "k_mg_resid_rows", "k_dense_gemv", "k_galerkin",
                "k_elem_scatter", "k_band_panel", "k_band_update", "k_band_solve", "k_vanka_gather",
                "k_row_sum_scale", "k_gjb(coarse inverse)"]


class GmMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("num_nodes", C.c_int), ("num_elems", C.c_int),
                ("coor", C.c_void_p), ("connect", C.c_void_p)]


class GmPhysics(C.Structure):
    _fields_ = [("energy", C.c_int), ("visc_model", C.c_int), ("rho", C.c_double),
                ("mu0", C.c_double), ("muinf", C.c_double), ("lam", C.c_double),
                ("nexp", C.c_double), ("aexp", C.c_double), ("Cp", C.c_double),
                ("kcond", C.c_double), ("hsrc", C.c_double), ("g", C.c_double * 3),
                ("beta", C.c_double), ("Tref", C.c_double), ("pspg", C.c_int),
                ("ps_scaling", C.c_double), ("energy_supg", C.c_double)]


class GmBC(C.Structure):
    _fields_ = [("num_dbc", C.c_int), ("node", C.c_void_p), ("var", C.c_void_p),
                ("value", C.c_void_p), ("datum_elem", C.c_int), ("datum_value", C.c_double)]


class GmSolverOpts(C.Structure):
    _fields_ = [("max_newton", C.c_int), ("eps_resid", C.c_double), ("eps_corr", C.c_double),
                ("damp", C.c_double), ("kspace", C.c_int), ("max_its", C.c_int),
                ("lin_tol", C.c_double), ("precond", C.c_int), ("patch_bins", C.c_int),
                ("precision", C.c_int), ("verbose", C.c_int), ("mg_nu", C.c_int),
                ("mg_omega", C.c_double), ("mg_coarse_sweeps", C.c_int), ("line_search", C.c_int),
                ("min_damp", C.c_double), ("pc_reuse", C.c_int), ("mg_smoother", C.c_int),
                ("mg_galerkin", C.c_int), ("lin_tol_adapt", C.c_int)]


MAX_REPORT = 64


class GmNewtonReport(C.Structure):
    _fields_ = [("newton_its", C.c_int), ("converged", C.c_int), ("lin_its_total", C.c_int),
                ("lin_its", C.c_int * MAX_REPORT), ("resid_L2", C.c_double * MAX_REPORT),
                ("resid_Loo", C.c_double * MAX_REPORT), ("unscaled_L2", C.c_double * MAX_REPORT),
                ("lin_relres", C.c_double * MAX_REPORT), ("damp", C.c_double * MAX_REPORT),
                ("t_assemble", C.c_double),
                ("t_scale", C.c_double), ("t_precond", C.c_double), ("t_solve", C.c_double),
                ("t_update", C.c_double), ("t_total", C.c_double), ("fell_back", C.c_int),
                ("discarded_its", C.c_int), ("pc_builds", C.c_int), ("resid_fills", C.c_int)]

    def summary(self) -> dict:
        n = min(self.newton_its, MAX_REPORT)
        return dict(newton_its=self.newton_its, converged=bool(self.converged),
                    lin_its_total=self.lin_its_total, lin_its=list(self.lin_its[:n]),
                    resid_L2=list(self.resid_L2[:n]), resid_Loo=list(self.resid_Loo[:n]),
                    unscaled_L2=list(self.unscaled_L2[:n]), damp=list(self.damp[:n]),
                    t_assemble=self.t_assemble,
                    t_scale=self.t_scale, t_precond=self.t_precond, t_solve=self.t_solve,
                    t_update=self.t_update, t_total=self.t_total, fell_back=bool(self.fell_back),
                    discarded_its=self.discarded_its, pc_builds=self.pc_builds,
                    resid_fills=self.resid_fills)


class GmTransientOpts(C.Structure):
    _fields_ = [("delta_t", C.c_double), ("theta", C.c_double), ("num_steps", C.c_int),
                ("startup_steps", C.c_int)]


class GmTransientReport(C.Structure):
    _fields_ = [("steps", C.c_int), ("time", C.c_double), ("newton_its_total", C.c_int),
                ("lin_its_total", C.c_int), ("newton_its", C.c_int * MAX_REPORT),
                ("t_total", C.c_double)]


MLP_MAX_IN = 8
MLP_MAX_OUT = 8


class GmMlpShape(C.Structure):
    _fields_ = [("n_in", C.c_int), ("n_out", C.c_int), ("hidden", C.c_int), ("width", C.c_int),
                ("seed", C.c_ulonglong), ("in_shift", C.c_double * MLP_MAX_IN),
                ("in_scale", C.c_double * MLP_MAX_IN)]


class GmMlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_double),
                ("lr_final", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("seed", C.c_ulonglong), ("verbose", C.c_int)]


class GmMlpQuery(C.Structure):
    _fields_ = [("param", C.c_double * MLP_MAX_IN), ("out_scale", C.c_double * MLP_MAX_OUT),
                ("out_shift", C.c_double * MLP_MAX_OUT)]


# every exported symbol of include/gm_capi.h
EXPORTS = [
    "gm_ctx_create", "gm_ctx_destroy", "gm_last_error", "gm_num_unknowns", "gm_msr_size",
    "gm_copy_msr_graph", "gm_copy_first_unknown", "gm_copy_dirichlet", "gm_num_elem_colours",
    "gm_copy_elem_colours", "gm_copy_patch_permutation", "gm_set_physics",
    "gm_matrix_fill_full_dev", "gm_matrix_fill_full", "gm_matrix_fill_full32_dev",
    "gm_row_sum_scale_dev", "gm_msr_matvec_dev",
    "gm_gmres_dev", "gm_ilu0_factor_dev", "gm_solve_nonlinear_dev", "gm_solve_nonlinear",
    "gm_default_solver_opts", "gm_prof_enable", "gm_prof_get", "gm_mg_add_level",
    "gm_mg_num_levels", "gm_mg_setup_dev", "gm_mg_copy_level_matrix", "gm_solve_nonlinear_guess_dev",
    "gm_mlp_create", "gm_mlp_destroy", "gm_mlp_num_params", "gm_mlp_get_params",
    "gm_mlp_set_params", "gm_mlp_default_train_opts", "gm_mlp_set_data", "gm_mlp_train",
    "gm_mlp_loss_grad", "gm_mlp_predict_dev", "gm_mlp_predict", "gm_mlp_guess_dev",
    "gm_mlp_set_basis", "gm_mlp_clear_basis", "gm_matrix_fill_full_transient_dev",
    "gm_add_qside", "gm_add_qconv", "gm_clear_qside", "gm_solve_nonlinear_transient_dev", "gm_solve_transient_dev",
    "gm_num_nodes", "gm_num_elems", "gm_mg_build", "gm_mg_level", "gm_mg_copy_prolongation",
    "gm_matrix_fill_full_assign_dev", "gm_mlp_predict_noalloc_dev", "gm_mg_copy_block_inverse",
    "gm_mg_copy_coarse_inverse",
    "gm_h_elem_avg", "gm_velocity_norm_dev", "gm_set_velocity_norm",
]

_lib = None


def lib():
    """Load libgm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with goma_amd/build.py "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, ip = C.c_void_p, C.c_int
    L.gm_last_error.restype = C.c_char_p
    L.gm_ctx_create.argtypes = [C.POINTER(GmMesh), C.POINTER(GmPhysics), C.POINTER(GmBC),
                                C.POINTER(vp)]
    L.gm_ctx_destroy.argtypes = [vp]
    for f in ("gm_num_unknowns", "gm_msr_size", "gm_num_elem_colours", "gm_num_nodes",
              "gm_num_elems"):
        getattr(L, f).argtypes = [vp]
    for f in ("gm_copy_msr_graph", "gm_copy_first_unknown", "gm_copy_elem_colours"):
        getattr(L, f).argtypes = [vp, vp]
    L.gm_copy_dirichlet.argtypes = [vp, vp, vp]
    L.gm_copy_patch_permutation.argtypes = [vp, ip, vp, vp]
    L.gm_set_physics.argtypes = [vp, C.POINTER(GmPhysics)]
    L.gm_h_elem_avg.argtypes = [vp, C.POINTER(C.c_double)]
    L.gm_velocity_norm_dev.argtypes = [vp, vp, C.POINTER(C.c_double)]
    L.gm_set_velocity_norm.argtypes = [vp, C.c_double]
    L.gm_matrix_fill_full_dev.argtypes = [vp, vp, vp, vp]
    L.gm_matrix_fill_full.argtypes = [vp, vp, vp, vp]
    L.gm_matrix_fill_full32_dev.argtypes = [vp, vp, vp, vp]
    L.gm_matrix_fill_full_assign_dev.argtypes = [vp, vp, vp, vp]
    L.gm_row_sum_scale_dev.argtypes = [vp, vp, vp, vp]
    L.gm_msr_matvec_dev.argtypes = [vp, vp, vp, vp]
    L.gm_gmres_dev.argtypes = [vp, vp, vp, vp, C.POINTER(GmSolverOpts), C.POINTER(ip),
                               C.POINTER(C.c_double)]
    L.gm_ilu0_factor_dev.argtypes = [vp, vp, ip, vp, vp, vp]
    L.gm_solve_nonlinear_dev.argtypes = [vp, vp, C.POINTER(GmSolverOpts),
                                         C.POINTER(GmNewtonReport)]
    L.gm_solve_nonlinear.argtypes = [vp, vp, C.POINTER(GmSolverOpts), C.POINTER(GmNewtonReport)]
    L.gm_default_solver_opts.argtypes = [C.POINTER(GmSolverOpts)]
    L.gm_default_solver_opts.restype = None
    L.gm_mg_add_level.argtypes = [vp, vp, vp, vp, vp, vp]
    L.gm_mg_num_levels.argtypes = [vp]
    L.gm_mg_build.argtypes = [vp, ip]
    L.gm_mg_level.argtypes = [vp, ip]
    L.gm_mg_level.restype = vp
    L.gm_mg_copy_prolongation.argtypes = [vp, ip, vp, vp, vp]
    L.gm_mg_setup_dev.argtypes = [vp, vp, vp, vp]
    L.gm_mg_copy_level_matrix.argtypes = [vp, ip, vp]
    L.gm_mg_copy_block_inverse.argtypes = [vp, ip, ip, vp]
    L.gm_mg_copy_coarse_inverse.argtypes = [vp, vp]
    L.gm_prof_enable.argtypes = [vp, ip]
    L.gm_prof_get.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.gm_solve_nonlinear_guess_dev.argtypes = [vp, vp, vp, C.POINTER(GmSolverOpts), ip,
                                               C.c_double, C.POINTER(GmNewtonReport)]
    L.gm_matrix_fill_full_transient_dev.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, vp, vp]
    L.gm_add_qside.argtypes = [vp, ip, vp, vp, C.c_double]
    L.gm_add_qconv.argtypes = [vp, ip, vp, vp, C.c_double, C.c_double]
    L.gm_clear_qside.argtypes = [vp]
    L.gm_solve_nonlinear_transient_dev.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double,
                                                   C.POINTER(GmSolverOpts),
                                                   C.POINTER(GmNewtonReport)]
    L.gm_solve_transient_dev.argtypes = [vp, vp, vp, C.POINTER(GmTransientOpts),
                                         C.POINTER(GmSolverOpts), C.POINTER(GmTransientReport)]
    L.gm_mlp_create.argtypes = [C.POINTER(GmMlpShape), C.POINTER(vp)]
    L.gm_mlp_destroy.argtypes = [vp]
    L.gm_mlp_destroy.restype = None
    L.gm_mlp_num_params.argtypes = [vp]
    L.gm_mlp_num_params.restype = C.c_longlong
    L.gm_mlp_get_params.argtypes = [vp, vp]
    L.gm_mlp_set_params.argtypes = [vp, vp]
    L.gm_mlp_default_train_opts.argtypes = [C.POINTER(GmMlpTrainOpts)]
    L.gm_mlp_default_train_opts.restype = None
    L.gm_mlp_set_data.argtypes = [vp, C.c_longlong, vp, vp, vp]
    L.gm_mlp_train.argtypes = [vp, C.POINTER(GmMlpTrainOpts), vp]
    L.gm_mlp_loss_grad.argtypes = [vp, ip, vp, vp, vp, C.POINTER(C.c_double), vp]
    L.gm_mlp_predict_dev.argtypes = [vp, C.c_longlong, vp, vp]
    L.gm_mlp_predict.argtypes = [vp, C.c_longlong, vp, vp]
    L.gm_mlp_predict_noalloc_dev.argtypes = [vp, C.c_longlong, vp, vp]
    L.gm_mlp_guess_dev.argtypes = [vp, vp, C.POINTER(GmMlpQuery), C.POINTER(GmMlpQuery), vp, vp]
    L.gm_mlp_set_basis.argtypes = [vp, vp, ip, vp, vp]
    L.gm_mlp_clear_basis.argtypes = [vp]
    L.gm_mlp_clear_basis.restype = None
    _lib = L
    return L


class GomaError(RuntimeError):
    pass


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().gm_last_error().decode(errors="replace")
        raise GomaError(f"{what} failed ({rc}): {msg}")
    return rc


def solver_opts(**kw) -> GmSolverOpts:
    o = GmSolverOpts()
    lib().gm_default_solver_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def physics_struct(ph) -> GmPhysics:
    return GmPhysics(ph.energy, ph.visc_model, ph.rho, ph.mu0, ph.muinf, ph.lam, ph.nexp, ph.aexp,
                     ph.Cp, ph.kcond, ph.hsrc, (C.c_double * 3)(*ph.g), ph.beta, ph.Tref,
                     int(getattr(ph, "pspg", 0)), float(getattr(ph, "ps_scaling", 0.0)),
                     float(getattr(ph, "energy_supg", 0.0)))


class Context:
    """A gm_ctx for one problem.  Device-pointer methods take integer addresses
    (e.g. ``tensor.data_ptr()`` of a torch CUDA tensor)."""

    def __init__(self, prob):
        L = lib()
        self.prob = prob
        m = prob.mesh
        self._coor = np.ascontiguousarray(m.coor, dtype=np.float64)
        self._conn = np.ascontiguousarray(m.connect, dtype=np.int32)
        self._bn = np.ascontiguousarray(prob.dbc_node, dtype=np.int32)
        self._bv = np.ascontiguousarray(prob.dbc_var, dtype=np.int32)
        self._bx = np.ascontiguousarray(prob.dbc_value, dtype=np.float64)
        gm = GmMesh(m.dim, m.num_nodes, m.num_elems, self._coor.ctypes.data, self._conn.ctypes.data)
        ph = physics_struct(prob.phys)
        bc = GmBC(len(self._bn), self._bn.ctypes.data, self._bv.ctypes.data, self._bx.ctypes.data,
                  prob.datum_elem, prob.datum_value)
        h = C.c_void_p()
        _check(L.gm_ctx_create(C.byref(gm), C.byref(ph), C.byref(bc), C.byref(h)), "gm_ctx_create")
        self.h = h
        self.N = L.gm_num_unknowns(h)
        self.nval = L.gm_msr_size(h)

    def close(self):
        # the level contexts belong to this one and go with it
        for cc in getattr(self, "_coarse", []):
            cc.h = None
        self._coarse = []
        if getattr(self, "h", None):
            lib().gm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- host-side views of the set-up
    def bindx(self) -> np.ndarray:
        out = np.zeros(self.nval, np.int32)
        _check(lib().gm_copy_msr_graph(self.h, out.ctypes.data), "gm_copy_msr_graph")
        return out

    def first_unknown(self) -> np.ndarray:
        out = np.zeros(lib().gm_num_nodes(self.h) + 1, np.int32)
        _check(lib().gm_copy_first_unknown(self.h, out.ctypes.data), "gm_copy_first_unknown")
        return out

    def dirichlet(self):
        mask = np.zeros(self.N, np.int8)
        val = np.zeros(self.N)
        _check(lib().gm_copy_dirichlet(self.h, mask.ctypes.data, val.ctypes.data), "gm_copy_dirichlet")
        return mask, val

    def elem_colours(self) -> np.ndarray:
        out = np.zeros(lib().gm_num_elems(self.h), np.int32)
        _check(lib().gm_copy_elem_colours(self.h, out.ctypes.data), "gm_copy_elem_colours")
        return out

    def patch_permutation(self, bins: int = 0):
        perm = np.zeros(self.N, np.int32)
        cptr = np.zeros(4097, np.int32)
        nc = _check(lib().gm_copy_patch_permutation(self.h, bins, perm.ctypes.data, cptr.ctypes.data),
                    "gm_copy_patch_permutation")
        return perm, cptr[: nc + 1].copy()

    def set_physics(self, phys):
        ph = physics_struct(phys)
        _check(lib().gm_set_physics(self.h, C.byref(ph)), "gm_set_physics")

    def h_elem_avg(self) -> float:
        """global_h_elem_siz of the mesh (the PSPG element size)."""
        h = C.c_double()
        _check(lib().gm_h_elem_avg(self.h, C.byref(h)), "gm_h_elem_avg")
        return h.value

    def velocity_norm_dev(self, x_ptr: int) -> float:
        """global_velocity_norm of a device state: mean square of the velocities."""
        u = C.c_double()
        _check(lib().gm_velocity_norm_dev(self.h, x_ptr, C.byref(u)), "gm_velocity_norm_dev")
        return u.value

    def set_velocity_norm(self, U: float):
        """The PSPG velocity norm the assembly entry points use (the Newton
        drivers set it from x at every pass)."""
        _check(lib().gm_set_velocity_norm(self.h, float(U)), "gm_set_velocity_norm")

    # ---- compute (device pointers)
    def fill_dev(self, x_ptr: int, resid_ptr: int, val_ptr: int) -> int:
        return lib().gm_matrix_fill_full_dev(self.h, x_ptr, resid_ptr, val_ptr)

    def fill_assign_dev(self, x_ptr: int, resid_ptr: int, val_ptr: int) -> int:
        """The assembly with the Jacobian stored, not accumulated (val needs no zeroing)."""
        return lib().gm_matrix_fill_full_assign_dev(self.h, x_ptr, resid_ptr, val_ptr)

    def fill32_dev(self, x_ptr: int, resid_ptr: int, val32_ptr: int) -> int:
        """The assembly with FP32 Jacobian values (mixed precision)."""
        return lib().gm_matrix_fill_full32_dev(self.h, x_ptr, resid_ptr, val32_ptr)

    def fill_host(self, x: np.ndarray, resid: np.ndarray, val: np.ndarray) -> int:
        for a in (x, resid, val):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        return lib().gm_matrix_fill_full(self.h, x.ctypes.data, resid.ctypes.data, val.ctypes.data)

    def fill_transient_dev(self, x_ptr, x_old_ptr, xdot_old_ptr, delta_t, theta, resid_ptr,
                           val_ptr) -> int:
        """The assembly of one theta-method time step (transient mass terms)."""
        return lib().gm_matrix_fill_full_transient_dev(self.h, x_ptr, x_old_ptr, xdot_old_ptr,
                                                       delta_t, theta, resid_ptr, val_ptr)

    def add_qside(self, elems, sides, q: float):
        """QSIDE weak heat flux q on the given (element, 1-based side) pairs."""
        e = np.ascontiguousarray(elems, dtype=np.int32)
        sd = np.ascontiguousarray(sides, dtype=np.int32)
        assert e.shape == sd.shape
        _check(lib().gm_add_qside(self.h, len(e), e.ctypes.data, sd.ctypes.data, float(q)),
               "gm_add_qside")

    def add_qconv(self, elems, sides, h: float, t_bath: float):
        """QCONV convective heat transfer h (t_bath - T) on the given (element,
        1-based side) pairs."""
        e = np.ascontiguousarray(elems, dtype=np.int32)
        sd = np.ascontiguousarray(sides, dtype=np.int32)
        assert e.shape == sd.shape
        _check(lib().gm_add_qconv(self.h, len(e), e.ctypes.data, sd.ctypes.data, float(h),
                                  float(t_bath)), "gm_add_qconv")

    def clear_qside(self):
        _check(lib().gm_clear_qside(self.h), "gm_clear_qside")

    def solve_transient_step_dev(self, x_ptr, x_old_ptr, xdot_old_ptr, delta_t, theta, **opts):
        """Newton for one theta-method time step; x holds the prediction."""
        o = solver_opts(**opts)
        rep = GmNewtonReport()
        rc = lib().gm_solve_nonlinear_transient_dev(self.h, x_ptr, x_old_ptr, xdot_old_ptr,
                                                    delta_t, theta, C.byref(o), C.byref(rep))
        _check(rc, "gm_solve_nonlinear_transient_dev")
        return rc, rep

    def solve_transient_dev(self, x_ptr, xdot_ptr, delta_t, theta, num_steps, startup_steps=3,
                            **opts):
        """The fixed-step time loop: x (initial condition -> final solution) and
        xdot (initial derivative -> final derivative) are device arrays."""
        t = GmTransientOpts(delta_t, theta, num_steps, startup_steps)
        o = solver_opts(**opts)
        rep = GmTransientReport()
        rc = lib().gm_solve_transient_dev(self.h, x_ptr, xdot_ptr, C.byref(t), C.byref(o),
                                          C.byref(rep))
        _check(rc, "gm_solve_transient_dev")
        return rc, rep

    def row_sum_scale_dev(self, val_ptr, b_ptr, scale_ptr):
        _check(lib().gm_row_sum_scale_dev(self.h, val_ptr, b_ptr, scale_ptr), "gm_row_sum_scale_dev")

    def matvec_dev(self, val_ptr, x_ptr, y_ptr):
        _check(lib().gm_msr_matvec_dev(self.h, val_ptr, x_ptr, y_ptr), "gm_msr_matvec_dev")

    def gmres_dev(self, val_ptr, b_ptr, x_ptr, **opts):
        o = solver_opts(**opts)
        its = C.c_int(0)
        rr = C.c_double(0.0)
        _check(lib().gm_gmres_dev(self.h, val_ptr, b_ptr, x_ptr, C.byref(o), C.byref(its),
                                  C.byref(rr)), "gm_gmres_dev")
        return its.value, rr.value

    def ilu0_factor_dev(self, val_ptr, bins: int = 0):
        nnz = self.nval - 1
        rowptr = np.zeros(self.N + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        lu = np.zeros(nnz)
        _check(lib().gm_ilu0_factor_dev(self.h, val_ptr, bins, rowptr.ctypes.data, col.ctypes.data,
                                        lu.ctypes.data), "gm_ilu0_factor_dev")
        return rowptr, col, lu

    def solve_dev(self, x_ptr: int, **opts):
        o = solver_opts(**opts)
        rep = GmNewtonReport()
        rc = lib().gm_solve_nonlinear_dev(self.h, x_ptr, C.byref(o), C.byref(rep))
        _check(rc, "gm_solve_nonlinear_dev")
        return rc, rep

    def solve_guess_dev(self, x_ptr: int, x_prev_ptr: int, watch_its: int = 3, drop: float = 10.0,
                        **opts):
        """Newton from the guess in x, falling back to x_prev when the scaled
        residual has not dropped by ``drop`` within ``watch_its`` passes."""
        o = solver_opts(**opts)
        rep = GmNewtonReport()
        rc = lib().gm_solve_nonlinear_guess_dev(self.h, x_ptr, x_prev_ptr, C.byref(o), watch_its,
                                                drop, C.byref(rep))
        _check(rc, "gm_solve_nonlinear_guess_dev")
        return rc, rep

    def solve_host(self, x: np.ndarray, **opts):
        o = solver_opts(**opts)
        rep = GmNewtonReport()
        rc = lib().gm_solve_nonlinear(self.h, x.ctypes.data, C.byref(o), C.byref(rep))
        _check(rc, "gm_solve_nonlinear")
        return rc, rep

    # ---- multigrid hierarchy
    def attach_multigrid(self, min_elems_per_side: int = 2):
        """Builds the multigrid hierarchy in the library (gm_mg_build): coarse
        contexts down to ``min_elems_per_side`` elements per side with their
        Q2/P1 prolongations.  Returns the number of levels."""
        if self.prob.mesh.n < 1:
            raise ValueError("geometric multigrid needs the structured mesh hierarchy; this mesh "
                             "has none (read from a file or renumbered, mesh.n == 0): use "
                             "ILU(0) or the banded LU")
        if getattr(self, "_coarse", None):
            # the library keeps pointers to the attached levels: dropping them here
            # would leave it with freed contexts
            raise RuntimeError("a multigrid hierarchy is already attached to this context")
        nl = _check(lib().gm_mg_build(self.h, min_elems_per_side), "gm_mg_build")
        self._coarse = [LevelContext(lib().gm_mg_level(self.h, k)) for k in range(1, nl)]
        return nl

    def mg_prolongation(self, level: int):
        """CSR (rowptr, col, val) of the prolongation into level ``level`` - 1
        from level ``level`` (>= 1) of the attached hierarchy."""
        L = lib()
        above = self if level == 1 else self._coarse[level - 2]
        nnz = _check(L.gm_mg_copy_prolongation(self.h, level, None, None, None),
                     "gm_mg_copy_prolongation")
        rp = np.zeros(above.N + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        cv = np.zeros(nnz)
        _check(L.gm_mg_copy_prolongation(self.h, level, rp.ctypes.data, ci.ctypes.data,
                                         cv.ctypes.data), "gm_mg_copy_prolongation")
        return rp, ci, cv

    def mg_level_matrix(self, level: int) -> np.ndarray:
        """FP64 host copy of the Jacobian of multigrid level ``level`` (1 = first
        coarse level) from the last set-up, in that level's MSR layout."""
        out = np.zeros(self._coarse[level - 1].nval)
        _check(lib().gm_mg_copy_level_matrix(self.h, level, out.ctypes.data), "gm_mg_copy_level_matrix")
        return out

    def mg_block_inverse(self, level: int, elem: int) -> np.ndarray:
        """(A_e^-1) of element ``elem`` on multigrid level ``level`` (0 = this
        mesh) from the last set-up, as an n x n array."""
        out = np.zeros(128 * 128)  # the largest block: 112 unknowns (HEX27 with energy)
        n = _check(lib().gm_mg_copy_block_inverse(self.h, level, elem, out.ctypes.data),
                   "gm_mg_copy_block_inverse")
        return out[: n * n].reshape(n, n).T.copy()

    def mg_coarse_inverse(self) -> np.ndarray:
        """Dense inverse of the coarsest level's Jacobian from the last set-up
        with an exact coarse solve, as an n x n array."""
        n = self._coarse[-1].N
        out = np.zeros(n * n)
        m = _check(lib().gm_mg_copy_coarse_inverse(self.h, out.ctypes.data), "gm_mg_copy_coarse_inverse")
        assert m == n
        return out.reshape(n, n)

    def mg_setup_dev(self, x_ptr, val_ptr, scale_ptr=None):
        _check(lib().gm_mg_setup_dev(self.h, x_ptr, val_ptr, scale_ptr), "gm_mg_setup_dev")

    # ---- kernel timing
    def prof_enable(self, on=True, every: int = 1):
        """Time every ``every``-th launch of each kernel (0/False: off)."""
        _check(lib().gm_prof_enable(self.h, (every if every > 1 else 1) if on else 0),
               "gm_prof_enable")

    def prof_get(self, kid: int):
        """(launches, total ms, total algorithmic bytes) of one kernel family."""
        cnt = C.c_int(0)
        ms = C.c_double(0.0)
        by = C.c_double(0.0)
        _check(lib().gm_prof_get(self.h, kid, C.byref(cnt), C.byref(ms), C.byref(by)), "gm_prof_get")
        return cnt.value, ms.value, by.value

    def prof_all(self) -> dict:
        out = {}
        for i, nm in enumerate(KERNEL_NAMES):
            cnt, ms, by = self.prof_get(i)
            if cnt:
                out[nm] = dict(launches=cnt, total_ms=ms, avg_ms=ms / cnt,
                               bytes_per_launch=by / cnt if by else None,
                               GBps=(by / (ms * 1e-3) / 1e9) if (by and ms) else None)
        return out


class LevelContext(Context):
    """A multigrid level's context, owned by the finest level's Context:
    the query methods work on it; it is released with its owner."""

    def __init__(self, handle):  # noqa: D107 - borrowed handle, no set-up
        self.h = C.c_void_p(handle)
        self.prob = None
        self.N = lib().gm_num_unknowns(self.h)
        self.nval = lib().gm_msr_size(self.h)

    def close(self):
        self.h = None


def mlp_query(param, out_scale, out_shift) -> GmMlpQuery:
    q = GmMlpQuery()
    for i, v in enumerate(param):
        q.param[i] = float(v)
    for i, v in enumerate(out_scale):
        q.out_scale[i] = float(v)
    for i, v in enumerate(out_shift):
        q.out_shift[i] = float(v)
    return q


class Mlp:
    """A gm_mlp: the learned-initial-guess network on the GPU."""

    def __init__(self, n_in, n_out, hidden=4, width=256, seed=20261202, in_shift=None,
                 in_scale=None):
        s = GmMlpShape()
        s.n_in, s.n_out, s.hidden, s.width, s.seed = n_in, n_out, hidden, width, seed
        for i in range(n_in):
            s.in_shift[i] = 0.0 if in_shift is None else float(in_shift[i])
            s.in_scale[i] = 1.0 if in_scale is None else float(in_scale[i])
        self.shape = s
        self.n_in, self.n_out = n_in, n_out
        h = C.c_void_p()
        _check(lib().gm_mlp_create(C.byref(s), C.byref(h)), "gm_mlp_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(lib().gm_mlp_num_params(self.h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        _check(lib().gm_mlp_get_params(self.h, p.ctypes.data), "gm_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().gm_mlp_set_params(self.h, p.ctypes.data), "gm_mlp_set_params")

    def set_data(self, x, t, m):
        self._data = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, m)]
        x, t, m = self._data
        _check(lib().gm_mlp_set_data(self.h, x.shape[0], x.ctypes.data, t.ctypes.data,
                                     m.ctypes.data), "gm_mlp_set_data")
        self._data = None

    def train(self, **kw) -> np.ndarray:
        o = GmMlpTrainOpts()
        lib().gm_mlp_default_train_opts(C.byref(o))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        hist = np.zeros(max(o.epochs, 1))
        _check(lib().gm_mlp_train(self.h, C.byref(o), hist.ctypes.data), "gm_mlp_train")
        return hist[: o.epochs]

    def loss_grad(self, x, t, m):
        x, t, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, m))
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_double(0.0)
        _check(lib().gm_mlp_loss_grad(self.h, x.shape[0], x.ctypes.data, t.ctypes.data,
                                      m.ctypes.data, C.byref(loss), g.ctypes.data),
               "gm_mlp_loss_grad")
        return loss.value, g

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        _check(lib().gm_mlp_predict(self.h, x.shape[0], x.ctypes.data, y.ctypes.data),
               "gm_mlp_predict")
        return y

    def guess_dev(self, ctx: Context, query: GmMlpQuery, x_ptr: int, prev_query=None,
                  x_prev_ptr=None):
        _check(lib().gm_mlp_guess_dev(self.h, ctx.h, C.byref(query),
                                      C.byref(prev_query) if prev_query is not None else None,
                                      x_prev_ptr, x_ptr), "gm_mlp_guess_dev")

    def set_basis(self, ctx: Context, snaps_ptr: int, nsnap: int, weight_ptr: int) -> int:
        """Snapshot subspace for the delta-form guess on ctx (device arrays:
        snapshots [nsnap][N], weights [N]); returns its dimension."""
        return _check(lib().gm_mlp_set_basis(self.h, ctx.h, nsnap, snaps_ptr, weight_ptr),
                      "gm_mlp_set_basis")

    def clear_basis(self):
        lib().gm_mlp_clear_basis(self.h)
