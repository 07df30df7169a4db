"""ctypes binding of the FP64 CPU oracle (gm_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgm_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class OmMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("num_nodes", C.c_int), ("num_elems", C.c_int),
                ("coor", C.c_void_p), ("connect", C.c_void_p)]


class OmPhysics(C.Structure):
    _fields_ = [("energy", C.c_int), ("visc_model", C.c_int), ("rho", C.c_double),
                ("mu0", C.c_double), ("muinf", C.c_double), ("lam", C.c_double),
                ("nexp", C.c_double), ("aexp", C.c_double), ("Cp", C.c_double),
                ("kcond", C.c_double), ("hsrc", C.c_double), ("g", C.c_double * 3),
                ("beta", C.c_double), ("Tref", C.c_double), ("pspg", C.c_int),
                ("ps_scaling", C.c_double), ("energy_supg", C.c_double)]


class OmBC(C.Structure):
    _fields_ = [("num_dbc", C.c_int), ("node", C.c_void_p), ("var", C.c_void_p),
                ("value", C.c_void_p), ("datum_elem", C.c_int), ("datum_value", C.c_double)]


class OmSolverOpts(C.Structure):
    _fields_ = [("max_newton", C.c_int), ("eps_resid", C.c_double), ("eps_corr", C.c_double),
                ("damp", C.c_double), ("kspace", C.c_int), ("max_its", C.c_int),
                ("lin_tol", C.c_double), ("precond", C.c_int), ("line_search", C.c_int),
                ("min_damp", C.c_double)]


class OmReport(C.Structure):
    _fields_ = [("newton_its", C.c_int), ("converged", C.c_int), ("lin_its_total", C.c_int),
                ("lin_its", C.c_int * 64), ("resid_L2", C.c_double * 64),
                ("resid_Loo", C.c_double * 64), ("unscaled_L2", C.c_double * 64),
                ("t_assemble", C.c_double), ("t_scale", C.c_double), ("t_precond", C.c_double),
                ("t_solve", C.c_double), ("t_total", C.c_double), ("damp", C.c_double * 64)]


class OmlpShape(C.Structure):
    _fields_ = [("n_in", C.c_int), ("n_out", C.c_int), ("hidden", C.c_int), ("width", C.c_int)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, plain C)."""
    srcs = [os.path.join(HERE, f) for f in ("gm_oracle.c", "gm_oracle.h", "gm_mlp_oracle.c",
                                            "gm_mlp_oracle.h", "Makefile")]
    newest = max(os.path.getmtime(s) for s in srcs)
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < newest:
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.om_create.restype = C.c_void_p
        L.om_create.argtypes = [C.POINTER(OmMesh), C.POINTER(OmPhysics), C.POINTER(OmBC)]
        L.om_destroy.argtypes = [C.c_void_p]
        L.om_num_unknowns.argtypes = [C.c_void_p]
        L.om_ija_size.argtypes = [C.c_void_p]
        L.om_copy_ija.argtypes = [C.c_void_p, _ip]
        L.om_copy_first_unknown.argtypes = [C.c_void_p, _ip]
        L.om_copy_unknown_var.argtypes = [C.c_void_p, _ip]
        L.om_copy_dbc.argtypes = [C.c_void_p, _cp, _dp]
        L.om_fill.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.om_fill_range.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int]
        L.om_fill_transient.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp]
        L.om_add_qside.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_double]
        L.om_add_qconv.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_double, C.c_double]
        L.om_element.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.om_row_sum_scale.argtypes = [C.c_int, _dp, _ip, _dp, _dp]
        for nm in ("om_norm_L2", "om_norm_L1", "om_norm_Loo"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [_dp, C.c_int]
        L.om_msr_matvec.argtypes = [C.c_int, _dp, _ip, _dp, _dp]
        L.om_node_colours.argtypes = [C.c_void_p, _ip]
        L.om_colour_permutation.argtypes = [C.c_void_p, _ip, _ip, C.POINTER(C.c_int)]
        L.om_gmres_msr.argtypes = [C.c_int, _dp, _ip, _dp, _dp, C.c_void_p, C.c_int, C.c_int,
                                   C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.om_ilu0_csr.argtypes = [C.c_int, _dp, _ip, C.c_void_p, _ip, _ip, _dp]
        L.om_shape.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.om_gauss.argtypes = [C.c_int, C.c_int, _dp]
        L.om_patch_permutation.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_void_p, C.c_void_p]
        L.om_gmres_perm.argtypes = [C.c_int, _dp, _ip, _dp, _dp, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.om_newton.argtypes = [C.c_void_p, _dp, C.POINTER(OmSolverOpts), C.POINTER(OmReport)]
        for nm in ("om_h_elem_avg", "om_pspg_tau"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [C.c_void_p]
        L.om_velocity_norm.restype = C.c_double
        L.om_velocity_norm.argtypes = [C.c_void_p, _dp]
        L.om_set_velocity_norm.argtypes = [C.c_void_p, C.c_double]
        _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.omlp_num_params.restype = C.c_longlong
        L.omlp_num_params.argtypes = [C.POINTER(OmlpShape)]
        L.omlp_bf16.restype = C.c_float
        L.omlp_bf16.argtypes = [C.c_float]
        L.omlp_forward.argtypes = [C.POINTER(OmlpShape), _fp, _dp, _dp, C.c_int, C.c_longlong, _fp,
                                   _dp]
        L.omlp_loss_grad.argtypes = [C.POINTER(OmlpShape), _fp, _dp, _dp, C.c_int, C.c_longlong,
                                     _fp, _fp, _fp, C.POINTER(C.c_double), _dp]
        _lib = L
    return _lib


class Oracle:
    """One problem instance in the oracle (mirrors the product's context)."""

    def __init__(self, prob):
        L = lib()
        self.prob = prob
        m = prob.mesh
        self._coor = np.ascontiguousarray(m.coor, dtype=np.float64)
        self._conn = np.ascontiguousarray(m.connect, dtype=np.int32)
        self._bn = np.ascontiguousarray(prob.dbc_node, dtype=np.int32)
        self._bv = np.ascontiguousarray(prob.dbc_var, dtype=np.int32)
        self._bx = np.ascontiguousarray(prob.dbc_value, dtype=np.float64)
        om = OmMesh(m.dim, m.num_nodes, m.num_elems, self._coor.ctypes.data, self._conn.ctypes.data)
        ph = prob.phys
        op = OmPhysics(ph.energy, ph.visc_model, ph.rho, ph.mu0, ph.muinf, ph.lam, ph.nexp, ph.aexp,
                       ph.Cp, ph.kcond, ph.hsrc, (C.c_double * 3)(*ph.g), ph.beta, ph.Tref,
                       int(getattr(ph, "pspg", 0)), float(getattr(ph, "ps_scaling", 0.0)),
                       float(getattr(ph, "energy_supg", 0.0)))
        ob = OmBC(len(self._bn), self._bn.ctypes.data, self._bv.ctypes.data, self._bx.ctypes.data,
                  prob.datum_elem, prob.datum_value)
        self.h = L.om_create(C.byref(om), C.byref(op), C.byref(ob))
        self.N = L.om_num_unknowns(self.h)
        self.nval = L.om_ija_size(self.h)
        self.ija = np.zeros(self.nval, np.int32)
        L.om_copy_ija(self.h, self.ija)
        self.first_unknown = np.zeros(m.num_nodes + 1, np.int32)
        L.om_copy_first_unknown(self.h, self.first_unknown)
        self.unk_var = np.zeros(self.N, np.int32)
        L.om_copy_unknown_var(self.h, self.unk_var)
        self.dbc_mask = np.zeros(self.N, np.int8)
        self.dbc_val = np.zeros(self.N)
        L.om_copy_dbc(self.h, self.dbc_mask, self.dbc_val)

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_destroy(self.h)
            self.h = None

    def fill(self, x, e0: int = 0, e1: int | None = None):
        """matrix_fill_full into fresh zeroed arrays: returns (rc, resid, val)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.N)
        v = np.zeros(self.nval)
        e1 = self.prob.mesh.num_elems if e1 is None else e1
        rc = lib().om_fill_range(self.h, x, r, v, e0, e1)
        return rc, r, v

    def h_elem_avg(self) -> float:
        """global_h_elem_siz: the PSPG element size."""
        return lib().om_h_elem_avg(self.h)

    def velocity_norm(self, x) -> float:
        """global_velocity_norm: the mean square of the velocity unknowns."""
        return lib().om_velocity_norm(self.h, np.ascontiguousarray(x, dtype=np.float64))

    def set_velocity_norm(self, U: float):
        """The PSPG velocity norm the next fills use (newton sets it per pass)."""
        lib().om_set_velocity_norm(self.h, float(U))

    def pspg_tau(self) -> float:
        return lib().om_pspg_tau(self.h)

    def add_qside(self, elems, sides, q):
        """QSIDE heat flux q (into the domain) on element sides (Exodus
        numbering, 1-based).  Returns the library's status (0 or -1)."""
        e = np.ascontiguousarray(elems, np.int32)
        s = np.ascontiguousarray(sides, np.int32)
        if e.shape != s.shape:
            raise ValueError("elems and sides differ in length")
        return lib().om_add_qside(self.h, int(e.size), e, s, float(q))

    def add_qconv(self, elems, sides, h, t_bath):
        """QCONV convective heat transfer h (T_bath - T) into the domain on
        element sides (Exodus numbering, 1-based).  Returns the library's
        status (0 or -1)."""
        e = np.ascontiguousarray(elems, np.int32)
        s = np.ascontiguousarray(sides, np.int32)
        if e.shape != s.shape:
            raise ValueError("elems and sides differ in length")
        return lib().om_add_qconv(self.h, int(e.size), e, s, float(h), float(t_bath))

    def fill_transient(self, x, xdot, cmass):
        """matrix_fill_full of a theta-method step: the steady terms plus the
        mass terms for time derivative xdot (d xdot / d x = cmass)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        xd = np.ascontiguousarray(xdot, dtype=np.float64)
        r = np.zeros(self.N)
        v = np.zeros(self.nval)
        rc = lib().om_fill_transient(self.h, x, xd, float(cmass), r, v)
        return rc, r, v

    def element(self, e, x):
        n = C.c_int(0)
        R = np.zeros(200)
        J = np.zeros(200 * 200)
        rc = lib().om_element(self.h, e, np.ascontiguousarray(x, np.float64), R, J, C.byref(n))
        k = n.value
        return rc, R[:k].copy(), J[: k * k].reshape(k, k).copy()

    def colours(self):
        c = np.zeros(self.prob.mesh.num_nodes, np.int32)
        nc = lib().om_node_colours(self.h, c)
        return nc, c

    def colour_permutation(self):
        perm = np.zeros(self.N, np.int32)
        cptr = np.zeros(self.N + 2, np.int32)
        nc = C.c_int(0)
        lib().om_colour_permutation(self.h, perm, cptr, C.byref(nc))
        return perm, cptr[: nc.value + 1].copy()

    def patch_permutation(self, bins):
        perm = np.zeros(self.N, np.int32)
        cptr = np.zeros(self.N + 2, np.int32)
        nc = lib().om_patch_permutation(self.h, bins, perm, cptr, None, None)
        return perm, cptr[: nc + 1].copy()

    def gmres(self, val, b, precond=1, kspace=30, max_its=500, tol=1e-6, x0=None):
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        rr = C.c_double(0.0)
        its = lib().om_gmres_msr(self.N, np.ascontiguousarray(val), self.ija,
                                 np.ascontiguousarray(b, np.float64), x, self.h, precond, kspace,
                                 max_its, tol, C.byref(rr))
        return its, x, rr.value

    def ilu0(self, val, perm=None):
        N = self.N
        nz = self.nval - 1
        rowptr = np.zeros(N + 1, np.int32)
        col = np.zeros(nz, np.int32)
        lu = np.zeros(nz)
        pp = None if perm is None else np.ascontiguousarray(perm, np.int32)
        rc = lib().om_ilu0_csr(N, np.ascontiguousarray(val), self.ija,
                               None if pp is None else pp.ctypes.data, rowptr, col, lu)
        return rc, rowptr, col, lu

    def newton(self, x0=None, max_newton=20, eps_resid=1e-10, eps_corr=1e10, damp=1.0,
               kspace=30, max_its=500, lin_tol=1e-6, precond=1, line_search=0, min_damp=0.05):
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        o = OmSolverOpts(max_newton, eps_resid, eps_corr, damp, kspace, max_its, lin_tol, precond,
                         line_search, min_damp)
        rep = OmReport()
        rc = lib().om_newton(self.h, x, C.byref(o), C.byref(rep))
        return rc, x, rep


def sensitivity(prob_at, x, lam, h):
    """soln_sens (src/mm_sol_nonlinear.c:3638-3930) restated on the oracle:
    dx/dlambda = -J^-1 dR/dlambda at the state x, with dR/dlambda the central
    difference (R(lam + h) - R(lam - h)) / 2h of the residuals of the problems
    prob_at(lam +- h), and J the Jacobian of prob_at(lam), solved exactly
    (SciPy sparse LU).  Test infrastructure only."""
    import scipy.sparse.linalg as spla

    x = np.ascontiguousarray(x, dtype=np.float64)
    rp = Oracle(prob_at(lam + h)).fill(x)[1]
    rm = Oracle(prob_at(lam - h)).fill(x)[1]
    o = Oracle(prob_at(lam))
    rc, r, v = o.fill(x)
    A = msr_to_scipy(o.ija, v, o.N).tocsc()
    return -spla.spsolve(A, (rp - rm) / (2.0 * h))


def msr_to_scipy(ija, val, N):
    """MSR (Goma ams->bindx/val) -> scipy CSR, for direct solves in tests."""
    import scipy.sparse as sp

    rows = [np.arange(N)]
    cols = [np.arange(N)]
    vals = [val[:N]]
    cnt = np.diff(ija[: N + 1])
    rows.append(np.repeat(np.arange(N), cnt))
    cols.append(ija[N + 1: ija[N]])
    vals.append(val[N + 1: ija[N]])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


def shape(dim, s, t, u=0.0):
    n = 9 if dim == 2 else 27
    phi = np.zeros(n)
    dphi = np.zeros(n * 3)
    lib().om_shape(dim, s, t, u, phi, dphi)
    return phi, dphi.reshape(n, 3)[:, :dim].copy()


def gauss(dim):
    nq = 9 if dim == 2 else 27
    out = np.zeros((nq, 4))
    for q in range(nq):
        v = np.zeros(4)
        lib().om_gauss(dim, q, v)
        out[q] = v
    return out


class MlpOracle:
    """The learned-guess network on the CPU (gm_mlp_oracle.h)."""

    def __init__(self, n_in, n_out, hidden, width, in_shift=None, in_scale=None):
        self.shape = OmlpShape(n_in, n_out, hidden, width)
        self.n_in, self.n_out = n_in, n_out
        self.in_shift = np.ascontiguousarray(
            np.zeros(n_in) if in_shift is None else in_shift, dtype=np.float64)
        self.in_scale = np.ascontiguousarray(
            np.ones(n_in) if in_scale is None else in_scale, dtype=np.float64)

    @property
    def num_params(self):
        return int(lib().omlp_num_params(C.byref(self.shape)))

    def forward(self, params, x, quant=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out))
        rc = lib().omlp_forward(C.byref(self.shape), np.ascontiguousarray(params, np.float32),
                                self.in_shift, self.in_scale, int(quant), x.shape[0], x, y)
        assert rc == 0, rc
        return y

    def loss_grad(self, params, x, t, m, quant=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        g = np.zeros(self.num_params)
        loss = C.c_double(0.0)
        rc = lib().omlp_loss_grad(C.byref(self.shape), np.ascontiguousarray(params, np.float32),
                                  self.in_shift, self.in_scale, int(quant), x.shape[0], x,
                                  np.ascontiguousarray(t, np.float32),
                                  np.ascontiguousarray(m, np.float32), C.byref(loss), g)
        assert rc == 0, rc
        return loss.value, g


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)
