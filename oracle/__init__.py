"""ctypes wrapper of the FP64 CPU oracle (oracle/fem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg
of bench.py, always as the checker or the timed CPU baseline. The shipped library never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfem_oracle.so")
_lib = None

TET_LINEAR, TET_COROTATIONAL, HEX = 0, 1, 2
EULER_IMPLICIT, LINEAR_EULER_IMPLICIT = 0, 1
STATIC, LINEAR_STATIC = 2, 3
EULER_EXPLICIT, LINEAR_EULER_EXPLICIT = 4, 5
EULER_EXPLICIT_MODIFIED, LINEAR_EULER_EXPLICIT_MODIFIED = 6, 7
RUNGE_KUTTA_4, LINEAR_RUNGE_KUTTA_4 = 8, 9
LU, CG = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_rotation_scaling.argtypes = [_dp, _dp, _dp]
        L.orc_tet_linear_km.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_cube_km.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_coro_element.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       _dp, _dp, _dp, _dp, _ip]
        L.orc_coro_element_given_rotation.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                      _dp, _dp, _dp, _dp, _ip]
        L.orc_ode_complex_nonlinear.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_double, _dp,
                                                ctypes.POINTER(ctypes.c_long)]
        L.orc_sim_last_newton_iters.restype = ctypes.c_long
        L.orc_sim_last_newton_iters.argtypes = [ctypes.c_void_p]
        L.orc_sim_set_newton_eps.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_sim_set_compliance_at_first_linear_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_compliance_matrix.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_create.restype = ctypes.c_void_p
        L.orc_sim_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, _dp]
        L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sim_add_bc_dof.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_set_gravity.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int]
        L.orc_sim_set_external_force.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_set_damping.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_sim_add_localized_force.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_sim_update_fmdk.argtypes = [ctypes.c_void_p]
        L.orc_sim_apply_compliance.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.orc_sim_set_solver.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.orc_sim_set_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_sim_get_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_sim_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.orc_sim_adjoint_step.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_sim_set_young.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_nnz.restype = ctypes.c_long
        L.orc_sim_nnz.argtypes = [ctypes.c_void_p]
        L.orc_sim_pattern.argtypes = [ctypes.c_void_p, _ip, _ip]
        L.orc_sim_system_values.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_matrix_values.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_sim_rhs.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_force.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_mass_per_node.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_last_cg_iters.restype = ctypes.c_long
        L.orc_sim_last_cg_iters.argtypes = [ctypes.c_void_p]
        L.orc_sim_last_cg_error.restype = ctypes.c_double
        L.orc_sim_last_cg_error.argtypes = [ctypes.c_void_p]
        L.orc_sim_total_cg_iters.restype = ctypes.c_long
        L.orc_sim_total_cg_iters.argtypes = [ctypes.c_void_p]
        L.orc_sim_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
        L.orc_sim_valid.argtypes = [ctypes.c_void_p]
        L.orc_sim_element.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_cg_solve.restype = ctypes.c_long
        L.orc_cg_solve.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, ctypes.c_double, ctypes.c_long, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def rotation_scaling(F):
    F = _f64(F).reshape(3, 3)
    R = np.zeros((3, 3))
    S = np.zeros((3, 3))
    lib().orc_rotation_scaling(_d(F), _d(R), _d(S))
    return R, S


def tet_linear_km(x0, E, nu, rho):
    x0 = _f64(x0).reshape(12)
    K = np.zeros((12, 12))
    M = np.zeros((12, 12))
    vol = np.zeros(1)
    _check(lib().orc_tet_linear_km(_d(x0), E, nu, rho, _d(K), _d(M), _d(vol)))
    return K, M, vol[0]


def cube_km(x0, E, nu, rho):
    x0 = _f64(x0).reshape(24)
    K = np.zeros((24, 24))
    M = np.zeros((24, 24))
    vol = np.zeros(1)
    _check(lib().orc_cube_km(_d(x0), E, nu, rho, _d(K), _d(M), _d(vol)))
    return K, M, vol[0]


def coro_element(x0, x, E, nu, rho):
    """Initialize a corotational tet at x0 and update it at x. Returns R, M, K, f, flags."""
    x0 = _f64(x0).reshape(12)
    x = _f64(x).reshape(12)
    R = np.zeros((3, 3))
    M = np.zeros((12, 12))
    K = np.zeros((12, 12))
    f = np.zeros(12)
    flags = ctypes.c_int(0)
    _check(lib().orc_coro_element(_d(x0), _d(x), E, nu, rho, _d(R), _d(M), _d(K), _d(f), ctypes.byref(flags)))
    return R, M, K, f, flags.value


def coro_element_given_rotation(x0, x, E, nu, rho, R):
    """coro_element with the polar factor R supplied (for configurations where it is not unique). Returns
    M, K, f, flags."""
    x0 = _f64(x0).reshape(12)
    x = _f64(x).reshape(12)
    R = np.ascontiguousarray(_f64(R).reshape(9))
    M = np.zeros((12, 12))
    K = np.zeros((12, 12))
    f = np.zeros(12)
    flags = ctypes.c_int(0)
    _check(lib().orc_coro_element_given_rotation(_d(x0), _d(x), E, nu, rho, _d(R), _d(M), _d(K), _d(f),
                                                 ctypes.byref(flags)))
    return M, K, f, flags.value


def ode_complex_nonlinear(linear_variant, max_iter, eps):
    """Two steps of the reference's OdeComplexNonLinear test equation (a = x |v|^2) through the oracle's
    Newton-Raphson loop (OdeSolverEulerImplicit) or its linear variant; dict of x1, v1, x2, v2, f2 = f(x2, v2)
    and the number of linear solves."""
    out = np.zeros(15)
    solves = ctypes.c_long(0)
    _check(lib().orc_ode_complex_nonlinear(int(linear_variant), int(max_iter), float(eps), _d(out),
                                           ctypes.byref(solves)))
    o = out.reshape(5, 3)
    return dict(x1=o[0], v1=o[1], x2=o[2], v2=o[3], f2=o[4], solves=solves.value)


def cg_solve(outer, inner, vals, b, tol, max_iter=-1):
    outer = np.ascontiguousarray(outer, dtype=np.int32)
    inner = np.ascontiguousarray(inner, dtype=np.int32)
    vals = _f64(vals)
    b = _f64(b)
    x = np.zeros_like(b)
    err = np.zeros(1)
    it = lib().orc_cg_solve(len(outer) - 1, _i(outer), _i(inner), _d(vals), _d(b), _d(x), tol, max_iter, _d(err))
    return x, it, err[0]


class Sim:
    """Oracle FemRepresentation + OdeSolver (implicit Euler) over one mesh."""

    def __init__(self, kind, x0, conn, E, nu, rho):
        x0 = _f64(x0).reshape(-1)
        n_nodes = x0.size // 3
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        n_el = conn.shape[0]
        E = _f64(np.broadcast_to(E, (n_el,)))
        nu = _f64(np.broadcast_to(nu, (n_el,)))
        rho = _f64(np.broadcast_to(rho, (n_el,)))
        self._keep = (x0, conn, E, nu, rho)
        self.n_nodes = n_nodes
        self.ndof = 3 * n_nodes
        self.h = lib().orc_sim_create(kind, n_nodes, _d(x0), n_el, _i(conn), _d(E), _d(nu), _d(rho))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def add_bc_node(self, node):
        for c in range(3):
            lib().orc_sim_add_bc_dof(self.h, 3 * int(node) + c)

    def add_bc_dof(self, dof):
        lib().orc_sim_add_bc_dof(self.h, int(dof))

    def set_gravity(self, g=(0.0, -9.81, 0.0), enabled=True):
        lib().orc_sim_set_gravity(self.h, g[0], g[1], g[2], int(enabled))

    def set_external_force(self, f):
        if f is None:
            lib().orc_sim_set_external_force(self.h, None)
        else:
            self._fext = _f64(f).reshape(-1)
            lib().orc_sim_set_external_force(self.h, _d(self._fext))

    def set_damping(self, rayleigh_mass=0.0, rayleigh_stiffness=0.0, lumping=False):
        lib().orc_sim_set_damping(self.h, float(rayleigh_mass), float(rayleigh_stiffness), int(bool(lumping)))

    def add_localized_force(self, element, coord, force, K=None, D=None):
        """Force (3) at natural coordinates coord of an element, with optional 3x3 stiffness K and damping
        D, for the next step only."""
        arrs = [_f64(coord).reshape(-1), _f64(force).reshape(-1)]
        k = _f64(K).reshape(-1) if K is not None else None
        d = _f64(D).reshape(-1) if D is not None else None
        rc = lib().orc_sim_add_localized_force(self.h, int(element), _d(arrs[0]), _d(arrs[1]),
                                               _d(k) if k is not None else None, _d(d) if d is not None else None)
        if rc < 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def update_fmdk(self):
        """Rebuild f, M, D, K at the current state as computeFMDK does, without stepping."""
        if lib().orc_sim_update_fmdk(self.h) < 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def apply_compliance(self, b):
        """C b (b of shape (ndof,) or (ndof, k)) with the last system matrix (applyCompliance)."""
        b = np.asarray(b, dtype=np.float64)
        vec = b.ndim == 1
        cols = np.ascontiguousarray((b[:, None] if vec else b).T)
        out = np.zeros_like(cols)
        if lib().orc_sim_apply_compliance(self.h, _d(cols), cols.shape[0], _d(out)) < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return out[0] if vec else out.T

    def set_solver(self, scheme=EULER_IMPLICIT, linsolver=CG, tol=2.220446049250313e-16, max_iter=-1,
                   newton_max_iter=1, faithful_bc=False, newton_eps=1e-5):
        lib().orc_sim_set_solver(self.h, scheme, linsolver, tol, max_iter, newton_max_iter, int(faithful_bc))
        lib().orc_sim_set_newton_eps(self.h, float(newton_eps))

    def set_state(self, x=None, v=None):
        xa = _f64(x).reshape(-1) if x is not None else None
        va = _f64(v).reshape(-1) if v is not None else None
        lib().orc_sim_set_state(self.h, _d(xa) if xa is not None else None, _d(va) if va is not None else None)

    def state(self):
        x = np.zeros(self.ndof)
        v = np.zeros(self.ndof)
        lib().orc_sim_get_state(self.h, _d(x), _d(v))
        return x, v

    def step(self, dt, n=1):
        rc = lib().orc_sim_step(self.h, dt, n)
        if rc < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return rc

    def set_young(self, E):
        self._E = _f64(np.broadcast_to(E, (self._keep[1].shape[0],)))
        if lib().orc_sim_set_young(self.h, _d(self._E)) < 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def adjoint_step(self, dt, xn, vn, vnext, lx, lv, gradE):
        """Reverse of one implicit step from (xn, vn) to vnext: returns (dL/dxn, dL/dvn) and adds dL/dE
        through the step to gradE (in place)."""
        arrs = [_f64(a).reshape(-1) for a in (xn, vn, vnext, lx, lv)]
        lxo = np.zeros(self.ndof)
        lvo = np.zeros(self.ndof)
        assert gradE.dtype == np.float64 and gradE.flags.c_contiguous
        if lib().orc_sim_adjoint_step(self.h, dt, *[_d(a) for a in arrs], _d(lxo), _d(lvo), _d(gradE)) < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return lxo, lvo

    def pattern(self):
        nnz = lib().orc_sim_nnz(self.h)
        outer = np.zeros(self.ndof + 1, dtype=np.int32)
        inner = np.zeros(nnz, dtype=np.int32)
        lib().orc_sim_pattern(self.h, _i(outer), _i(inner))
        return outer, inner

    def matrix_values(self, which):
        """which: 0 = M, 1 = K, 2 = system matrix A (after boundary conditions), 3 = D."""
        vals = np.zeros(lib().orc_sim_nnz(self.h))
        lib().orc_sim_matrix_values(self.h, which, _d(vals))
        return vals

    def rhs(self):
        r = np.zeros(self.ndof)
        lib().orc_sim_rhs(self.h, _d(r))
        return r

    def force(self):
        r = np.zeros(self.ndof)
        lib().orc_sim_force(self.h, _d(r))
        return r

    def mass_per_node(self):
        r = np.zeros(self.n_nodes)
        lib().orc_sim_mass_per_node(self.h, _d(r))
        return r

    def last_cg(self):
        return lib().orc_sim_last_cg_iters(self.h), lib().orc_sim_last_cg_error(self.h)

    def total_cg_iters(self):
        return lib().orc_sim_total_cg_iters(self.h)

    def set_compliance_at_first_linear_step(self, on=True):
        lib().orc_sim_set_compliance_at_first_linear_step(self.h, int(on))

    def compliance_matrix(self):
        """Dense compliance of the last system matrix, boundary rows and columns zero (D x D)."""
        C = np.zeros((self.ndof, self.ndof), order="F")
        _check(lib().orc_sim_compliance_matrix(self.h, C.ctypes.data_as(_dp)))
        return C

    def last_newton_iters(self):
        return lib().orc_sim_last_newton_iters(self.h)

    def counters(self):
        a = ctypes.c_long(0)
        b = ctypes.c_long(0)
        lib().orc_sim_counters(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def valid(self):
        return bool(lib().orc_sim_valid(self.h))

    def element(self, i, nn=4):
        R = np.zeros((3, 3))
        K = np.zeros((3 * nn, 3 * nn))
        f = np.zeros(3 * nn)
        lib().orc_sim_element(self.h, i, _d(R), _d(K), _d(f))
        return R, K, f
