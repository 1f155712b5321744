"""ctypes binding of libfem_gpu.so (include/fem_gpu.h).

The library is loaded from the package's _lib/ directory; there is no fallback: a missing or
unloadable library raises ImportError / OSError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libfem_gpu.so")

FEM_OK, FEM_INACTIVE = 0, 1
FEM_ERR_INVALID, FEM_ERR_HIP, FEM_ERR_ORDER, FEM_ERR_NOT_CONVERGED = -1, -2, -3, -4
TET_LINEAR, TET_COROTATIONAL, HEX = 0, 1, 2
EULER_IMPLICIT, LINEAR_EULER_IMPLICIT = 0, 1
STATIC, LINEAR_STATIC = 2, 3
EULER_EXPLICIT, LINEAR_EULER_EXPLICIT = 4, 5
EULER_EXPLICIT_MODIFIED, LINEAR_EULER_EXPLICIT_MODIFIED = 6, 7
RUNGE_KUTTA_4, LINEAR_RUNGE_KUTTA_4 = 8, 9
FP64, FP32_REFINED = 0, 1
LINEAR_SOLVER_LU, LINEAR_SOLVER_CG = 0, 1

# Names of the exported entry points (tests/test_abi.py compares this list with include/fem_gpu.h).
EXPORTS = [
    "fem_create", "fem_destroy", "fem_last_error", "fem_set_stream", "fem_upload_mesh",
    "fem_add_boundary_condition_node", "fem_add_boundary_condition_dof", "fem_default_options",
    "fem_set_options", "fem_set_external_force", "fem_add_external_generalized_force", "fem_apply_compliance",
    "fem_apply_compliance_device", "fem_apply_compliance_device_noalloc", "fem_constraint_compliance",
    "fem_get_compliance_matrix", "fem_get_compliance_matrix_device",
    "fem_set_state", "fem_get_state", "fem_set_state_device",
    "fem_get_state_device", "fem_step", "fem_synchronize", "fem_query", "fem_num_nonzeros", "fem_get_pattern",
    "fem_get_system_matrix", "fem_get_rhs", "fem_evaluate_elements", "fem_enable_phase_timers",
    "fem_get_phase_times", "fem_probe_copy", "fem_time_operator", "fem_get_node_masses",
    "fem_set_young_modulus", "fem_adjoint_step", "fem_adjoint_steps", "fem_step_record", "fem_set_instances",
    "fem_query_instances",
    "fem_set_batch_tuning", "fem_get_batch_tuning", "fem_hex_distinct_matrices",
]

PHASES = ["element", "assemble", "rhs", "cg_init", "spmv", "cg_vector", "refine", "update", "spmv_skipped"]


class FemOptions(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("gravity_enabled", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("newton_max_iterations", ctypes.c_int32),
        ("cg_tolerance", ctypes.c_double),
        ("cg_max_iterations", ctypes.c_int64),
        ("inner_tolerance", ctypes.c_double),
        ("max_refinements", ctypes.c_int32),
        ("fail_on_nonconvergence", ctypes.c_int32),
        ("rayleigh_mass", ctypes.c_double),
        ("rayleigh_stiffness", ctypes.c_double),
        ("mass_lumping", ctypes.c_int32),
        ("newton_epsilon", ctypes.c_double),
        ("linear_solver", ctypes.c_int32),
    ]


class FemStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("last_cg_iterations", ctypes.c_int64),
        ("last_cg_error", ctypes.c_double),
        ("total_cg_iterations", ctypes.c_int64),
        ("last_refinements", ctypes.c_int64),
        ("det_r_invalid", ctypes.c_int64),
        ("g_singular", ctypes.c_int64),
        ("fp64_polar_fallbacks", ctypes.c_int64),
        ("active", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("last_newton_iterations", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FemInstanceStats(ctypes.Structure):
    _fields_ = [("last_iterations", ctypes.c_int64), ("last_error", ctypes.c_double),
                ("unconverged_steps", ctypes.c_int64), ("converged", ctypes.c_int32), ("active", ctypes.c_int32),
                ("solved_in_fp64", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class FemPhaseTimes(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 9), ("launches", ctypes.c_int64 * 9)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def load():
    """Load libfem_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libfem_gpu.so is not built (run opensurgsim_amd/build.py): " + LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.fem_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.fem_destroy.argtypes = [_vp]
    L.fem_destroy.restype = None
    L.fem_last_error.argtypes = [_vp]
    L.fem_last_error.restype = ctypes.c_char_p
    L.fem_set_stream.argtypes = [_vp, _vp]
    L.fem_upload_mesh.argtypes = [_vp, ctypes.c_int, _i64, _dp, _i64, _i32p, _dp, _dp, _dp]
    L.fem_add_boundary_condition_node.argtypes = [_vp, _i64]
    L.fem_add_boundary_condition_dof.argtypes = [_vp, _i64]
    L.fem_default_options.argtypes = [ctypes.POINTER(FemOptions)]
    L.fem_default_options.restype = None
    L.fem_set_options.argtypes = [_vp, ctypes.POINTER(FemOptions)]
    L.fem_set_external_force.argtypes = [_vp, _dp]
    L.fem_add_external_generalized_force.argtypes = [_vp, ctypes.c_int64, _dp, _dp, _dp, _dp]
    L.fem_apply_compliance.argtypes = [_vp, _dp, ctypes.c_int64, _dp]
    L.fem_apply_compliance_device.argtypes = [_vp, _vp, ctypes.c_int64, _vp]
    L.fem_apply_compliance_device_noalloc.argtypes = [_vp, _vp, ctypes.c_int64, _vp]
    L.fem_get_compliance_matrix.argtypes = [_vp, _dp]
    L.fem_get_compliance_matrix_device.argtypes = [_vp, _vp]
    L.fem_constraint_compliance.argtypes = [_vp, _i64, _dp, ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
    L.fem_set_state.argtypes = [_vp, _dp, _dp]
    L.fem_get_state.argtypes = [_vp, _dp, _dp]
    L.fem_set_state_device.argtypes = [_vp, _vp, _vp]
    L.fem_get_state_device.argtypes = [_vp, _vp, _vp]
    L.fem_step.argtypes = [_vp, ctypes.c_int32]
    L.fem_synchronize.argtypes = [_vp]
    L.fem_query.argtypes = [_vp, ctypes.POINTER(FemStats)]
    L.fem_num_nonzeros.argtypes = [_vp]
    L.fem_num_nonzeros.restype = _i64
    L.fem_hex_distinct_matrices.argtypes = [_vp]
    L.fem_hex_distinct_matrices.restype = ctypes.c_int
    L.fem_get_pattern.argtypes = [_vp, _i32p, _i32p]
    L.fem_get_system_matrix.argtypes = [_vp, _dp]
    L.fem_get_rhs.argtypes = [_vp, _dp]
    L.fem_evaluate_elements.argtypes = [_vp, _dp, _i64, _i64, _dp, _dp, _dp]
    L.fem_enable_phase_timers.argtypes = [_vp, ctypes.c_int]
    L.fem_get_phase_times.argtypes = [_vp, ctypes.POINTER(FemPhaseTimes)]
    L.fem_get_node_masses.argtypes = [_vp, _dp]
    L.fem_set_young_modulus.argtypes = [_vp, _dp]
    L.fem_set_instances.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), _dp]
    L.fem_query_instances.argtypes = [_vp, ctypes.POINTER(FemInstanceStats)]
    L.fem_set_batch_tuning.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.fem_get_batch_tuning.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32), _dp]
    L.fem_adjoint_step.argtypes = [_vp] + [_vp] * 8
    L.fem_adjoint_steps.argtypes = [_vp, ctypes.c_int32] + [_vp] * 5
    L.fem_step_record.argtypes = [_vp, ctypes.c_int32, _vp, _vp]
    L.fem_time_operator.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.fem_probe_copy.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


class FemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("[%d] %s" % (code, msg))
        self.code = code


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def default_options(**kw):
    o = FemOptions()
    load().fem_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if k == "gravity":
            for i in range(3):
                o.gravity[i] = v[i]
        else:
            setattr(o, k, v)
    return o


def probe_copy(device, nbytes, reps=5):
    """Average milliseconds of a plain device copy of nbytes (counter calibration, achievable bandwidth)."""
    L = load()
    ms = ctypes.c_double(0.0)
    rc = L.fem_probe_copy(int(device), int(nbytes), int(reps), ctypes.byref(ms))
    if rc != FEM_OK:
        raise FemError("fem_probe_copy failed (%d)" % rc)
    return ms.value


def _torch_first():
    """libfem_gpu.so uses the system ROCm runtime; PyTorch ships its own HIP runtime, which fails to find
    the GPU when it starts after the other one in the same process. When torch is already imported, let it
    initialise the GPU before the library does."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and hasattr(torch, "cuda") and torch.cuda.is_available():
        torch.cuda.init()


class FemGpu:
    """One GPU representation (context) of the C ABI."""

    def __init__(self, device=0):
        _torch_first()
        self._L = load()
        h = _vp()
        rc = self._L.fem_create(int(device), ctypes.byref(h))
        if rc != FEM_OK:
            raise FemError(rc, "fem_create failed (no HIP device?)")
        self.h = h
        self.ndof = 0
        self.nn = 4
        self.num_elements = 0

    def close(self):
        if getattr(self, "h", None):
            self._L.fem_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, allow_inactive=False):
        if rc == FEM_OK or (allow_inactive and rc == FEM_INACTIVE):
            return rc
        raise FemError(rc, self._L.fem_last_error(self.h).decode())

    def upload_mesh(self, kind, x0, conn, young, poisson, density):
        x0 = _f64(x0).reshape(-1)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        ne = conn.shape[0]
        E = _f64(np.broadcast_to(young, (ne,)))
        nu = _f64(np.broadcast_to(poisson, (ne,)))
        rho = _f64(np.broadcast_to(density, (ne,)))
        self._check(self._L.fem_upload_mesh(self.h, int(kind), x0.size // 3, _ptr(x0), ne,
                                            conn.ctypes.data_as(_i32p), _ptr(E), _ptr(nu), _ptr(rho)))
        self.ndof = x0.size
        self.nn = conn.shape[1]
        self.num_elements = ne

    def add_bc_node(self, node):
        self._check(self._L.fem_add_boundary_condition_node(self.h, int(node)))

    def add_bc_nodes(self, nodes):
        for n in np.asarray(nodes).reshape(-1):
            self.add_bc_node(n)

    def add_bc_dof(self, dof):
        self._check(self._L.fem_add_boundary_condition_dof(self.h, int(dof)))

    def set_options(self, opt=None, **kw):
        if opt is None:
            opt = default_options(**kw)
        self._check(self._L.fem_set_options(self.h, ctypes.byref(opt)))
        return opt

    def set_external_force(self, f):
        if f is None:
            self._check(self._L.fem_set_external_force(self.h, None))
        else:
            f = _f64(f).reshape(-1)
            self._check(self._L.fem_set_external_force(self.h, _ptr(f)))

    def add_external_generalized_force(self, element, coordinate, force, K=None, D=None):
        """Localized force (3) at the natural coordinates of an element, with optional 3x3 stiffness K and
        damping D, for the next step only (Fem3DRepresentation::addExternalGeneralizedForce)."""
        arrs = [_f64(coordinate).reshape(-1), _f64(force).reshape(-1)]
        k = _f64(K).reshape(-1) if K is not None else None
        d = _f64(D).reshape(-1) if D is not None else None
        self._check(self._L.fem_add_external_generalized_force(self.h, int(element), _ptr(arrs[0]), _ptr(arrs[1]),
                                                               _ptr(k), _ptr(d)))

    def apply_compliance(self, b):
        """C b for b of shape (ndof,) or (ndof, k): the last step's system matrix solved per column with
        boundary rows zeroed (DeformableRepresentation::applyCompliance)."""
        b = np.asarray(b, dtype=np.float64)
        vec = b.ndim == 1
        cols = np.ascontiguousarray((b[:, None] if vec else b).T)  # column-major: one row per column here
        out = np.zeros_like(cols)
        self._check(self._L.fem_apply_compliance(self.h, _ptr(cols), cols.shape[0], _ptr(out)))
        return out[0] if vec else out.T

    def apply_compliance_device(self, b_ptr, num_columns, out_ptr):
        """C b with device pointers (column-major ndof x num_columns FP64), ordered on the context stream."""
        self._check(self._L.fem_apply_compliance_device(self.h, ctypes.c_void_p(b_ptr), int(num_columns),
                                                        ctypes.c_void_p(out_ptr)))

    def apply_compliance_device_noalloc(self, b_ptr, num_columns, out_ptr):
        """apply_compliance_device through the entry point that allocates nothing (fixed per-GPU workspace)."""
        self._check(self._L.fem_apply_compliance_device_noalloc(self.h, ctypes.c_void_p(b_ptr), int(num_columns),
                                                                ctypes.c_void_p(out_ptr)))

    def compliance_matrix(self):
        """Dense compliance of the last system matrix (ndof x ndof), boundary rows and columns zero
        (OdeSolver::getComplianceMatrix)."""
        C = np.zeros((self.ndof, self.ndof), order="F")
        self._check(self._L.fem_get_compliance_matrix(self.h, C.ctypes.data_as(_dp)))
        return C

    def compliance_matrix_device(self, out_ptr):
        """The same into device memory (ndof x ndof FP64, column-major), ordered on the context stream."""
        self._check(self._L.fem_get_compliance_matrix_device(self.h, ctypes.c_void_p(out_ptr)))

    def constraint_compliance(self, element, coordinate, directions, scale, want_cht=True):
        """(C H^T of shape (ndof, m) or None, H C H^T of shape (m, m)) for a constraint localized in one element:
        row j of H is scale * coordinate[a] * directions[j] at node a's DOFs (FemConstraintFixedPoint: the three
        axes, FemConstraintFrictionlessContact: the normal; scale = dt * the builder's scale)."""
        coord = _f64(coordinate).reshape(-1)
        if coord.size != self.nn:
            raise ValueError(f"coordinate needs {self.nn} values, one per element node")
        dirs = np.ascontiguousarray(_f64(directions).reshape(-1, 3))
        m = dirs.shape[0]
        cht = np.zeros((m, self.ndof)) if want_cht else None
        hcht = np.zeros((m, m))
        self._check(self._L.fem_constraint_compliance(self.h, int(element), _ptr(coord), m, _ptr(dirs), float(scale),
                                                      _ptr(cht), _ptr(hcht)))
        return (cht.T if want_cht else None), hcht

    def set_state(self, x=None, v=None):
        xa = _f64(x).reshape(-1) if x is not None else None
        va = _f64(v).reshape(-1) if v is not None else None
        self._check(self._L.fem_set_state(self.h, _ptr(xa), _ptr(va)))

    def state(self):
        x = np.zeros(self.ndof)
        v = np.zeros(self.ndof)
        self._check(self._L.fem_get_state(self.h, _ptr(x), _ptr(v)))
        return x, v

    def set_state_device(self, x_ptr, v_ptr):
        self._check(self._L.fem_set_state_device(self.h, x_ptr, v_ptr))

    def get_state_device(self, x_ptr, v_ptr):
        self._check(self._L.fem_get_state_device(self.h, x_ptr, v_ptr))

    def step(self, n=1):
        return self._check(self._L.fem_step(self.h, int(n)), allow_inactive=True)

    def step_record(self, n, x_hist, v_hist):
        """n steps; device pointers (ints) x_hist, v_hist ((n + 1) x ndof) receive the start and each step's
        state (fem_step_record)."""
        return self._check(self._L.fem_step_record(self.h, int(n), _vp(int(x_hist)), _vp(int(v_hist))),
                           allow_inactive=True)

    def hex_distinct_matrices(self):
        """Distinct FP32 element matrices the hexahedral inner solve reads from its table (0: per element)."""
        return int(self._L.fem_hex_distinct_matrices(self.h))

    def synchronize(self):
        self._check(self._L.fem_synchronize(self.h))

    def stats(self):
        s = FemStats()
        self._check(self._L.fem_query(self.h, ctypes.byref(s)))
        return s.as_dict()

    def enable_phase_timers(self, every=1):
        """Bracket every kernel of one step in `every` with HIP events (0 turns them off)."""
        self._check(self._L.fem_enable_phase_timers(self.h, int(every)))

    def phase_times(self):
        t = FemPhaseTimes()
        self._check(self._L.fem_get_phase_times(self.h, ctypes.byref(t)))
        return {name: (t.ms[i], t.launches[i]) for i, name in enumerate(PHASES)}

    def set_instances(self, node_offsets, dt=None):
        """Declare the uploaded mesh as independent instances (fem_set_instances)."""
        off = np.ascontiguousarray(node_offsets, dtype=np.int64)
        self._inst_dt = None if dt is None else _f64(dt)
        self._check(self._L.fem_set_instances(self.h, off.size - 1, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              _ptr(self._inst_dt) if self._inst_dt is not None else None))
        self.num_instances = off.size - 1

    def instance_stats(self):
        """Per-instance (last iterations, last error, converged) arrays; also sets last_instance_active,
        last_instance_unconverged_steps and last_instance_fp64."""
        m = self.num_instances
        st = (FemInstanceStats * m)()
        self._check(self._L.fem_query_instances(self.h, st))
        it = np.array([s.last_iterations for s in st], dtype=np.int64)
        err = np.array([s.last_error for s in st])
        conv = np.array([s.converged for s in st], dtype=np.int32)
        self.last_instance_active = np.array([s.active for s in st], dtype=np.int32)
        self.last_instance_unconverged_steps = np.array([s.unconverged_steps for s in st], dtype=np.int64)
        self.last_instance_fp64 = np.array([s.solved_in_fp64 for s in st], dtype=np.int32)
        return it, err, conv

    def set_batch_tuning(self, chunk_rows=0, lanes_per_row=0, use_graphs=True):
        self._check(self._L.fem_set_batch_tuning(self.h, int(chunk_rows), int(lanes_per_row), int(bool(use_graphs))))

    def batch_tuning(self):
        r, l, k = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_double()
        self._check(self._L.fem_get_batch_tuning(self.h, ctypes.byref(r), ctypes.byref(l), ctypes.byref(k),
                                                 ctypes.byref(ms)))
        return dict(chunk_rows=r.value, lanes_per_row=l.value, chunks=k.value, ms_per_iteration=ms.value)

    def set_young_modulus(self, young):
        E = _f64(np.broadcast_to(young, (self.num_elements,)))
        self._check(self._L.fem_set_young_modulus(self.h, _ptr(E)))

    def set_stream(self, stream_handle):
        """Run on an existing hipStream_t (an int handle, e.g. torch.cuda.current_stream().cuda_stream)."""
        self._check(self._L.fem_set_stream(self.h, _vp(int(stream_handle)) if stream_handle else None))

    def adjoint_step(self, x_n, v_n, v_next, lam_x, lam_v, lam_x_out, lam_v_out, grad_young):
        """Device pointers (ints): reverse of one step; grad_young accumulated (fem_adjoint_step)."""
        args = [_vp(int(p)) for p in (x_n, v_n, v_next, lam_x, lam_v, lam_x_out, lam_v_out, grad_young)]
        self._check(self._L.fem_adjoint_step(self.h, *args))

    def adjoint_steps(self, n, x_hist, v_hist, lam_x, lam_v, grad_young):
        """Device pointers (ints): n adjoint steps backwards through a recorded history, lam_x, lam_v updated
        in place, grad_young accumulated (fem_adjoint_steps)."""
        args = [_vp(int(p)) for p in (x_hist, v_hist, lam_x, lam_v, grad_young)]
        self._check(self._L.fem_adjoint_steps(self.h, int(n), *args))

    def node_masses(self):
        m = np.zeros(self.ndof // 3)
        self._check(self._L.fem_get_node_masses(self.h, m.ctypes.data_as(_dp)))
        return m

    def time_operator(self, reps=50):
        """Average milliseconds of one CG operator launch (back-to-back replay on CG scratch buffers)."""
        ms = ctypes.c_double(0.0)
        self._check(self._L.fem_time_operator(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    def num_nonzeros(self):
        """Scalar nonzeros of the sparsity pattern (9 per 3x3 node block)."""
        return int(self._L.fem_num_nonzeros(self.h))

    def pattern(self):
        nnz = self._L.fem_num_nonzeros(self.h)
        rowptr = np.zeros(self.ndof + 1, dtype=np.int32)
        colidx = np.zeros(nnz, dtype=np.int32)
        self._check(self._L.fem_get_pattern(self.h, rowptr.ctypes.data_as(_i32p), colidx.ctypes.data_as(_i32p)))
        return rowptr, colidx

    def system_matrix(self):
        vals = np.zeros(self._L.fem_num_nonzeros(self.h))
        self._check(self._L.fem_get_system_matrix(self.h, _ptr(vals)))
        return vals

    def rhs(self):
        r = np.zeros(self.ndof)
        self._check(self._L.fem_get_rhs(self.h, _ptr(r)))
        return r

    def evaluate_elements(self, x, first=0, count=None):
        if count is None:
            count = self.num_elements - first
        x = _f64(x).reshape(-1)
        n3 = 3 * self.nn
        R = np.zeros((count, 3, 3))
        K = np.zeros((count, n3, n3))
        f = np.zeros((count, n3))
        self._check(self._L.fem_evaluate_elements(self.h, _ptr(x), int(first), int(count), _ptr(R), _ptr(K),
                                                  _ptr(f)))
        return R, K, f
