"""Shared test helpers: build the oracle and the GPU representation from one configuration."""
import numpy as np
import scipy.sparse as sp

import oracle
from opensurgsim_amd import femgpu


def make_oracle(cfg, scheme=oracle.EULER_IMPLICIT, linsolver=oracle.CG, tol=None, faithful_bc=False):
    s = oracle.Sim(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    for n in cfg.bc_nodes:
        s.add_bc_node(n)
    s.set_gravity(cfg.gravity, True)
    if cfg.fext is not None:
        s.set_external_force(cfg.fext)
    s.set_solver(scheme=scheme, linsolver=linsolver, tol=cfg.cg_tol if tol is None else tol,
                 faithful_bc=faithful_bc)
    s.set_state(v=cfg.v0)
    return s


def make_gpu(cfg, precision=femgpu.FP64, scheme=femgpu.EULER_IMPLICIT, tol=None, **opts):
    g = femgpu.FemGpu(0)
    g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    g.add_bc_nodes(cfg.bc_nodes)
    g.set_options(dt=cfg.dt, cg_tolerance=cfg.cg_tol if tol is None else tol, precision=precision,
                  scheme=scheme, gravity=cfg.gravity, **opts)
    if cfg.fext is not None:
        g.set_external_force(cfg.fext)
    g.set_state(v=cfg.v0)
    return g


def lower_view_csr(outer, inner, vals):
    """The lower-triangle self-adjoint view (Eigen selfadjointView<Lower>) of a CSC matrix, as CSR
    over the same (symmetric) pattern."""
    n = len(outer) - 1
    A = sp.csc_matrix((vals, inner, outer), shape=(n, n))
    L = sp.tril(A, format="csr")
    S = (L + sp.tril(A, -1, format="csr").T).tocsr()
    # restore the full pattern order (explicit zeros included) by sampling S at the pattern
    P = sp.csc_matrix((np.ones_like(vals), inner, outer), shape=(n, n)).tocsr()
    P.sort_indices()
    rows = np.repeat(np.arange(n), np.diff(P.indptr))
    out = np.asarray(S[rows, P.indices]).ravel()
    return P.indptr, P.indices, out


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
