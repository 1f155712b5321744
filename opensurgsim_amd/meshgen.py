"""Synthetic block meshes, seeded random inputs and the benchmark configurations.

The reference ships a hex block generator (Physics/PerformanceTests/DivisibleCubeRepresentation.cpp:30-157)
and a 5-tetrahedra split of one cube (Physics/RenderTests/RenderTestFem3DCorotational.cpp:53-74).
The block tet meshes here extend that split to a grid with parity-alternating cubes so that
neighbouring cubes share face diagonals.

Conventions (SI units):
  * node id(i, j, k) = i*(ny+1)*(nz+1) + j*(nz+1) + k, i along x, j along y, k along z
    (generalises DivisibleCubeRepresentation::get1DIndexFrom3D, :69-72);
  * y is vertical (reference gravity is (0, -9.81, 0), Physics/Representation.cpp:36);
    "bottom" is the y = 0 face, "top" the y = max face;
  * cube corner L = a + 2b + 4c for offsets (a, b, c) along (x, y, z);
  * even cubes (i+j+k even) use the reference split {4,7,1,2},{4,1,7,5},{4,2,1,0},{4,7,2,6},{1,2,7,3};
    odd cubes mirror it (L -> L xor 1) and swap the last two ids of every tet to restore positive
    orientation (Physics/Fem3DElementTetrahedron.h:48-52 requires ABC counter-clockwise seen from D);
  * hex node order (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k), then the same at k+1: the reference's
    {0,1,3,2,4,5,7,6} permutation (DivisibleCubeRepresentation.cpp:134-145).
Random inputs come from Philox4x32-10 (Salmon et al., SC'11) with a Box-Muller transform, so the
host oracle and every GPU run see identical inputs.
"""
import math

import numpy as np

_REF_SPLIT = np.array([[4, 7, 1, 2], [4, 1, 7, 5], [4, 2, 1, 0], [4, 7, 2, 6], [1, 2, 7, 3]], dtype=np.int64)
_ODD_SPLIT = (_REF_SPLIT ^ 1)[:, [0, 1, 3, 2]]

# ------------------------------------------------------------------------------------------
# Philox4x32-10
# ------------------------------------------------------------------------------------------
_PHILOX_M0 = np.uint64(0xD2511F53)
_PHILOX_M1 = np.uint64(0xCD9E8D57)
_PHILOX_W0 = 0x9E3779B9
_PHILOX_W1 = 0xBB67AE85
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32(counter_lo, seed, stream=0):
    """Philox4x32-10 of counters (counter_lo, stream, 0, 0) under key (seed lo, seed hi).

    Returns a (n, 4) uint32 array.
    """
    c0 = np.asarray(counter_lo, dtype=np.uint64) & _MASK32
    c1 = np.full_like(c0, np.uint64(stream) & _MASK32)
    c2 = np.zeros_like(c0)
    c3 = np.zeros_like(c0)
    k0 = int(seed) & 0xFFFFFFFF
    k1 = (int(seed) >> 32) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _PHILOX_M0 * c0
        p1 = _PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0)), lo1, (hi0 ^ c3 ^ np.uint64(k1)), lo0
        k0 = (k0 + _PHILOX_W0) & 0xFFFFFFFF
        k1 = (k1 + _PHILOX_W1) & 0xFFFFFFFF
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def uniform(n, seed, stream=0):
    """n uniforms in (0, 1): (u + 0.5) / 2^32 of the Philox output words."""
    blocks = (n + 3) // 4
    w = philox4x32(np.arange(blocks, dtype=np.uint64), seed, stream).reshape(-1)[:n]
    return (w.astype(np.float64) + 0.5) / 4294967296.0


def normal(n, seed, stream=0):
    """n standard normals by Box-Muller on Philox uniforms (pairs (u1, u2) -> cos branch)."""
    u = uniform(2 * n, seed, stream).reshape(n, 2)
    return np.sqrt(-2.0 * np.log(u[:, 0])) * np.cos(2.0 * math.pi * u[:, 1])


# ------------------------------------------------------------------------------------------
# Meshes
# ------------------------------------------------------------------------------------------
def grid_nodes(nx, ny, nz, h):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    return np.stack([i.ravel() * h, j.ravel() * h, k.ravel() * h], axis=1).astype(np.float64)


def node_id(i, j, k, ny, nz):
    return (np.asarray(i) * (ny + 1) + np.asarray(j)) * (nz + 1) + np.asarray(k)


def _cube_corners(nx, ny, nz):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    L = np.arange(8)
    a, b, c = L & 1, (L >> 1) & 1, (L >> 2) & 1
    corners = node_id(i[:, None] + a, j[:, None] + b, k[:, None] + c, ny, nz)
    return i, j, k, corners


def block_tet_mesh(nx, ny, nz, h=0.01):
    """(x0 [N,3], tets [5*nx*ny*nz, 4] int32, cube index of every tet)."""
    x0 = grid_nodes(nx, ny, nz, h)
    i, j, k, corners = _cube_corners(nx, ny, nz)
    odd = ((i + j + k) & 1).astype(bool)
    split = np.where(odd[:, None, None], _ODD_SPLIT[None], _REF_SPLIT[None])
    tets = np.take_along_axis(corners[:, None, :].repeat(5, axis=1), split, axis=2).reshape(-1, 4)
    cube_of_tet = np.repeat(np.arange(corners.shape[0]), 5)
    return x0, tets.astype(np.int32), cube_of_tet


def block_hex_mesh(nx, ny, nz, h=0.01):
    """(x0 [N,3], hexes [nx*ny*nz, 8] int32) in the reference's cube node order."""
    x0 = grid_nodes(nx, ny, nz, h)
    _, _, _, corners = _cube_corners(nx, ny, nz)
    hexes = corners[:, [0, 1, 3, 2, 4, 5, 7, 6]]
    return x0, hexes.astype(np.int32)


def tet_volumes(x0, tets):
    p = x0[tets]
    return np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]) / 6.0


def face_nodes(nx, ny, nz, axis, side):
    """Node ids on a block face: axis 0/1/2, side 0 (min) or 1 (max)."""
    n = [nx, ny, nz]
    rng = [np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1)]
    rng[axis] = np.array([0 if side == 0 else n[axis]])
    i, j, k = np.meshgrid(*rng, indexing="ij")
    return node_id(i.ravel(), j.ravel(), k.ravel(), ny, nz).astype(np.int64)


def top_face_triangles_tet(nx, ny, nz):
    """Triangles of the y = max face of a block tet mesh (the faces of the tets there)."""
    i, k = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, k = i.ravel(), k.ravel()
    j = ny - 1
    odd = ((i + j + k) & 1).astype(bool)
    # even cube: diagonal 2-7 -> triangles (2,7,6), (2,7,3); odd: diagonal 3-6 -> (3,6,7), (3,6,2)
    tri_even = np.array([[2, 7, 6], [2, 7, 3]])
    tri_odd = np.array([[3, 6, 7], [3, 6, 2]])
    L = np.where(odd[:, None, None], tri_odd[None], tri_even[None])
    a, b, c = L & 1, (L >> 1) & 1, (L >> 2) & 1
    return node_id(i[:, None, None] + a, j + b, k[:, None, None] + c, ny, nz).reshape(-1, 3)


def top_face_triangles_hex(nx, ny, nz):
    """Two triangles per y = max hex face, split along (i,k)-(i+1,k+1)."""
    i, k = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, k = i.ravel(), k.ravel()
    j = ny
    n00 = node_id(i, j, k, ny, nz)
    n10 = node_id(i + 1, j, k, ny, nz)
    n11 = node_id(i + 1, j, k + 1, ny, nz)
    n01 = node_id(i, j, k + 1, ny, nz)
    return np.concatenate([np.stack([n00, n10, n11], 1), np.stack([n00, n11, n01], 1)])


def pressure_forces(x0, tris, pressure_fn):
    """Nodal forces of a pressure on triangles: each triangle adds p(centroid) * A / 3 along -y."""
    n = x0.shape[0]
    p = x0[tris]
    area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
    cen = p.mean(axis=1)
    pr = pressure_fn(cen)
    f = np.zeros((n, 3))
    share = pr * area / 3.0
    for c in range(3):
        np.add.at(f[:, 1], tris[:, c], -share)
    return f.reshape(-1)


# ------------------------------------------------------------------------------------------
# Configurations
# ------------------------------------------------------------------------------------------
class Config(dict):
    """A workload: mesh, materials, boundary conditions, loads, initial state and solver settings."""

    __getattr__ = dict.__getitem__


def _tet_config(name, n, kind, E, nu, rho, dt, steps, tol, seed, h=0.01):
    nx, ny, nz = n
    x0, tets, cube = block_tet_mesh(nx, ny, nz, h)
    Ne = tets.shape[0]
    return Config(name=name, kind=kind, nx=nx, ny=ny, nz=nz, h=h, x0=x0, conn=tets, cube_of_elem=cube,
                  E=np.full(Ne, float(E)) if np.isscalar(E) else np.asarray(E, dtype=np.float64),
                  nu=np.full(Ne, float(nu)), rho=np.full(Ne, float(rho)),
                  bc_nodes=face_nodes(nx, ny, nz, 1, 0), v0=np.zeros(x0.size), fext=None,
                  gravity=(0.0, -9.81, 0.0), dt=dt, steps=steps, cg_tol=tol, seed=seed)


def config_linear_cpu(n=(10, 10, 10)):
    """Linear tets, E=1e5, nu=0.45, rho=1000; bottom fixed, gravity; linear implicit Euler, dt 1 ms, 100 steps."""
    return _tet_config("linear-tet-block", n, 0, 1e5, 0.45, 1000.0, 1e-3, 100, 1e-12, 0)


def config_corotational(n=(40, 40, 40), seed=2024):
    """Corotational tets, E=5e4, nu=0.45, rho=1000; bottom fixed; top-face nodes get tangential
    (x, z) velocities ~ N(0, 0.01 m/s); gravity; implicit Euler dt 1 ms, Jacobi-CG 1e-8, 200 steps."""
    cfg = _tet_config("corotational-tet-block", n, 1, 5e4, 0.45, 1000.0, 1e-3, 200, 1e-8, seed)
    top = face_nodes(cfg.nx, cfg.ny, cfg.nz, 1, 1)
    z = normal(2 * top.size, seed).reshape(-1, 2) * 0.01
    v0 = np.zeros((cfg.x0.shape[0], 3))
    v0[top, 0] = z[:, 0]
    v0[top, 2] = z[:, 1]
    cfg["v0"] = v0.reshape(-1)
    return cfg


def smoothed_lognormal_E(nx, ny, nz, median, sigma, seed):
    """Per-cube modulus exp(box-mean over the clamped 3x3x3 neighbourhood of ln(median) + sigma z)."""
    z = normal(nx * ny * nz, seed, stream=7).reshape(nx, ny, nz)
    lnE = math.log(median) + sigma * z
    pad = np.pad(lnE, 1, mode="edge")
    acc = np.zeros_like(lnE)
    cnt = np.zeros_like(lnE)
    ones = np.pad(np.ones_like(lnE), 1, mode="constant")
    padz = np.pad(lnE, 1, mode="constant")
    for di in range(3):
        for dj in range(3):
            for dk in range(3):
                acc += padz[di:di + nx, dj:dj + ny, dk:dk + nz]
                cnt += ones[di:di + nx, dj:dj + ny, dk:dk + nz]
    del pad
    return np.exp(acc / cnt).reshape(-1)


def config_adjoint(n=(20, 20, 20), seed=77):
    """Corotational tets with per-cube E (log-normal, median 3e4, sigma 0.5, 3-cube box filter), nu=0.45;
    bottom fixed; top pressure 200 Pa; gravity; dt 2 ms; 50 steps."""
    nx, ny, nz = n
    Ecube = smoothed_lognormal_E(nx, ny, nz, 3e4, 0.5, seed)
    cfg = _tet_config("adjoint-tet-block", n, 1, 1.0, 0.45, 1000.0, 2e-3, 50, 1e-8, seed)
    cfg["E"] = Ecube[cfg.cube_of_elem]
    cfg["E_cube"] = Ecube
    tris = top_face_triangles_tet(nx, ny, nz)
    cfg["fext"] = pressure_forces(cfg.x0, tris, lambda c: np.full(c.shape[0], 200.0))
    cfg["obs_nodes"] = observation_nodes(nx, ny, nz, 500, seed)
    return cfg


def surface_nodes(nx, ny, nz, exclude_bottom=True):
    """Node ids on the block's boundary (optionally without the y = 0 face)."""
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    on = (i == 0) | (i == nx) | (j == ny) | (k == 0) | (k == nz)
    if not exclude_bottom:
        on |= j == 0
    else:
        on &= j > 0
    return node_id(i[on], j[on], k[on], ny, nz).astype(np.int64)


def observation_nodes(nx, ny, nz, count, seed):
    """`count` distinct seeded surface nodes (the observed nodes of the identification problem)."""
    surf = surface_nodes(nx, ny, nz)
    key = uniform(surf.size, seed, stream=9)
    return np.sort(surf[np.argsort(key, kind="stable")[:min(count, surf.size)]])


def config_hex(n=(128, 64, 128), seed=5):
    """Hexahedra, E=2e4, nu=0.48, rho=1050; bottom fixed; gravity plus a Gaussian indentation
    p(r) = 500 exp(-r^2 / (2 sigma^2)) Pa, sigma = 8 cubes, seeded centre; dt 1 ms, 50 steps."""
    nx, ny, nz = n
    h = 0.01
    x0, hexes = block_hex_mesh(nx, ny, nz, h)
    Ne = hexes.shape[0]
    u = uniform(2, seed, stream=3)
    cx = (0.25 + 0.5 * u[0]) * nx * h
    cz = (0.25 + 0.5 * u[1]) * nz * h
    sig = 8 * h
    tris = top_face_triangles_hex(nx, ny, nz)
    fext = pressure_forces(x0, tris, lambda c: 500.0 * np.exp(-((c[:, 0] - cx) ** 2 + (c[:, 2] - cz) ** 2)
                                                              / (2 * sig * sig)))
    return Config(name="hex-block", kind=2, nx=nx, ny=ny, nz=nz, h=h, x0=x0, conn=hexes,
                  E=np.full(Ne, 2e4), nu=np.full(Ne, 0.48), rho=np.full(Ne, 1050.0),
                  bc_nodes=face_nodes(nx, ny, nz, 1, 0), v0=np.zeros(x0.size), fext=fext,
                  gravity=(0.0, -9.81, 0.0), dt=1e-3, steps=50, cg_tol=1e-6, seed=seed)


def concat_instances(cfgs):
    """One block-diagonal mesh from independent instance configs: node offsets, global connectivity,
    per-element materials, global boundary nodes, per-instance time steps."""
    off = [0]
    for c in cfgs:
        off.append(off[-1] + c.x0.shape[0])
    conn = np.concatenate([c.conn + off[i] for i, c in enumerate(cfgs)]).astype(np.int32)
    fext = None
    if any(c.fext is not None for c in cfgs):
        fext = np.concatenate([c.fext if c.fext is not None else np.zeros(c.x0.size) for c in cfgs])
    return Config(name="batch-%d" % len(cfgs), kind=cfgs[0].kind, x0=np.concatenate([c.x0 for c in cfgs]),
                  conn=conn, E=np.concatenate([c.E for c in cfgs]), nu=np.concatenate([c.nu for c in cfgs]),
                  rho=np.concatenate([c.rho for c in cfgs]),
                  bc_nodes=np.concatenate([c.bc_nodes + off[i] for i, c in enumerate(cfgs)]),
                  v0=np.concatenate([c.v0 for c in cfgs]), fext=fext, node_offsets=np.asarray(off, dtype=np.int64),
                  dt=np.asarray([c.dt for c in cfgs]), gravity=cfgs[0].gravity, steps=cfgs[0].steps,
                  cg_tol=cfgs[0].cg_tol, instances=cfgs)


def config_batched(count=64, seed=11, lo=6, hi=48):
    """Independent corotational block meshes with per-axis cube counts ~ U{lo..hi}, E ~ logU[1e3, 1e6],
    nu ~ U[0.30, 0.49], rho 1000, dt in {0.5, 1, 2} ms, one random fixed face each."""
    u = uniform(count * 8, seed, stream=5).reshape(count, 8)
    out = []
    for m in range(count):
        n = tuple(int(lo + min(hi - lo, math.floor(u[m, a] * (hi - lo + 1)))) for a in range(3))
        E = math.exp(math.log(1e3) + u[m, 3] * (math.log(1e6) - math.log(1e3)))
        nu = 0.30 + 0.19 * u[m, 4]
        dt = [0.5e-3, 1e-3, 2e-3][min(2, int(u[m, 5] * 3))]
        face = min(5, int(u[m, 6] * 6))
        cfg = _tet_config("batched-%d" % m, n, 1, E, nu, 1000.0, dt, 100, 1e-8, seed + m)
        cfg["bc_nodes"] = face_nodes(n[0], n[1], n[2], face // 2, face % 2)
        cfg["fixed_face"] = face
        out.append(cfg)
    return out
