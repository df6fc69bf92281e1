"""Synthetic voxel problems: geometry, fixed nodes and load cases.

The reference reads its problems from files (src/inout.cpp:6-128); these
generators produce the same three inputs -- a voxel occupancy grid (value 1 =
solid, x fastest, inout.cpp:16-36), a set of fixed *compact* node ids
(inout.cpp:95) and per-node load vectors (inout.cpp:113-125) -- for the
workloads this project measures.

Compact node ids follow FEM.cpp:88-99,179-184: the corners of solid voxels,
numbered in ascending grid-linear order x + y*(X+1) + z*(X+1)*(Y+1).

Random draws use numpy's PCG64 (``np.random.default_rng(seed)``), in the order
documented on each function, so a (seed, size) pair always gives the same
problem on any host.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    voxels: np.ndarray            # (nz, ny, nx) int32, 1 = solid
    fixed: np.ndarray             # (nfixed,) int64 compact node ids
    loads: np.ndarray             # (nrhs, ndof) float64 load vectors in compact DOF order
    name: str = ""
    info: dict = field(default_factory=dict)

    @property
    def ndof(self):
        return self.loads.shape[1]


def ke0(E=1.0, nu=0.3):
    """24x24 element stiffness of the unit hex8 voxel (trilinear shape
    functions, 2x2x2 Gauss, isotropic E, nu) in the DOF frame of the
    reference's data/Ke0.dat: local node L = x + 2y + 4z, and each node's DOF
    triple is (u_y, -u_x, u_z) of the grid axes (scripts/outputToC.m:13-14
    reorders the exporter's nodes into this frame)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.full((3, 3), lam) + 2 * mu * np.eye(3)
    D = np.block([[D, np.zeros((3, 3))], [np.zeros((3, 3)), mu * np.eye(3)]])
    g = 0.5 / np.sqrt(3.0)
    pts = np.array([0.5 - g, 0.5 + g])
    corner = np.array([[L & 1, (L >> 1) & 1, (L >> 2) & 1] for L in range(8)], float)
    K = np.zeros((24, 24))
    for q in np.array(np.meshgrid(pts, pts, pts, indexing="ij")).reshape(3, -1).T:
        f = np.where(corner == 1, q, 1 - q)             # per-axis linear factors (8, 3)
        s = np.where(corner == 1, 1.0, -1.0)            # their derivatives
        dN = np.stack([s[:, 0] * f[:, 1] * f[:, 2], f[:, 0] * s[:, 1] * f[:, 2], f[:, 0] * f[:, 1] * s[:, 2]], 1)
        B = np.zeros((6, 24))
        for a in range(3):
            B[a, a::3] = dN[:, a]
        for row, (a, b) in zip((3, 4, 5), ((0, 1), (1, 2), (0, 2))):
            B[row, a::3] = dN[:, b]
            B[row, b::3] = dN[:, a]
        K += B.T @ D @ B / 8.0
    R = np.array([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    T = np.kron(np.eye(8), R)
    return T @ K @ T.T


def node_occupancy(voxels):
    """(nz+1, ny+1, nx+1) bool: grid nodes that are a corner of a solid voxel."""
    v = np.asarray(voxels) == 1
    nz, ny, nx = v.shape
    occ = np.zeros((nz + 1, ny + 1, nx + 1), bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                occ[dz:dz + nz, dy:dy + ny, dx:dx + nx] |= v
    return occ


def compact_ids(voxels):
    """(nz+1, ny+1, nx+1) int64 compact id of every used grid node, -1 elsewhere."""
    occ = node_occupancy(voxels)
    ids = np.full(occ.shape, -1, np.int64)
    flat = occ.reshape(-1)
    ids.reshape(-1)[flat] = np.arange(int(flat.sum()), dtype=np.int64)
    return ids


def apply_boundary_conditions(ndof, loaded_nodes, load_vectors):
    """f[3i+c] = load of compact node i (FEM.cpp:568-584, MATRIX_FREE branch)."""
    f = np.zeros(ndof)
    loaded_nodes = np.asarray(loaded_nodes, np.int64)
    load_vectors = np.asarray(load_vectors, np.float64).reshape(-1, 3)
    for c in range(3):
        f[3 * loaded_nodes + c] = load_vectors[:, c]
    return f


def cantilever_box(nx=32, ny=16, nz=16, total_force=1.0):
    """Solid nx*ny*nz box; every node of the x=0 face fixed; a total force
    `total_force` in -z split equally over the nodes of the x=nx face.

    The z displacement is DOF component 2 in the frame of data/Ke0.dat."""
    vox = np.ones((nz, ny, nx), np.int32)
    ids = compact_ids(vox)
    fixed = np.sort(ids[:, :, 0].reshape(-1))
    loaded = np.sort(ids[:, :, nx].reshape(-1))
    ndof = 3 * int(ids.max() + 1)
    lv = np.zeros((loaded.size, 3))
    lv[:, 2] = -total_force / loaded.size
    f = apply_boundary_conditions(ndof, loaded, lv)
    return Problem(vox, fixed, f[None, :], name=f"box{nx}x{ny}x{nz}",
                   info={"nfixed": int(fixed.size), "nloaded": int(loaded.size)})


def _label6(mask):
    """6-connected components of a 3-D bool array (scipy.ndimage.label)."""
    from scipy import ndimage
    structure = ndimage.generate_binary_structure(3, 1)
    return ndimage.label(mask, structure=structure)


def ellipsoid_union(n, count, amin, amax, seed):
    """Union of `count` axis-aligned ellipsoids in an n^3 voxel grid.

    Draw order (PCG64, default_rng(seed)), per ellipsoid i = 0..count-1:
    centre (cx, cy, cz) ~ U[0.1 n, 0.9 n]^3, then semi-axes (ax, ay, az) ~
    U[amin, amax]^3.  Voxel (x, y, z) is solid when
    sum_d ((p_d + 0.5 - c_d) / a_d)^2 <= 1.  Only the 6-connected components
    that touch the z = 0 face are kept: every other component would float and
    make the stiffness matrix singular."""
    rng = np.random.default_rng(seed)
    vox = np.zeros((n, n, n), bool)
    ax = np.arange(n) + 0.5
    for _ in range(count):
        c = rng.uniform(0.1 * n, 0.9 * n, 3)
        a = rng.uniform(amin, amax, 3)
        lo = np.maximum(np.floor(c - a - 1).astype(int), 0)
        hi = np.minimum(np.ceil(c + a + 1).astype(int), n)
        xs = ((ax[lo[0]:hi[0]] - c[0]) / a[0]) ** 2
        ys = ((ax[lo[1]:hi[1]] - c[1]) / a[1]) ** 2
        zs = ((ax[lo[2]:hi[2]] - c[2]) / a[2]) ** 2
        inside = (zs[:, None, None] + ys[None, :, None] + xs[None, None, :]) <= 1.0
        vox[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] |= inside
    labels, ncomp = _label6(vox)
    touching = np.unique(labels[0][labels[0] > 0])
    keep = np.isin(labels, touching)
    return keep.astype(np.int32), {"components": int(ncomp), "kept_components": int(touching.size),
                                   "fill": float(keep.mean())}


def ellipsoid_problem(n, count, amin, amax, seed, total_force=1.0):
    """Ellipsoid-union geometry; every used node on z = 0 fixed; a total force
    `total_force` in -z spread equally over the used nodes of the two topmost
    node layers that contain used nodes."""
    vox, info = ellipsoid_union(n, count, amin, amax, seed)
    if vox.sum() == 0:
        raise ValueError("no solid component touches z = 0")
    ids = compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    zs = np.nonzero((ids >= 0).any(axis=(1, 2)))[0]
    ztop = zs.max()
    top = ids[max(ztop - 1, 0):ztop + 1]
    loaded = np.sort(top[top >= 0])
    ndof = 3 * int(ids.max() + 1)
    lv = np.zeros((loaded.size, 3))
    lv[:, 2] = -total_force / loaded.size
    f = apply_boundary_conditions(ndof, loaded, lv)
    info.update({"nfixed": int(fixed.size), "nloaded": int(loaded.size), "ndof": ndof})
    return Problem(vox, fixed, f[None, :], name=f"ellipsoids{n}_s{seed}", info=info)


def patch_load_cases(voxels, nrhs, seed, patch=8, total_force=1.0):
    """`nrhs` load cases on one geometry.  Draw order (PCG64, default_rng(seed)),
    per case: direction d ~ N(0, I_3) normalised (components in the DOF frame of
    Ke0), then the patch corner (x0, y0) ~ U{0..nx+1-patch} x U{0..ny+1-patch}.
    The loaded nodes are, for each node column (x, y) of the patch x patch
    window, the topmost used node of that column; the force `total_force * d`
    is split equally over them.  A window without used nodes is redrawn."""
    rng = np.random.default_rng(seed)
    ids = compact_ids(voxels)
    nzn, nyn, nxn = ids.shape
    used = ids >= 0
    has = used.any(axis=0)
    if not has.any():
        raise ValueError("the voxel model has no solid voxel")
    ztop = nzn - 1 - np.argmax(used[::-1], axis=0)
    ndof = 3 * int(ids.max() + 1)
    F = np.zeros((nrhs, ndof))
    for k in range(nrhs):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        while True:
            x0 = int(rng.integers(0, nxn - patch + 1))
            y0 = int(rng.integers(0, nyn - patch + 1))
            sel = has[y0:y0 + patch, x0:x0 + patch]
            if sel.any():
                break
        yy, xx = np.nonzero(sel)
        yy += y0
        xx += x0
        nodes = ids[ztop[yy, xx], yy, xx]
        lv = np.tile(d * total_force / nodes.size, (nodes.size, 1))
        F[k] = apply_boundary_conditions(ndof, nodes, lv)
    return F


def batched_problem(n, count, amin, amax, seed, nrhs=16, rhs_seed=None):
    """Ellipsoid geometry with the bottom fixed and `nrhs` patch load cases."""
    vox, info = ellipsoid_union(n, count, amin, amax, seed)
    if vox.sum() == 0:
        raise ValueError("no solid component touches z = 0")
    ids = compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    F = patch_load_cases(vox, nrhs, seed if rhs_seed is None else rhs_seed)
    info.update({"nfixed": int(fixed.size), "ndof": int(F.shape[1]), "nrhs": nrhs})
    return Problem(vox, fixed, F, name=f"batched{n}_s{seed}_r{nrhs}", info=info)


# The measured workloads, by the sizes that define them.
def workload(name):
    if name == "box32x16x16":
        return cantilever_box(32, 16, 16)
    if name == "ellipsoids128":
        return ellipsoid_problem(128, 60, 6.0, 30.0, seed=1)
    if name == "ellipsoids256":
        return ellipsoid_problem(256, 120, 12.0, 60.0, seed=2)
    if name == "batched128x16":
        return batched_problem(128, 60, 6.0, 30.0, seed=3, nrhs=16)
    raise KeyError(name)


WORKLOAD_LEVELS = {"box32x16x16": 3, "ellipsoids128": 5, "ellipsoids256": 6, "batched128x16": 5}
