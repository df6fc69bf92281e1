"""Writes the golden fixtures under tests/golden/.  Run from the repo root:

    python tests/golden/make_fixtures.py

Two kinds of fixture are written:

* reference_tables.npz -- data the reference ships, read as data:
  data/Ke0.dat (the 24x24 element stiffness, read as CSV like inout.cpp:60-69)
  and the numeric arrays of include/RestrictionOperator.hpp, reshaped with the
  Eigen Map semantics the reference applies to them (FEM.cpp:228-233).
  Needs the reference checkout (default /root/reference, or $VFEM_REFERENCE).
* scipy_goldens.npz -- outputs of an independent scipy assembly of the global
  stiffness matrix (plain triplet assembly of the element matrix over solid
  voxels, written here, sharing no code with the oracle or the product):
  matvec goldens K x (fixed rows zeroed, MatrixFreeSparse.hpp:210-213) on small
  seeded geometries, and the direct sparse solve (SuperLU) of the 32x16x16
  cantilever, the reference's own problem class.
"""
import os
import re
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vulkanfem_amd import problems  # noqa: E402  (input generation only)

REF = os.environ.get("VFEM_REFERENCE", "/root/reference")


def parse_arrays(text):
    out = {}
    for m in re.finditer(r"std::array<\s*(int|double)\s*,\s*(\d+)\s*>\s+(\w+)\s*=\s*\{([^}]*)\}", text):
        typ, n, name, body = m.group(1), int(m.group(2)), m.group(3), m.group(4)
        vals = [v for v in re.split(r"[\s,]+", body.strip()) if v]
        arr = np.array([float(v) for v in vals])
        assert arr.size == n, (name, arr.size, n)
        out[name] = arr.astype(np.int32) if typ == "int" else arr
    return out


def reference_tables():
    ke0 = np.loadtxt(os.path.join(REF, "data", "Ke0.dat"), delimiter=",")
    with open(os.path.join(REF, "include", "RestrictionOperator.hpp")) as fh:
        arr = parse_arrays(fh.read())
    # Eigen::Map of a column-major 8x8 array, used as (child, local node)
    map8 = arr["restrictionOnElementValues"].reshape(8, 8).T
    # Map<8x64>.transpose() -> (child, local node) = data[L + 8 child]
    map64 = arr["restrictionOnElement4Values"].reshape(64, 8)
    r27 = arr["restrictionOperatorValues"].reshape(27, 8)
    r125 = arr["restrictionOperator4Values"].reshape(125, 8)
    rdof = arr["restrictionOperatorDoFValues"].reshape(81, 24)
    rdof4 = arr["restrictionOperatorDoF4Values"].reshape(375, 24)
    np.savez_compressed(os.path.join(HERE, "reference_tables.npz"), ke0=ke0, map8=map8, map64=map64, r27=r27,
                        r125=r125, rdof=rdof, rdof4=rdof4)


def assemble(voxels, ke):
    """Global K over solid voxels; local node L = x + 2y + 4z, DOF 3*id + c."""
    ids = problems.compact_ids(voxels)
    ez, ey, ex = np.nonzero(np.asarray(voxels) == 1)
    nodes = np.stack([ids[ez + ((L >> 2) & 1), ey + ((L >> 1) & 1), ex + (L & 1)] for L in range(8)], axis=1)
    dofs = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, 24)).reshape(-1)
    vals = np.tile(ke.reshape(-1), dofs.shape[0])
    n = 3 * int(ids.max() + 1)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def scipy_goldens(ke):
    sym = np.triu(ke) + np.triu(ke, 1).T  # upper triangle is what the reference reads (FEM.cpp:200-210)
    out = {}
    # matvec goldens on small seeded geometries
    cases = [(10, 5, 2.0, 5.0, 11), (12, 8, 2.0, 6.0, 12), (9, 4, 2.0, 4.0, 13)]
    for k, (n, cnt, a0, a1, seed) in enumerate(cases):
        vox, _ = problems.ellipsoid_union(n, cnt, a0, a1, seed)
        ids = problems.compact_ids(vox)
        fixed = np.sort(ids[0][ids[0] >= 0])
        K = assemble(vox, sym)
        rng = np.random.default_rng(100 + k)
        x = rng.standard_normal(K.shape[0])
        y = K @ x
        for c in range(3):
            y[3 * fixed + c] = 0.0
        out[f"mv{k}_voxels"] = vox
        out[f"mv{k}_fixed"] = fixed
        out[f"mv{k}_x"] = x
        out[f"mv{k}_y"] = y
    # direct solve of the 32x16x16 cantilever
    p = problems.cantilever_box(32, 16, 16)
    K = assemble(p.voxels, sym)
    f = p.loads[0]
    isfixed = np.zeros(K.shape[0], bool)
    for c in range(3):
        isfixed[3 * p.fixed + c] = True
    free = np.nonzero(~isfixed)[0]
    u = np.zeros(K.shape[0])
    u[free] = spla.spsolve(K[free][:, free].tocsc(), f[free])
    res = np.linalg.norm(K[free][:, free] @ u[free] - f[free]) / np.linalg.norm(f[free])
    out["box_u"] = u
    out["box_f"] = f
    out["box_fixed"] = p.fixed
    out["box_relres"] = np.array(res)
    out["box_compliance"] = np.array(f @ u)
    np.savez_compressed(os.path.join(HERE, "scipy_goldens.npz"), **out)
    print("box direct solve: rel residual %.3e, compliance %.10e" % (res, f @ u))


if __name__ == "__main__":
    reference_tables()
    tabs = np.load(os.path.join(HERE, "reference_tables.npz"))
    scipy_goldens(tabs["ke0"])
    print("wrote", os.listdir(HERE))
