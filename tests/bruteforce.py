"""Brute-force restatements used to pin the oracle: explicit sparse matrices
built with numpy/scipy from the definitions (global triplet assembly, explicit
trilinear prolongation matrices), sharing no code with oracle/vfem_oracle.c or
the HIP library.  Small grids only."""
import numpy as np
import scipy.sparse as sp

from vulkanfem_amd import problems


def sym_upper(ke):
    return np.triu(ke) + np.triu(ke, 1).T


def assemble(occ, ke, scale=1.0, ids=None):
    """Global K over the elements occ == 1 (local node L = x + 2y + 4z); node
    numbering `ids` (default: the compact numbering of occ itself)."""
    if ids is None:
        ids = problems.compact_ids(occ)
    ez, ey, ex = np.nonzero(np.asarray(occ) == 1)
    nodes = np.stack([ids[ez + ((L >> 2) & 1), ey + ((L >> 1) & 1), ex + (L & 1)] for L in range(8)], axis=1)
    dofs = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, 24)).reshape(-1)
    vals = np.tile(scale * ke.reshape(-1), dofs.shape[0])
    n = 3 * int(ids.max() + 1)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def coarsen(occ, div=2):
    """Coarse elements: floor(fine / div), present if any child is (FEM.cpp:274-283)."""
    occ = np.asarray(occ) == 1
    nz, ny, nx = occ.shape
    cz, cy, cx = (nz + div - 1) // div, (ny + div - 1) // div, (nx + div - 1) // div
    pad = np.zeros((cz * div, cy * div, cx * div), bool)
    pad[:nz, :ny, :nx] = occ
    return pad.reshape(cz, div, cy, div, cx, div).any(axis=(1, 3, 5)).astype(np.int32)


def node_coords(occ):
    ids = problems.compact_ids(occ)
    z, y, x = np.nonzero(ids >= 0)
    order = np.argsort(ids[z, y, x])
    return np.stack([x[order], y[order], z[order]], axis=1), ids


def prolongation(fine_occ, coarse_occ, div=2):
    """Explicit trilinear P (fine DOFs x coarse DOFs) on used nodes:
    P[f, N] = prod_d (1 - |f_d - div N_d| / div) for |f - div N|_inf < div."""
    fc, _ = node_coords(fine_occ)
    _, cids = node_coords(coarse_occ)
    rows, cols, vals = [], [], []
    for i, (x, y, z) in enumerate(fc):
        cand = []
        for p in (x, y, z):
            lo = p // div
            cand.append([lo] if p % div == 0 else [lo, lo + 1])
        for Z in cand[2]:
            for Y in cand[1]:
                for X in cand[0]:
                    if Z >= cids.shape[0] or Y >= cids.shape[1] or X >= cids.shape[2]:
                        continue
                    j = cids[Z, Y, X]
                    if j < 0:
                        continue
                    w = (1 - abs(x - div * X) / div) * (1 - abs(y - div * Y) / div) * (1 - abs(z - div * Z) / div)
                    if w > 0:
                        rows.append(i)
                        cols.append(j)
                        vals.append(w)
    Pn = sp.csr_matrix((vals, (rows, cols)), shape=(fc.shape[0], int(cids.max() + 1)))
    return sp.kron(Pn, sp.identity(3), format="csr")


def fixed_dof_mask(nnodes, fixed_nodes):
    m = np.zeros(3 * nnodes, bool)
    for c in range(3):
        m[3 * np.asarray(fixed_nodes, np.int64) + c] = True
    return m


def coarse_fixed(fine_occ, coarse_occ, fixed_fine, div=2):
    """Coarse nodes with a fixed fine node strictly inside their support."""
    fc, _ = node_coords(fine_occ)
    _, cids = node_coords(coarse_occ)
    out = set()
    for f in np.asarray(fixed_fine, np.int64):
        x, y, z = fc[f]
        cand = []
        for p in (x, y, z):
            lo = p // div
            cand.append([lo] if p % div == 0 else [lo, lo + 1])
        for Z in cand[2]:
            for Y in cand[1]:
                for X in cand[0]:
                    if Z < cids.shape[0] and Y < cids.shape[1] and X < cids.shape[2] and cids[Z, Y, X] >= 0:
                        out.add(int(cids[Z, Y, X]))
    return np.array(sorted(out), np.int64)


def full_parents(fine_occ, coarse_occ, div):
    """Coarse elements all of whose div^3 children are present fine elements."""
    f = np.asarray(fine_occ) == 1
    cz, cy, cx = np.asarray(coarse_occ).shape
    pad = np.zeros((cz * div, cy * div, cx * div), bool)
    pad[:f.shape[0], :f.shape[1], :f.shape[2]] = f
    return pad.reshape(cz, div, cy, div, cx, div).all(axis=(1, 3, 5))


class Hierarchy:
    """Explicit-matrix hierarchy: occupancies, P_l, fixed sets, the level
    operators and the reference coarsest matrix Kc (FEM.cpp:513).  Level 0: K.
    Level 1: P^T K P.  Level l >= 2, the reference's element rule
    (FEM.cpp:364-395,463): a full coarse element is 2^(l+skip) Ke, a partial
    one the Galerkin sum of its present children taken at 2^(l-1) Ke.  skip = 1
    makes level 1 a x4 coarsening (FEM.cpp:266-272)."""

    def __init__(self, voxels, ke, fixed, nlev, skip=0):
        ke = sym_upper(ke)
        self.occ = [np.asarray(voxels, np.int32)]
        self.fixed = [np.asarray(fixed, np.int64)]
        self.P = []
        self.A = [assemble(self.occ[0], ke)]
        for l in range(1, nlev):
            div = 4 if (l == 1 and skip) else 2
            fine = self.occ[-1]
            c = coarsen(fine, div)
            P = prolongation(fine, c, div)
            self.P.append(P)
            if l == 1:
                A = P.T @ self.A[0] @ P
            else:
                full = full_parents(fine, c, div)
                kids = np.array(fine == 1)
                nz, ny, nx = kids.shape
                par = full[np.arange(nz)[:, None, None] // div, np.arange(ny)[None, :, None] // div,
                           np.arange(nx)[None, None, :] // div]
                kids &= ~par
                A = (P.T @ assemble(kids.astype(np.int32), ke, 2.0 ** (l - 1), problems.compact_ids(fine)) @ P
                     + assemble(full.astype(np.int32), ke, 2.0 ** (l + skip), problems.compact_ids(c)))
            self.A.append(A.tocsr())
            self.fixed.append(coarse_fixed(fine, c, self.fixed[-1], div))
            self.occ.append(c)
        self.nlev = nlev
        self.kc_full = assemble(self.occ[-1], ke, 2.0 ** (nlev - 1 + skip)).toarray()
        nn = self.kc_full.shape[0] // 3
        self.free = np.nonzero(~fixed_dof_mask(nn, self.fixed[-1]))[0]

    def invdiag(self, l):
        return 1.0 / self.A[l].diagonal()

    def kc(self):
        return self.kc_full[np.ix_(self.free, self.free)]

    def vcycle_additive(self, r):
        """z = Pi [sum_l P_{0<-l} 2 D_l^-1 R_{l<-0} r + P_{0<-L-1} Kc^-1 R r], R = P^T
        (MultigridPreconditioner.hpp:158-193 with mu cancelled)."""
        L = self.nlev
        rs = [r]
        for l in range(L - 1):
            rs.append(self.P[l].T @ rs[-1])
        xc = np.zeros_like(rs[-1])
        xc[self.free] = np.linalg.solve(self.kc(), rs[-1][self.free])
        x = xc
        for l in range(L - 2, -1, -1):
            x = self.P[l] @ x + 2 * self.invdiag(l) * rs[l]
        x[fixed_dof_mask(len(x) // 3, self.fixed[0])] = 0
        return x

    def level_operator(self, l):
        """A_l with the fixed rows zeroed: the operator of the damped cycle on
        level l (MatrixFreeSparse.hpp:210-213 applied on every level)."""
        keep = ~fixed_dof_mask(self.A[l].shape[0] // 3, self.fixed[l])
        return (sp.diags(keep.astype(np.float64)) @ self.A[l]).tocsr()

    def kc_galerkin(self):
        """Free-DOF block of the coarsest Galerkin operator (the damped cycle's Kc)."""
        return self.A[-1].toarray()[np.ix_(self.free, self.free)]

    def chebyshev_error(self, l, e, degree, ratio, lmax):
        """Error after `degree` Chebyshev smoothing steps on level l, in closed form:
        e -> T_m((theta I - M) / delta) e / T_m(theta / delta) on the free DOFs,
        M = D^-1 A, [theta - delta, theta + delta] = [lmax / ratio, lmax] (the
        error-reduction polynomial of Chebyshev iteration, Saad Sec. 12.1), built
        with the polynomial's own three-term recurrence T_k+1 = 2 y T_k - T_k-1."""
        free = ~fixed_dof_mask(len(e) // 3, self.fixed[l])
        Aff = self.A[l].tocsr()[free][:, free]
        dff = self.invdiag(l)[free]
        lo, hi = lmax / ratio, lmax
        theta, delta = (hi + lo) / 2, (hi - lo) / 2
        Y = lambda v: (theta * v - dff * (Aff @ v)) / delta  # noqa: E731
        t0, t1 = e[free], Y(e[free])
        for _ in range(degree - 1):
            t0, t1 = t1, 2 * Y(t1) - t0
        tm = t1 if degree >= 1 else t0
        out = np.zeros_like(e)
        out[free] = tm / np.polynomial.chebyshev.chebval(theta / delta, [0] * degree + [1])
        return out

    def vcycle_damped(self, r, omega, cheb=None):
        """Multiplicative V(1,1) with damped Jacobi written out with explicit
        matrices: x = w D^-1 r; x += P (cycle of P^T (r - A x)); x += w D^-1 (r - A x);
        fixed DOFs held at zero on every level; exact solve on the coarsest free DOFs.
        cheb = (level, degree, ratio, lmax per level): on levels >= level the
        smoothing before and after the correction is `degree` Chebyshev steps,
        applied through their error polynomial (chebyshev_error) and the exact
        level solution x* = A_ff^-1 r_f."""
        L = self.nlev

        def level(l, rl):
            fm = fixed_dof_mask(len(rl) // 3, self.fixed[l])
            if l == L - 1:
                x = np.zeros_like(rl)
                x[self.free] = np.linalg.solve(self.kc_galerkin(), rl[self.free])
                return x
            A = self.level_operator(l)
            d = self.invdiag(l)
            ch = cheb is not None and cheb[0] > 0 and l >= cheb[0]
            if ch:
                xs = np.zeros_like(rl)
                xs[~fm] = np.linalg.solve(self.A[l].toarray()[np.ix_(~fm, ~fm)], rl[~fm])
                smooth = lambda x0: xs - self.chebyshev_error(l, xs - x0, cheb[1], cheb[2], cheb[3][l])  # noqa: E731
                x = smooth(np.zeros_like(rl))
            else:
                x = omega * d * rl
            x[fm] = 0
            rc = self.P[l].T @ (rl - A @ x)
            rc[fixed_dof_mask(len(rc) // 3, self.fixed[l + 1])] = 0
            t = self.P[l] @ level(l + 1, rc)
            t[fm] = 0
            x = x + t
            x = smooth(x) if ch else x + omega * d * (rl - A @ x)
            x[fm] = 0
            return x

        r0 = np.array(r, dtype=np.float64)
        r0[fixed_dof_mask(len(r0) // 3, self.fixed[0])] = 0
        return level(0, r0)
