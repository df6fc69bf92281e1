"""The CPU oracle (oracle/vfem_oracle.c) pinned against the reference's own data
files, closed forms and brute-force explicit-matrix restatements (CPU only)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import bruteforce as BF
from vulkanfem_amd import problems


def test_ke0_closed_form_matches_reference_data(ke0):
    # data/Ke0.dat is the E=1, nu=0.3 trilinear hex8 matrix in the rotated DOF frame
    assert np.abs(O.ke0_closed_form(1.0, 0.3) - ke0).max() < 1e-15
    # closed-form invariants: symmetric, rigid translations in the null space
    K = O.ke0_closed_form(2.5, 0.2)
    assert np.allclose(K, K.T, atol=1e-15)
    for c in range(3):
        t = np.zeros(24)
        t[c::3] = 1
        assert np.abs(K @ t).max() < 1e-14
    assert np.sum(np.abs(np.linalg.eigvalsh(K)) < 1e-12) == 6


def test_transfer_tables_match_reference(tables):
    assert np.array_equal(O.transfer_weights(2), tables["r27"])
    assert np.array_equal(O.transfer_weights(4), tables["r125"])
    assert np.array_equal(O.child_lattice_map(2), tables["map8"])
    assert np.array_equal(O.child_lattice_map(4), tables["map64"])
    assert np.array_equal(np.kron(O.transfer_weights(2), np.eye(3)), tables["rdof"])
    assert np.array_equal(np.kron(O.transfer_weights(4), np.eye(3)), tables["rdof4"])


@pytest.mark.parametrize("k", [0, 1, 2])
def test_matvec_matches_scipy_assembly(goldens, ke0, k):
    s = O.OracleSystem(goldens[f"mv{k}_voxels"], ke0, goldens[f"mv{k}_fixed"], 2)
    y = s.matvec(goldens[f"mv{k}_x"])
    ref = goldens[f"mv{k}_y"]
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_matvec_rigid_modes_and_symmetry(ke0):
    vox, _ = problems.ellipsoid_union(12, 6, 2.0, 6.0, 22)
    ids = problems.compact_ids(vox)
    # rigid translations are in the null space of K, so also of K with fixed rows zeroed
    s = O.OracleSystem(vox, ke0, np.sort(ids[0][ids[0] >= 0]), 2)
    for c in range(3):
        t = np.zeros(s.ndof)
        t[c::3] = 1
        assert np.abs(s.matvec(t)).max() < 1e-13
    # symmetric on vectors that vanish at fixed DOFs (the CG subspace)
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(s.ndof), rng.standard_normal(s.ndof)
    mask = np.repeat(s.fixed(0), 3)
    a[mask] = 0
    b[mask] = 0
    assert abs(a @ s.matvec(b) - b @ s.matvec(a)) < 1e-12 * abs(a @ s.matvec(b))


@pytest.fixture(scope="module")
def small_case(ke0):
    vox, _ = problems.ellipsoid_union(14, 8, 2.0, 6.0, 32)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    return vox, fixed


def test_hierarchy_matches_brute_force(ke0, small_case):
    vox, fixed = small_case
    s = O.OracleSystem(vox, ke0, fixed, 3)
    bf = BF.Hierarchy(vox, ke0, fixed, 3)
    assert s.num_levels == 3
    for l in range(3):
        assert s.nodes(l) == bf.A[l].shape[0] // 3
        # invDiag: level 0 = diag(K); level 1 = diag(P^T K P) (Galerkin, exact);
        # level 2 = Galerkin of the level-1 elements taken full (the reference's unique-K rule)
        np.testing.assert_allclose(s.invdiag(l), bf.invdiag(l), rtol=1e-13)
        fx = np.zeros(s.nodes(l), bool)
        fx[bf.fixed[l]] = True
        assert np.array_equal(s.fixed(l), fx)
    np.testing.assert_allclose(s.kc(0), bf.kc(), rtol=1e-12, atol=1e-12 * np.abs(bf.kc()).max())


def test_transfers_adjoint_and_reproduce_linear_fields(ke0, small_case):
    vox, fixed = small_case
    s = O.OracleSystem(vox, ke0, fixed, 3)
    rng = np.random.default_rng(1)
    for l in range(2):
        r = rng.standard_normal(3 * s.nodes(l))
        xc = rng.standard_normal(3 * s.nodes(l + 1))
        assert abs(s.restrict(l, r) @ xc - r @ s.interpolate(l, xc)) < 1e-12 * np.abs(r).sum()
    # interpolation of a coarse linear field is the fine linear field (pins mu = 1/multiplicity)
    bf = BF.Hierarchy(vox, ke0, fixed, 2)
    fc, _ = BF.node_coords(bf.occ[0])
    cc, _ = BF.node_coords(bf.occ[1])
    a = np.array([0.3, -1.2, 0.7])
    xc = np.repeat(cc @ a + 2.0, 3)
    xf = s.interpolate(0, xc)
    np.testing.assert_allclose(xf, np.repeat(fc @ a / 2 + 2.0, 3), rtol=1e-13, atol=1e-13)


def test_additive_vcycle_matches_explicit_formula(ke0, small_case):
    vox, fixed = small_case
    s = O.OracleSystem(vox, ke0, fixed, 3)
    bf = BF.Hierarchy(vox, ke0, fixed, 3)
    r = np.random.default_rng(2).standard_normal(s.ndof)
    z = s.vcycle(r, O.REF_ADDITIVE)
    zb = bf.vcycle_additive(r)
    np.testing.assert_allclose(z, zb, rtol=1e-10, atol=1e-12 * np.abs(zb).max())


@pytest.fixture(scope="module")
def damped_cases(ke0, small_case):
    """The 14^3 ellipsoid union (3 levels) and a skip-free 16^3 one (3 levels)."""
    vox16, _ = problems.ellipsoid_union(16, 9, 2.0, 6.5, 77)
    ids = problems.compact_ids(vox16)
    return [small_case, (vox16, np.sort(ids[0][ids[0] >= 0]))]


def test_damped_level_operators_match_galerkin_brute_force(ke0, damped_cases):
    """The damped cycle's coarse operators: level l >= 1 sums, per element, the
    unique Galerkin K of its children (taken full with 2^(l-1) Ke,
    FEM.cpp:364-395,463) or 2^l Ke on a full element; both must equal the
    explicit P^T A P of the brute-force hierarchy, fixed rows zeroed.  A wrong
    coarse scale, child pattern or transfer pair fails here."""
    rng = np.random.default_rng(11)
    for vox, fixed in damped_cases:
        s = O.OracleSystem(vox, ke0, fixed, 3)
        bf = BF.Hierarchy(vox, ke0, fixed, 3)
        for l in range(1, 3):
            x = rng.standard_normal(3 * s.nodes(l))
            ref = bf.level_operator(l) @ x
            y = s.level_matvec(l, x)
            assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max(), l
        kc = bf.kc_galerkin()
        assert np.abs(s.kc(1) - kc).max() <= 1e-12 * np.abs(kc).max()
        s.close()


@pytest.mark.parametrize("omega", [0.6, 0.5])
def test_damped_vcycle_matches_explicit_v11(ke0, damped_cases, omega):
    """vcycle_damped_level against V(1,1) written out with the brute-force
    P, A and 1/diag(A) (no shared code)."""
    rng = np.random.default_rng(12)
    for vox, fixed in damped_cases:
        s = O.OracleSystem(vox, ke0, fixed, 3)
        bf = BF.Hierarchy(vox, ke0, fixed, 3)
        r = rng.standard_normal(s.ndof)
        z = s.vcycle(r, O.DAMPED, omega)
        zb = bf.vcycle_damped(r, omega)
        assert np.abs(z - zb).max() <= 1e-12 * np.abs(zb).max()
        s.close()


def exact_lmax(bf, l):
    """lambda_max(D^-1 A) on the free DOFs of level l, by a dense symmetric eigensolve
    of D^-1/2 A D^-1/2."""
    free = ~BF.fixed_dof_mask(bf.A[l].shape[0] // 3, bf.fixed[l])
    A = bf.A[l].toarray()[np.ix_(free, free)]
    s = 1 / np.sqrt(A.diagonal())
    return float(np.linalg.eigvalsh(s[:, None] * A * s[None, :]).max())


@pytest.mark.parametrize("degree,ratio", [(2, 30.0), (3, 8.0)])
def test_damped_chebyshev_cycle_matches_closed_form(ke0, damped_cases, degree, ratio):
    """Chebyshev smoothing on the coarse levels of the damped cycle: the
    oracle's three-term Chebyshev iteration (Saad Alg. 12.1) against the
    closed-form error polynomial T_m((theta - D^-1 A) / delta) / T_m(theta / delta)
    applied with explicit matrices (4 levels, Chebyshev on levels 1 and 2).  A
    wrong recurrence coefficient or a missed first-step shortcut fails here."""
    rng = np.random.default_rng(14)
    for vox, fixed in damped_cases:
        s = O.OracleSystem(vox, ke0, fixed, 4)
        bf = BF.Hierarchy(vox, ke0, fixed, 4)
        lmax = [0.0] + [1.1 * exact_lmax(bf, l) for l in (1, 2)] + [0.0]
        s.set_chebyshev(1, degree, ratio, lmax)
        r = rng.standard_normal(s.ndof)
        z = s.vcycle(r, O.DAMPED, 0.5)
        zb = bf.vcycle_damped(r, 0.5, cheb=(1, degree, ratio, lmax))
        assert np.abs(z - zb).max() <= 1e-10 * np.abs(zb).max()
        # the Jacobi cycle is a different preconditioner
        s.set_chebyshev(0)
        assert np.abs(s.vcycle(r, O.DAMPED, 0.5) - zb).max() > 1e-3 * np.abs(zb).max()
        s.close()


def test_damped_chebyshev_cycle_is_symmetric_positive(ke0, small_case):
    """With the same polynomial before and after the correction the cycle stays
    a symmetric positive definite preconditioner (plain CG needs that)."""
    vox, fixed = small_case
    s = O.OracleSystem(vox, ke0, fixed, 3)
    bf = BF.Hierarchy(vox, ke0, fixed, 3)
    s.set_chebyshev(1, 2, 30.0, [0.0, 1.1 * exact_lmax(bf, 1), 0.0])
    rng = np.random.default_rng(15)
    mask = np.repeat(s.fixed(0), 3)
    a, b = rng.standard_normal(s.ndof), rng.standard_normal(s.ndof)
    a[mask] = 0
    b[mask] = 0
    Ma, Mb = s.vcycle(a, O.DAMPED, 0.5), s.vcycle(b, O.DAMPED, 0.5)
    assert abs(b @ Ma - a @ Mb) < 1e-10 * abs(a @ Mb)
    assert a @ Ma > 0 and b @ Mb > 0
    s.close()


def test_damped_skip_level_cycle_matches_brute_force(ke0):
    """The damped cycle over a x4 (skip) level 1 (FEM.cpp:266-272): the oracle's
    level operators, coarsest Galerkin matrix and V(1,1) cycle against the
    brute-force hierarchy with the same element rule."""
    vox, _ = problems.ellipsoid_union(18, 9, 2.0, 7.0, 41)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    s = O.OracleSystem(vox, ke0, fixed, 3, skip=1)
    bf = BF.Hierarchy(vox, ke0, fixed, 3, skip=1)
    assert s.num_levels == 3 and s.rmap(1).shape[1] == 125
    rng = np.random.default_rng(13)
    for l in range(1, 3):
        np.testing.assert_allclose(s.invdiag(l), bf.invdiag(l), rtol=1e-12)
        x = rng.standard_normal(3 * s.nodes(l))
        ref = bf.level_operator(l) @ x
        assert np.abs(s.level_matvec(l, x) - ref).max() <= 1e-12 * np.abs(ref).max(), l
    kc = bf.kc_galerkin()
    assert np.abs(s.kc(1) - kc).max() <= 1e-12 * np.abs(kc).max()
    np.testing.assert_allclose(s.kc(0), bf.kc(), rtol=1e-12, atol=1e-12 * np.abs(bf.kc()).max())
    r = rng.standard_normal(s.ndof)
    z, zb = s.vcycle(r, O.DAMPED, 0.5), bf.vcycle_damped(r, 0.5)
    assert np.abs(z - zb).max() <= 1e-12 * np.abs(zb).max()
    s.close()


def test_damped_vcycle_is_symmetric_positive(ke0, small_case):
    vox, fixed = small_case
    s = O.OracleSystem(vox, ke0, fixed, 3)
    rng = np.random.default_rng(3)
    mask = np.repeat(s.fixed(0), 3)
    a, b = rng.standard_normal(s.ndof), rng.standard_normal(s.ndof)
    a[mask] = 0
    b[mask] = 0
    Ma, Mb = s.vcycle(a, O.DAMPED, 0.6), s.vcycle(b, O.DAMPED, 0.6)
    assert abs(b @ Ma - a @ Mb) < 1e-10 * abs(a @ Mb)
    assert a @ Ma > 0 and b @ Mb > 0


@pytest.mark.parametrize("mode", [O.REF_ADDITIVE, O.REF_ADDITIVE_FP32, O.DAMPED, O.IDENTITY])
def test_pcg_box_matches_direct_solve(goldens, ke0, mode):
    s = O.OracleSystem(np.ones((16, 16, 32), np.int32), ke0, goldens["box_fixed"], 3)
    assert [(s.nodes(l), s.elems(l), s.nfixed(l)) for l in range(3)] == [(9537, 8192, 289), (1377, 1024, 81),
                                                                          (225, 128, 25)]
    assert s.free_dofs().size == 600
    x, it, rr = s.pcg(goldens["box_f"], tol=1e-12, precond=mode)
    assert rr < 1e-12
    err = np.linalg.norm(x - goldens["box_u"]) / np.linalg.norm(goldens["box_u"])
    assert err < 1e-9, err


def test_pcg_eigen_semantics(ke0, goldens):
    box = np.ones((4, 4, 8), np.int32)
    ids = problems.compact_ids(box)
    fixed = np.sort(ids[:, :, 0].reshape(-1))
    b = np.zeros(3 * int(ids.max() + 1))
    b[3 * ids[:, :, 8].reshape(-1) + 2] = -1.0
    # one level: the cycle is the exact inverse -> converged in the first step, 0 iterations counted
    s1 = O.OracleSystem(box, ke0, fixed, 1)
    x, it, rr = s1.pcg(b, tol=1e-10)
    assert it == 0 and rr < 1e-10
    # zero right-hand side: x = 0 and no iteration
    x, it, rr = s1.pcg(np.zeros_like(b))
    assert it == 0 and not x.any()
    # iteration cap
    s3 = O.OracleSystem(box, ke0, fixed, 2)
    x, it, rr = s3.pcg(b, tol=1e-14, max_iter=2, precond=O.IDENTITY)
    assert it == 2 and rr > 1e-14


def test_skip_level_hierarchy(ke0):
    vox, _ = problems.ellipsoid_union(16, 8, 2.0, 7.0, 41)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    s = O.OracleSystem(vox, ke0, fixed, 3, skip=1)
    # level 1 is a x4 coarsening with the 125-node lattice (FEM.cpp:266-272)
    assert s.rmap(1).shape[1] == 125
    assert s.extent(1) == (4, 4, 4)
    fine = BF.Hierarchy(vox, ke0, fixed, 1)
    c = BF.coarsen(vox, 4)
    P = BF.prolongation(vox, c, 4)
    np.testing.assert_allclose(s.invdiag(1), 1.0 / (P.T @ fine.A[0] @ P).diagonal(), rtol=1e-12)
    r = np.random.default_rng(5).standard_normal(s.ndof)
    np.testing.assert_allclose(s.restrict(0, r), P.T @ r, rtol=1e-12, atol=1e-12)


def test_flexible_cg_tolerates_inexact_preconditioner(goldens, ke0):
    """Why the reduced-precision solves use Polak-Ribiere beta: with FP32 storage
    of p, A p and z and a cycle inexact at the 1e-6 level (as an FP32 cycle is),
    Fletcher-Reeves PCG with FP64 residual replacement every 20 iterations
    stalls, while flexible PCG converges (emulated with the FP64 oracle)."""
    s = O.OracleSystem(np.ones((16, 16, 32), np.int32), ke0, goldens["box_fixed"], 3)
    f = goldens["box_f"]
    rhs = np.linalg.norm(f)

    def f32(v):
        return v.astype(np.float32).astype(np.float64)

    def run(flex, seed):
        rng = np.random.default_rng(seed)
        M = lambda r: f32(s.vcycle(r, O.DAMPED, 0.6) * (1 + 1e-6 * rng.standard_normal(f.size)))  # noqa: E731
        x, r = np.zeros_like(f), f.copy()
        z = M(r)
        p, zold, absnew = f32(z), z, r @ z
        for it in range(150):
            Ap = f32(s.matvec(f32(p)) * (1 + 1e-7 * rng.standard_normal(f.size)))
            a = absnew / (p @ Ap)
            x += a * p
            r -= a * Ap
            if it % 20 == 19:
                r = f - s.matvec(x)
            if np.linalg.norm(f - s.matvec(x)) < 1e-10 * rhs:
                break
            z = M(r)
            absold, absnew = absnew, r @ z
            beta = (absnew - (r @ zold if flex else 0.0)) / absold
            zold = z
            p = f32(z + beta * p)
        return it, np.linalg.norm(f - s.matvec(x)) / rhs

    it_flex, res_flex = run(True, 0)
    it_fr, res_fr = run(False, 0)
    assert res_flex < 1e-10 and it_flex < 40
    assert res_fr > 1e-9  # stalls


def test_default_omega_keeps_damped_jacobi_convergent(ke0):
    """The damped V(1,1) cycle is a positive definite preconditioner only while
    the smoother converges in the energy norm, omega * lambda_max(D^-1 A) < 2 on
    every level.  On an ellipsoid-union geometry (the generator of the 128^3 and
    256^3 workloads) lambda_max is ~3.7-3.8, set by surface nodes: the library's
    default omega = 0.5 stays inside the bound, the configurations' 0.6 does not
    (there the mixed and FP64 iteration counts become erratic)."""
    from scipy.sparse.linalg import LinearOperator, eigsh

    n = 32
    vox, _ = problems.ellipsoid_union(n, 60, 6 * n / 128, 30 * n / 128, 3)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    s = O.OracleSystem(vox, ke0, fixed, 3)
    lams = []
    for lvl in range(2):
        free = ~np.repeat(s.fixed(lvl), 3)
        h = np.sqrt(s.invdiag(lvl))

        def op(v, lvl=lvl, free=free, h=h):
            x = np.zeros(free.size)
            x[free] = h[free] * np.ravel(v)
            y = s.matvec(x) if lvl == 0 else s.level_matvec(lvl, x)
            return h[free] * y[free]

        # D^-1/2 A D^-1/2 on the free DOFs: symmetric, same spectrum as D^-1 A
        L = LinearOperator((int(free.sum()),) * 2, matvec=op, dtype=np.float64)
        lams.append(float(eigsh(L, k=1, which="LA", tol=1e-6, return_eigenvectors=False)[0]))
    s.close()
    lam = max(lams)
    assert 3.4 < lam < 4.0, lams
    assert 0.5 * lam < 2.0 < 0.6 * lam, lams


def test_pcg_warm_start(goldens, ke0):
    """Eigen's solveWithGuess: r = b - A x0.  Started at the direct solution it
    stops before any iteration; started from a perturbed one it reaches the
    same solution as the cold start, in fewer iterations."""
    s = O.OracleSystem(np.ones((16, 16, 32), np.int32), ke0, goldens["box_fixed"], 3)
    f, u = goldens["box_f"], goldens["box_u"]
    _, it, rr = s.pcg(f, tol=1e-8, precond=O.DAMPED, x0=u)
    assert it == 0
    x_cold, it_cold, _ = s.pcg(f, tol=1e-12, precond=O.DAMPED)
    x0 = u * (1 + 1e-4 * np.random.default_rng(0).standard_normal(u.size))
    x0[np.repeat(s.fixed(0), 3)] = 0
    x_warm, it_warm, rr = s.pcg(f, tol=1e-12, precond=O.DAMPED, x0=x0)
    assert rr < 1e-12 and it_warm < it_cold
    assert np.linalg.norm(x_warm - u) < 1e-9 * np.linalg.norm(u)
    s.close()
