"""GPU parity of the damped cycle with Chebyshev smoothing on its coarse levels:
the library's lambda_max(D^-1 A) bounds against a sparse eigensolve of the
oracle's level operators, one cycle against the oracle's twin given the same
bounds (FP64 to 1e-12, FP32 to 1e-4), and full solves against the oracle's PCG
(FP64 iteration counts equal, displacements to 1e-6)."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import vulkanfem_amd as V
from oracle import oracle as O
from vulkanfem_amd import problems

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def system(ke0):
    vox, _ = problems.ellipsoid_union(33, 12, 3.0, 11.0, 3)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    A = V.assemble_system_matrix(vox, ke0, fixed, 4)
    ora = O.OracleSystem(vox, ke0, fixed, 4)
    zs = np.nonzero((ids >= 0).any(axis=(1, 2)))[0]
    top = ids[zs.max()]
    f = np.zeros(A.n)
    f[3 * top[top >= 0] + 2] = -1.0 / np.count_nonzero(top >= 0)
    yield A, ora, f
    A.close()
    ora.close()


def level_lambda_max(ora, l):
    """lambda_max(D^-1 A_l) on the free DOFs: the largest eigenvalue of the
    symmetric D^-1/2 A_ff D^-1/2, by ARPACK on the oracle's level operator."""
    n = 3 * ora.nodes(l)
    free = ~np.repeat(ora.fixed(l), 3)
    sq = np.sqrt(ora.invdiag(l))

    def mv(v):
        x = np.zeros(n)
        x[free] = sq[free] * v
        return sq[free] * ora.level_matvec(l, x)[free]

    op = sla.LinearOperator((int(free.sum()),) * 2, matvec=mv, dtype=np.float64)
    return float(sla.eigsh(op, k=1, which="LA", tol=1e-10, return_eigenvectors=False)[0])


def test_lmax_bounds(system):
    """The interval end the library formed at assembly bounds lambda_max from
    above, by at most 25%; levels 0 and the coarsest have none."""
    A, ora, _ = system
    nl = A.num_levels
    assert nl == 4
    for l in range(1, nl - 1):
        lam, lm = level_lambda_max(ora, l), A.level_lmax(l)
        say("level %d: lambda_max %.4f, bound %.4f" % (l, lam, lm))
        assert lam <= lm <= 1.25 * lam
    assert A.level_lmax(0) == 0 and A.level_lmax(nl - 1) == 0


@pytest.mark.parametrize("level,degree,ratio", [(1, 2, 30.0), (2, 3, 8.0)])
def test_chebyshev_cycle_matches_oracle(system, level, degree, ratio):
    A, ora, _ = system
    lmax = [A.level_lmax(l) for l in range(A.num_levels)]
    r = np.random.default_rng(21).standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    z_jac = ora.vcycle(r, O.DAMPED, 0.6)
    ora.set_chebyshev(level, degree, ratio, lmax)
    try:
        z_ref = ora.vcycle(r, O.DAMPED, 0.6)
    finally:
        ora.set_chebyshev(0)
    kw = dict(cycle=V.CYCLE_DAMPED, omega=0.6, cheb_level=level, cheb_degree=degree, cheb_ratio=ratio)
    z64 = V.MultigridPreconditioner(A, **kw).solve(r)
    z32 = V.MultigridPreconditioner(A, precision=V.FP32, **kw).solve(r)
    say("chebyshev cycle from level %d, degree %d, ratio %g: fp64 %.2e fp32 %.2e (jacobi differs by %.2e)"
        % (level, degree, ratio, rel(z64, z_ref), rel(z32, z_ref), rel(z_jac, z_ref)))
    assert rel(z64, z_ref) < 1e-12
    assert rel(z32, z_ref) < 1e-4
    assert rel(z_jac, z_ref) > 1e-3


def test_chebyshev_solve_matches_oracle(system):
    """FP64 PCG with the Chebyshev-smoothed cycle takes the oracle's iteration
    count; the FP32 cycle under the FP64 outer CG reaches the same displacement,
    and the columns of a two-column block equal their single solves."""
    A, ora, f = system
    lmax = [A.level_lmax(l) for l in range(A.num_levels)]
    for level in (1, 2):
        ora.set_chebyshev(level, 2, 60.0, lmax)
        try:
            u_ref, it_ref, _ = ora.pcg(f, tol=1e-8, precond=O.DAMPED, omega=0.5)
            u_tight, _, _ = ora.pcg(f, tol=1e-13, precond=O.DAMPED, omega=0.5)
        finally:
            ora.set_chebyshev(0)
        kw = dict(cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=level)
        u, st = A.solve(f, tol=1e-8, **kw)
        say("chebyshev from level %d: %d iterations (oracle %d)" % (level, st["iterations"], it_ref))
        assert st["iterations"] == it_ref
        assert rel(u, u_ref) < 1e-8
        u32, st32 = A.solve(f, tol=1e-10, precision=V.FP32, **kw)
        assert st32["converged"] == 1 and st32["rel_residual"] < 1e-10, st32
        assert rel(u32, u_tight) < 1e-6
    # a block of columns: FP64 columns equal their single solves bit for bit
    rng = np.random.default_rng(22)
    F = np.stack([f, rng.standard_normal(A.n)])
    F[1][np.repeat(ora.fixed(0), 3)] = 0
    U, stb = A.solve_batched(F, tol=1e-10, cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=2)
    for k in range(2):
        u1, _ = A.solve(F[k], tol=1e-10, cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=2)
        assert np.array_equal(U[k], u1)


def test_chebyshev_options_checked(system):
    A, _, f = system
    with pytest.raises(V.VfemError):
        A.solve(f, cheb_level=2)  # the reference additive cycle has no Chebyshev smoothing
    for bad in (dict(cheb_degree=9), dict(cheb_ratio=0.5), dict(cheb_level=-1)):
        with pytest.raises(V.VfemError):
            A.solve(f, **{"cycle": V.CYCLE_DAMPED, "cheb_level": 2, **bad})


def test_chebyshev_fused_and_separate_steps_agree(system):
    """The Chebyshev steps run fused into the element-form product (L0_AUTO on
    these small coarse levels) or into the stencil form's surface-pass launch
    (L0_STENCIL: the surface nodes in its epilogue, the other DOFs in extra
    blocks): both against the oracle's twin, for an even and an odd number of
    steps (an odd count ends in the level's t buffer)."""
    A, ora, _ = system
    lmax = [A.level_lmax(l) for l in range(A.num_levels)]
    r = np.random.default_rng(23).standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    for degree in (2, 3):
        ora.set_chebyshev(1, degree, 60.0, lmax)
        try:
            z_ref = ora.vcycle(r, O.DAMPED, 0.5)
        finally:
            ora.set_chebyshev(0)
        out = {}
        try:
            for kind in (V.L0_AUTO, V.L0_STENCIL):
                A.set_level0_kernel(kind)
                forms = [A.level_operator_form(l) for l in range(1, A.num_levels - 1)]
                out[kind] = V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=1,
                                                      cheb_degree=degree).solve(r)
                say("degree %d, form %d (coarse forms %s): %.2e" % (degree, kind, forms, rel(out[kind], z_ref)))
                assert rel(out[kind], z_ref) < 1e-12
        finally:
            A.set_level0_kernel(V.L0_AUTO)
        assert rel(out[V.L0_AUTO], out[V.L0_STENCIL]) < 1e-12


def test_chebyshev_with_skip_level_and_fp64_parts(ke0):
    """Chebyshev smoothing on the x4 skip level (element-table operator) and the
    levels below it matches the oracle's cycle; the FP32 cycle with FP64 coarse
    levels or an FP64 CG operator reaches the oracle's displacement."""
    vox, _ = problems.ellipsoid_union(32, 10, 3.0, 10.0, 5)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    A = V.assemble_system_matrix(vox, ke0, fixed, 4, skip_levels=1)
    ora = O.OracleSystem(vox, ke0, fixed, 4, skip=1)
    lmax = [A.level_lmax(l) for l in range(A.num_levels)]
    assert A.num_levels == 4 and all(lm > 0 for lm in lmax[1:3])
    ora.set_chebyshev(1, 2, 60.0, lmax)
    r = np.random.default_rng(24).standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    z_ref = ora.vcycle(r, O.DAMPED, 0.5)
    f = np.zeros(A.n)
    f[2::3] = -1.0 / (A.n // 3)
    f[np.repeat(ora.fixed(0), 3)] = 0
    u_ref, it_ref, _ = ora.pcg(f, tol=1e-12, precond=O.DAMPED, omega=0.5)
    ora.close()
    kw = dict(cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=1)
    z = V.MultigridPreconditioner(A, **kw).solve(r)
    say("skip level with Chebyshev from level 1: cycle %.2e" % rel(z, z_ref))
    assert rel(z, z_ref) < 1e-12
    for extra in (dict(coarse_fp64=1), dict(op_fp64=1)):
        u, st = A.solve(f, tol=1e-10, precision=V.FP32, **kw, **extra)
        say("skip level, Chebyshev, FP32 cycle", extra, st["iterations"], "iterations, err %.2e" % rel(u, u_ref))
        assert st["converged"] == 1 and st["rel_residual"] < 1e-10
        assert rel(u, u_ref) < 1e-6
    A.close()
