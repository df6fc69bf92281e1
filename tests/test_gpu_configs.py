"""GPU parity at the measured configurations' full sizes, through the C-ABI,
against the CPU oracle on the same seeded inputs:

    ellipsoids128   128^3, 60 ellipsoids, 5 levels, plain FP64 (~2.98M DOF)
    batched128x16   128^3 geometry (seed 3), 16 patch load cases as one block
    ellipsoids256   256^3, seed 2, 120 ellipsoids, 6 levels (~32.6M DOF): the
                    level-0 operator and one damped cycle (damped Jacobi, and
                    Chebyshev smoothing from level 2) against the oracle, mixed
                    precision against FP64
    learned         the same 256^3 solve with the trained U-Net at level 3

Tolerances: FP64 operators to 1e-14 and cycles to 1e-12 relative; every solve
reaches the FP64 true relative residual 1e-8; displacements agree with the FP64
oracle (solved to 1e-10) to 1e-6 relative L2; FP64 iteration counts agree with
the oracle's to 2 % (same algorithm, rounding drift only)."""
import os

import numpy as np
import pytest

import vulkanfem_amd as V
from oracle import oracle as O
from vulkanfem_amd import problems

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*a):
    print("[gpu-test]", *a, flush=True)


def within(it, it_ref, frac=0.02):
    return abs(it - it_ref) <= max(1, int(round(frac * it_ref)))


@pytest.fixture(scope="module")
def ell128(ke0):
    p = problems.workload("ellipsoids128")
    nl = problems.WORKLOAD_LEVELS["ellipsoids128"]
    say("assemble ellipsoids128", p.ndof, "DOF")
    A = V.assemble_system_matrix(p.voxels, ke0, p.fixed, nl)
    ora = O.OracleSystem(p.voxels, ke0, p.fixed, nl)
    u_ref, it_ref, rr = ora.pcg(p.loads[0], tol=1e-10, precond=O.DAMPED, omega=0.5)
    say("oracle FP64 damped solve to 1e-10: %d iterations, relres %.2e" % (it_ref, rr))
    yield p, A, ora, u_ref
    ora.close()
    A.close()


@pytest.mark.parametrize("omega", [0.6, 0.5])
def test_config2_damped_fp64_matches_oracle(ell128, omega):
    """Config 2 as stated (plain FP64 MGPCG, damped V(1,1), omega 0.6) and at the
    library's default omega 0.5: the GPU iteration count equals the oracle's
    damped twin's to 2 %, the displacement the oracle's solution to 1e-6."""
    p, A, ora, u_ref = ell128
    f = p.loads[0]
    _, it_ora, _ = ora.pcg(f, tol=1e-8, precond=O.DAMPED, omega=omega)
    u, st = A.solve(f, tol=1e-8, precision=V.FP64, cycle=V.CYCLE_DAMPED, omega=omega)
    err = rel(u, u_ref)
    say("config 2 omega %.1f: GPU %d iterations, oracle %d, true relres %.2e, displacement err %.2e, %.1f ms"
        % (omega, st["iterations"], it_ora, st["rel_residual"], err, st["solve_ms"]))
    assert st["converged"] == 1 and st["rel_residual"] < 1e-8
    assert within(st["iterations"], it_ora), (st["iterations"], it_ora)
    assert err < 1e-6


def test_config2_chebyshev_matches_oracle(ell128):
    """Config 2 with Chebyshev smoothing (degree 2, ratio 60) on levels >= 2 of the
    damped cycle, FP64: the oracle's twin, given the library's lambda_max bounds,
    takes the same iteration count to 2 %, fewer than damped Jacobi everywhere,
    and both reach the oracle's displacement to 1e-6."""
    p, A, ora, u_ref = ell128
    f = p.loads[0]
    lmax = [A.level_lmax(l) for l in range(A.num_levels)]
    ora.set_chebyshev(2, 2, 60.0, lmax)
    try:
        _, it_ora, _ = ora.pcg(f, tol=1e-8, precond=O.DAMPED, omega=0.5)
    finally:
        ora.set_chebyshev(0)
    u, st = A.solve(f, tol=1e-8, precision=V.FP64, cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=2)
    _, st_j = A.solve(f, tol=1e-8, precision=V.FP64, cycle=V.CYCLE_DAMPED, omega=0.5)
    err = rel(u, u_ref)
    say("config 2 chebyshev from level 2: GPU %d iterations (damped Jacobi %d), oracle %d, displacement err %.2e, "
        "%.1f ms" % (st["iterations"], st_j["iterations"], it_ora, err, st["solve_ms"]))
    assert st["converged"] == 1 and st["rel_residual"] < 1e-8
    assert within(st["iterations"], it_ora), (st["iterations"], it_ora)
    assert st["iterations"] < st_j["iterations"]
    assert err < 1e-6


def test_config2_reference_cycle_and_mixed(ell128):
    """The reference's additive cycle (FP64) and the mixed plan (FP32 damped cycle,
    FP64 outer CG) on config 2 reach the oracle's displacement."""
    p, A, ora, u_ref = ell128
    f = p.loads[0]
    for kw in (dict(precision=V.FP64, cycle=V.CYCLE_REFERENCE),
               dict(precision=V.FP32, cycle=V.CYCLE_DAMPED, omega=0.5)):
        u, st = A.solve(f, tol=1e-8, **kw)
        err = rel(u, u_ref)
        say("config 2", kw, "iterations", st["iterations"], "relres %.2e err %.2e" % (st["rel_residual"], err))
        assert st["converged"] == 1 and st["rel_residual"] < 1e-8
        assert err < 1e-6


def test_config4_batched_block(ke0):
    """Config 4: 16 patch load cases on the seed-3 128^3 geometry as one block.
    Every column reaches relative residual 1e-8; columns 0, 7 and 15 equal their
    single solves (FP64 bit for bit, the FP32 cycle to 1e-6); column 0 matches
    the oracle's FP64 solution to 1e-6."""
    p = problems.workload("batched128x16")
    nl = problems.WORKLOAD_LEVELS["batched128x16"]
    A = V.assemble_system_matrix(p.voxels, ke0, p.fixed, nl)
    F = p.loads
    assert F.shape[0] == 16
    ora = O.OracleSystem(p.voxels, ke0, p.fixed, nl)
    u_ref, it_ref, _ = ora.pcg(F[0], tol=1e-10, precond=O.DAMPED, omega=0.5)
    ora.close()
    for prec in (V.FP64, V.FP32):
        U, stats = A.solve_batched(F, tol=1e-8, precision=prec, cycle=V.CYCLE_DAMPED, omega=0.5)
        its = [s["iterations"] for s in stats]
        say("config 4 precision", prec, "iterations", its, "solve %.1f ms" % stats[0]["solve_ms"])
        assert all(s["converged"] == 1 and s["rel_residual"] < 1e-8 for s in stats)
        for k in (0, 7, 15):
            u, st = A.solve(F[k], tol=1e-8, precision=prec, cycle=V.CYCLE_DAMPED, omega=0.5)
            if prec == V.FP64:
                assert st["iterations"] == its[k] and np.array_equal(U[k], u), k
            else:
                assert rel(U[k], u) < 1e-6, (k, rel(U[k], u))
        err = rel(U[0], u_ref)
        say("config 4 column 0 displacement err vs oracle %.2e (oracle %d iterations to 1e-10)" % (err, it_ref))
        assert err < 1e-6
    A.close()


@pytest.fixture(scope="module")
def ell256(ke0):
    p = problems.workload("ellipsoids256")
    nl = problems.WORKLOAD_LEVELS["ellipsoids256"]
    say("assemble ellipsoids256", p.ndof, "DOF")
    A = V.assemble_system_matrix(p.voxels, ke0, p.fixed, nl)
    yield p, A, nl
    A.close()


def test_config3_operator_and_cycle_match_oracle(ell256, ke0):
    """The 256^3 rough geometry (10.9M nodes: the 12-plane stencil chunks) against
    the oracle: the level-0 operator in stencil and
    element form (FP64 to 1e-14, FP32 to 2e-6) and one damped V(1,1) cycle in
    both forms (FP64 to 1e-12)."""
    p, A, nl = ell256
    assert A.n // 3 >= 4_000_000
    ora = O.OracleSystem(p.voxels, ke0, p.fixed, nl)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(A.n)
    r = rng.standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    y_ref = ora.matvec(x)
    z_ref = ora.vcycle(r, O.DAMPED, 0.5)
    # the cycle with Chebyshev smoothing on levels >= 2, given the library's bounds
    ora.set_chebyshev(2, 2, 60.0, [A.level_lmax(l) for l in range(A.num_levels)])
    zc_ref = ora.vcycle(r, O.DAMPED, 0.5)
    ora.close()
    ec = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=2).solve(r), zc_ref)
    say("256^3 damped cycle with Chebyshev smoothing from level 2: %.2e" % ec)
    assert ec < 1e-12
    try:
        for kind in (V.L0_STENCIL, V.L0_ELEMENT):
            A.set_level0_kernel(kind)
            e64, e32 = rel(A.matvec(x, V.FP64), y_ref), rel(A.matvec(x, V.FP32), y_ref)
            ez = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.5).solve(r), z_ref)
            say("256^3 form %d: operator fp64 %.2e fp32 %.2e, damped cycle %.2e" % (kind, e64, e32, ez))
            assert e64 < 1e-14 and e32 < 2e-6 and ez < 1e-12
    finally:
        A.set_level0_kernel(V.L0_AUTO)


@pytest.mark.parametrize("cheb_level", [0, 2])
def test_config3_mixed_matches_fp64(ell256, cheb_level):
    """Config 3: the mixed plan (FP32 damped cycle, FP64 outer CG, true residual
    every 20 iterations) and plain FP64 both reach true relative residual 1e-8 and
    agree to 1e-6 (the bench reports both against the CPU oracle's solution), with
    damped Jacobi on every level and with Chebyshev smoothing from level 2 (the
    measured plan)."""
    p, A, _ = ell256
    f = p.loads[0]
    kw = dict(cycle=V.CYCLE_DAMPED, omega=0.5, cheb_level=cheb_level)
    u64, s64 = A.solve(f, tol=1e-8, precision=V.FP64, **kw)
    u32, s32 = A.solve(f, tol=1e-8, precision=V.FP32, true_res_every=20, **kw)
    say("config 3 (cheb_level %d): fp64 %d iterations %.1f ms, mixed %d iterations %.1f ms, relres %.2e / %.2e, "
        "diff %.2e" % (cheb_level, s64["iterations"], s64["solve_ms"], s32["iterations"], s32["solve_ms"],
                       s64["rel_residual"], s32["rel_residual"], rel(u32, u64)))
    assert s64["converged"] == 1 and s64["rel_residual"] < 1e-8
    assert s32["converged"] == 1 and s32["rel_residual"] < 1e-8 and s32["true_res_count"] >= 1
    assert rel(u32, u64) < 1e-6


def test_config5_learned_correction(ell256):
    """Config 5: the trained U-Net as the level-3 coarse correction in flexible
    PCG on the 256^3 solve, under the FP64 cycle and under the mixed plan; it
    converges with the network in use and reaches the plain solve's displacement."""
    p, A, nl = ell256
    path = os.path.join(ROOT, "vulkanfem_amd", "data", "unet_level3.npy")
    f = p.loads[0]
    net = V.UNet(0)
    net.set_weights(np.load(path))
    u0, s0 = A.solve(f, tol=1e-8, precision=V.FP64, cycle=V.CYCLE_DAMPED, omega=0.5)
    A.attach_unet(net, 3)
    try:
        for prec in (V.FP64, V.FP32):
            u, st = A.solve(f, tol=1e-8, precision=prec, cycle=V.CYCLE_DAMPED, omega=0.5, learned=1)
            say("config 5 precision %d: %d iterations (%d with the network, fell back %d), plain FP64 %d; err %.2e"
                % (prec, st["iterations"], st["learned_iters"], st["fell_back"], s0["iterations"], rel(u, u0)))
            assert st["converged"] == 1 and st["rel_residual"] < 1e-8 and st["learned_iters"] > 0
            assert rel(u, u0) < 1e-6
    finally:
        A.attach_unet(None, 0)


def test_skip_level_damped_cycle_at_256(ke0):
    """The 256^3 workload with the reference's skip level (x4 first coarsening,
    src/FEM.cpp:266-272): the damped cycle with the plane-tiled x4 transfers
    against the oracle's twin (FP64 to 1e-12), and an FP32 solve to relative
    residual 1e-8 that matches the plain FP64 solve."""
    p = problems.workload("ellipsoids256")
    A = V.assemble_system_matrix(p.voxels, ke0, p.fixed, 6, skip_levels=1)
    ora = O.OracleSystem(p.voxels, ke0, p.fixed, 6, skip=1)
    r = np.random.default_rng(21).standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    z_ref = ora.vcycle(r, O.DAMPED, 0.5)
    ora.close()
    ez = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.5).solve(r), z_ref)
    u32, s32 = A.solve(p.loads[0], tol=1e-8, precision=V.FP32, cycle=V.CYCLE_DAMPED, omega=0.5)
    u64, s64 = A.solve(p.loads[0], tol=1e-8, precision=V.FP64, cycle=V.CYCLE_DAMPED, omega=0.5)
    say("256^3 skip: cycle err %.2e; fp32 %d iterations %.1f ms, fp64 %d iterations %.1f ms, diff %.2e"
        % (ez, s32["iterations"], s32["solve_ms"], s64["iterations"], s64["solve_ms"], rel(u32, u64)))
    assert ez < 1e-12
    assert s32["converged"] == 1 and s64["converged"] == 1 and s32["rel_residual"] < 1e-8
    assert rel(u32, u64) < 1e-6
    A.close()
