"""GPU parity: the HIP library, called through the C-ABI, against the CPU oracle
on the same seeded inputs.

Tolerances: FP64 kernels agree with the oracle to rounding (1e-12 relative);
FP32 kernels to 1e-5 relative; full solves must reach the requested true
relative residual and agree with the FP64 oracle/direct solution to 1e-6
relative L2 displacement error."""
import numpy as np
import pytest

import vulkanfem_amd as V
from oracle import oracle as O
from vulkanfem_amd import problems

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*a):
    print("[gpu-test]", *a, flush=True)


def geometry(n, cnt, a0, a1, seed):
    vox, _ = problems.ellipsoid_union(n, cnt, a0, a1, seed)
    ids = problems.compact_ids(vox)
    return vox, np.sort(ids[0][ids[0] >= 0])


CASES = {
    "box": (np.ones((16, 16, 32), np.int32), None, 3),
    "ell14": geometry(14, 8, 2.0, 6.0, 32) + (3,),
    "ell24": geometry(24, 10, 3.0, 9.0, 7) + (3,),
    "ell33": geometry(33, 12, 3.0, 11.0, 3) + (4,),
}


@pytest.fixture(scope="module")
def systems(ke0, goldens):
    out = {}
    for name, (vox, fixed, nl) in CASES.items():
        if fixed is None:
            fixed = goldens["box_fixed"]
        say("assemble", name, vox.shape)
        out[name] = (V.assemble_system_matrix(vox, ke0, fixed, nl), O.OracleSystem(vox, ke0, fixed, nl), vox, fixed)
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_matvec_parity(systems, name):
    A, ora, _, _ = systems[name]
    x = np.random.default_rng(1).standard_normal(A.n)
    say("matvec", name)
    ref = ora.matvec(x)
    e64, e32, e16 = rel(A.matvec(x, V.FP64), ref), rel(A.matvec(x, V.FP32), ref), rel(A.matvec(x, V.BF16), ref)
    say("matvec rel err fp64 %.2e fp32(3-pass bf16 MFMA) %.2e bf16(1 pass) %.2e" % (e64, e32, e16))
    assert e64 < 1e-13
    assert e32 < 1e-4
    assert e16 < 1e-2  # one BF16 product of element differences (measured 1.6e-3)
    emx = rel(A.matvec(x, V.MIXED), ref)
    say("matvec rel err mixed (6-pass MFMA on FP64 differences) %.2e" % emx)
    assert emx < 1e-6


@pytest.mark.parametrize("name", list(CASES))
def test_stencil_form_matches_element_form(systems, name):
    """The assembled-stencil level-0 operator plus its surface correction against
    the element form and the oracle (FP64 to rounding, FP32 to single precision).
    The row-wave and the plane-staged sum-factorised kernels do the same
    arithmetic at every node the surface pass keeps: their products are equal bit
    for bit."""
    A, ora, _, _ = systems[name]
    assert A.level0_kernel == V.L0_STENCIL  # Ke0's stencil has the tensor form: sum-factorised kernel
    x = np.random.default_rng(5).standard_normal(A.n)
    ref = ora.matvec(x)
    out = {}
    try:
        for kind in (V.L0_ELEMENT, V.L0_STENCIL, V.L0_STENCIL_DIRECT, V.L0_STENCIL_PLANES):
            A.set_level0_kernel(kind)
            assert A.level0_kernel == kind
            out[kind] = (A.matvec(x, V.FP64), A.matvec(x, V.FP32))
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    e64, e32 = out[V.L0_ELEMENT]
    assert np.array_equal(out[V.L0_STENCIL][0], out[V.L0_STENCIL_PLANES][0])
    assert np.array_equal(out[V.L0_STENCIL][1], out[V.L0_STENCIL_PLANES][1])
    for kind in (V.L0_STENCIL, V.L0_STENCIL_DIRECT, V.L0_STENCIL_PLANES):
        s64, s32 = out[kind]
        say("stencil kind %d %s: fp64 %.2e (element %.2e), fp32 %.2e (element %.2e)"
            % (kind, name, rel(s64, ref), rel(e64, ref), rel(s32, ref), rel(e32, ref)))
        assert rel(s64, ref) < 1e-14
        assert rel(s64, e64) < 1e-14
        assert rel(s32, ref) < 2e-6


def test_large_level_kernels_match_element_form(ke0):
    """A level 0 of at least 4M nodes, as at 256^3, runs the 12-plane stencil
    kernels (smaller levels run 4 planes): its operator and damped cycle against
    the element form on the same system, which the smaller cases pin to the
    oracle."""
    p = problems.cantilever_box(160, 160, 160)
    A = V.assemble_system_matrix(p.voxels, ke0, p.fixed, 6)
    assert A.n // 3 >= 4_000_000
    rng = np.random.default_rng(7)
    x = rng.standard_normal(A.n)
    r = rng.standard_normal(A.n)
    r[3 * np.repeat(np.asarray(p.fixed), 3) + np.tile([0, 1, 2], len(p.fixed))] = 0
    out = {}
    try:
        for kind in (V.L0_ELEMENT, V.L0_STENCIL, V.L0_STENCIL_PLANES):
            A.set_level0_kernel(kind)
            say("large system, operator form", kind)
            out[kind] = (A.matvec(x, V.FP64), A.matvec(x, V.FP32),
                         V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED).solve(r))
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    e64, _, ze = out[V.L0_ELEMENT]
    s64, s32, zs = out[V.L0_STENCIL]
    say("large system: fp64 %.2e, fp32 %.2e, cycle %.2e" % (rel(s64, e64), rel(s32, s64), rel(zs, ze)))
    assert rel(s64, e64) < 1e-14
    assert rel(s32, s64) < 2e-6
    assert rel(zs, ze) < 1e-12
    assert np.array_equal(s64, out[V.L0_STENCIL_PLANES][0]) and np.array_equal(s32, out[V.L0_STENCIL_PLANES][1])


@pytest.mark.parametrize("name", ["ell24", "ell33"])
def test_stencil_cycle_matches_element_cycle(systems, name):
    """The damped cycle with the stencil form on every level (the coarse levels'
    full Galerkin elements included) against the element-form cycle and the
    oracle's twin."""
    A, ora, _, _ = systems[name]
    forms = [A.level_operator_form(l) for l in range(A.num_levels)]
    say("operator forms (auto)", name, forms)
    assert forms[0] == V.L0_STENCIL  # coarse levels this small keep the element form under L0_AUTO
    r = np.random.default_rng(4).standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    z_d = ora.vcycle(r, O.DAMPED, 0.6)
    out = {}
    try:
        for kind in (V.L0_ELEMENT, V.L0_STENCIL, V.L0_STENCIL_DIRECT, V.L0_STENCIL_PLANES):
            A.set_level0_kernel(kind)
            if kind != V.L0_ELEMENT:  # every div-2 level with full elements
                assert all(A.level_operator_form(l) == kind for l in range(A.num_levels - 1))
            out[kind] = (V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.6).solve(r),
                         V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.6, precision=V.FP32).solve(r))
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    say("cycle %s: fp64 %.2e fp32 %.2e (element fp32 %.2e)" % (name, rel(out[2][0], z_d), rel(out[2][1], z_d),
                                                               rel(out[1][1], z_d)))
    for kind in (V.L0_STENCIL, V.L0_STENCIL_DIRECT, V.L0_STENCIL_PLANES):
        assert rel(out[kind][0], z_d) < 1e-12
        assert rel(out[kind][0], out[V.L0_ELEMENT][0]) < 1e-12
        assert rel(out[kind][1], z_d) < 1e-5


def test_unstructured_ke_uses_element_form(ke0):
    """A Ke without the stencil structure keeps the element-form kernels, and
    asking for the stencil form is refused."""
    rng = np.random.default_rng(3)
    B = rng.standard_normal((24, 24)) * 1e-2
    ke = ke0 + B @ B.T
    vox = np.ones((6, 7, 8), np.int32)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    A = V.assemble_system_matrix(vox, ke, fixed, 2)
    assert A.level0_kernel == V.L0_ELEMENT
    with pytest.raises(V.VfemError):
        A.set_level0_kernel(V.L0_STENCIL)
    x = rng.standard_normal(A.n)
    ref = O.OracleSystem(vox, ke, fixed, 2).matvec(x)
    assert rel(A.matvec(x, V.FP64), ref) < 1e-13


def test_matvec_fp32_on_smooth_field(systems, goldens):
    # K u = f on the direct solution: extreme cancellation |K||u| >> |K u|
    A, ora, _, _ = systems["box"]
    u = goldens["box_u"]
    say("matvec fp32 smooth")
    ref = ora.matvec(u)
    assert rel(A.matvec(u, V.FP64), ref) < 1e-11
    # FP32 storage of u alone costs ~1e-4 here; the kernel's arithmetic must add
    # little on top: compare with the exact product of the FP32-rounded input
    ref32 = ora.matvec(u.astype(np.float32).astype(np.float64))
    e32 = rel(A.matvec(u, V.FP32), ref32)
    say("smooth field fp32 matvec rel err vs exact product of the fp32 input %.2e" % e32)
    assert e32 < 1e-3
    # FP64 input differences: the error scales with the local differences of u, so
    # the product stays accurate where FP32 storage of u alone costs ~1e-4
    emx = rel(A.matvec(u, V.MIXED), ref)
    say("smooth field mixed matvec rel err %.2e" % emx)
    assert emx < 1e-5


def test_matvec_golden(goldens, ke0):
    for k in range(3):
        say("matvec golden", k)
        A = V.assemble_system_matrix(goldens[f"mv{k}_voxels"], ke0, goldens[f"mv{k}_fixed"], 2)
        assert rel(A.matvec(goldens[f"mv{k}_x"]), goldens[f"mv{k}_y"]) < 1e-13


@pytest.mark.parametrize("name", list(CASES))
def test_precondition_parity(systems, name):
    A, ora, _, _ = systems[name]
    r = np.random.default_rng(2).standard_normal(A.n)
    say("precondition ref", name)
    z_ref = ora.vcycle(r, O.REF_ADDITIVE)
    assert rel(V.MultigridPreconditioner(A).solve(r), z_ref) < 1e-12
    assert rel(V.MultigridPreconditioner(A, precision=V.FP32).solve(r), z_ref) < 1e-5
    say("precondition damped", name)
    mask = np.repeat(ora.fixed(0), 3)
    r[mask] = 0
    z_d = ora.vcycle(r, O.DAMPED, 0.6)
    assert rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.6).solve(r), z_d) < 1e-12
    e32 = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.6, precision=V.FP32).solve(r), z_d)
    e16 = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=0.6, precision=V.BF16).solve(r), z_d)
    say("damped cycle rel err fp32 %.2e bf16 %.2e" % (e32, e16))
    assert e32 < 1e-4 and e16 < 2e-2


def test_dot(systems):
    A, _, _, _ = systems["ell24"]
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
    assert abs(A.dot(a, b) - a @ b) < 1e-12 * np.abs(a * b).sum()


def test_box_solve_matches_direct_solution(systems, goldens):
    A, ora, _, _ = systems["box"]
    f, u_direct = goldens["box_f"], goldens["box_u"]
    for cycle in (V.CYCLE_REFERENCE, V.CYCLE_DAMPED):
        for prec in (V.FP64, V.FP32, V.MIXED):
            say("box solve", cycle, prec)
            u, st = A.solve(f, tol=1e-10, precision=prec, cycle=cycle)
            assert st["converged"] == 1 and st["rel_residual"] < 1e-10, st
            assert rel(u, u_direct) < 1e-6, (cycle, prec, rel(u, u_direct))
    # FP64 reference cycle: same iteration count as the oracle (Eigen semantics)
    _, it_ora, _ = ora.pcg(f, tol=1e-8, precond=O.REF_ADDITIVE)
    _, st = A.solve(f, tol=1e-8)
    assert st["iterations"] == it_ora
    # the damped cycle at the configuration's omega = 0.6 and at the default 0.5
    for om in (0.6, 0.5):
        _, it_ora, _ = ora.pcg(f, tol=1e-8, precond=O.DAMPED, omega=om)
        _, st = A.solve(f, tol=1e-8, cycle=V.CYCLE_DAMPED, omega=om)
        assert st["iterations"] == it_ora, (om, st["iterations"], it_ora)


@pytest.mark.parametrize("name", ["ell24", "ell33"])
def test_solve_parity(systems, name):
    A, ora, vox, fixed = systems[name]
    ids = problems.compact_ids(vox)
    zs = np.nonzero((ids >= 0).any(axis=(1, 2)))[0]
    top = ids[zs.max()]
    f = np.zeros(A.n)
    f[3 * top[top >= 0] + 2] = -1.0 / np.count_nonzero(top >= 0)
    u_ref, _, _ = ora.pcg(f, tol=1e-13, precond=O.DAMPED)
    for cycle in (V.CYCLE_REFERENCE, V.CYCLE_DAMPED):
        for prec in (V.FP64, V.FP32, V.MIXED):
            say("solve", name, cycle, prec)
            u, st = A.solve(f, tol=1e-10, precision=prec, cycle=cycle)
            assert st["converged"] == 1 and st["rel_residual"] < 1e-10, st
            assert rel(u, u_ref) < 1e-6
    # FP32 cycle with the CG operator and direction in FP64
    u, st = A.solve(f, tol=1e-10, precision=V.FP32, cycle=V.CYCLE_DAMPED, op_fp64=1)
    assert st["converged"] == 1 and st["rel_residual"] < 1e-10, st
    assert rel(u, u_ref) < 1e-6


def test_edge_cases(ke0):
    box = np.ones((4, 4, 8), np.int32)
    ids = problems.compact_ids(box)
    fixed = np.sort(ids[:, :, 0].reshape(-1))
    f = np.zeros(3 * int(ids.max() + 1))
    f[3 * ids[:, :, 8].reshape(-1) + 2] = -1.0
    say("edge: single level")
    A1 = V.assemble_system_matrix(box, ke0, fixed, 1)
    assert A1.num_levels == 1
    u, st = A1.solve(f, tol=1e-10)
    assert st["iterations"] == 0 and st["rel_residual"] < 1e-10
    say("edge: zero rhs")
    A2 = V.assemble_system_matrix(box, ke0, fixed, 2)
    u, st = A2.solve(np.zeros_like(f))
    assert st["iterations"] == 0 and not u.any() and st["converged"] == 1
    say("edge: iteration cap")
    u, st = A2.solve(f, tol=1e-15, max_iter=3)
    assert st["iterations"] == 3 and st["converged"] == 0
    assert st["status"] == V.api.ERR_NOT_CONVERGED
    say("edge: warm start from the solution")
    u, st = A2.solve(f, tol=1e-10)
    u2, st2 = A2.solve(f, u=u.copy(), tol=1e-8, warm_start=1)
    assert st2["iterations"] == 0 and rel(u2, u) < 1e-12
    say("edge: early level stop")
    A3 = V.assemble_system_matrix(box, ke0, fixed, 6)
    assert A3.num_levels == 2  # (8,4,4) -> (4,2,2): a dim <= 2 stops the hierarchy (FEM.cpp:465-470)
    say("edge: bad options")
    with pytest.raises(V.VfemError):
        A2.solve(f, precision=7)


def test_skip_level_solve(ke0):
    vox, fixed = geometry(24, 10, 3.0, 9.0, 7)
    say("skip level")
    A = V.assemble_system_matrix(vox, ke0, fixed, 3, skip_levels=1)
    ora = O.OracleSystem(vox, ke0, fixed, 3, skip=1)
    r = np.random.default_rng(4).standard_normal(A.n)
    assert rel(V.MultigridPreconditioner(A).solve(r), ora.vcycle(r, O.REF_ADDITIVE)) < 1e-12
    f = np.zeros(A.n)
    f[2::3] = -1.0 / (A.n // 3)
    f[np.repeat(ora.fixed(0), 3)] = 0  # a loaded fixed DOF can never converge (nor in the reference)
    u, st = A.solve(f, tol=1e-10)
    u_ref, it, _ = ora.pcg(f, tol=1e-10)
    assert st["converged"] == 1 and rel(u, u_ref) < 1e-8 and st["iterations"] == it
    # the damped V(1,1) cycle over the x4 level (RestrictDouble.comp / InterpolateDouble.comp)
    r[np.repeat(ora.fixed(0), 3)] = 0
    # L0_AUTO keeps the row-segment x4 transfers on a level this small; L0_STENCIL
    # forces the plane-tiled ones (k_restrict4_tile, k_prolong4_tile)
    try:
        for kind in (V.L0_AUTO, V.L0_STENCIL):
            A.set_level0_kernel(kind)
            for om in (0.6, 0.5):
                z_ref = ora.vcycle(r, O.DAMPED, om)
                ez = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=om).solve(r), z_ref)
                e32 = rel(V.MultigridPreconditioner(A, cycle=V.CYCLE_DAMPED, omega=om, precision=V.FP32).solve(r), z_ref)
                say("skip level damped cycle form %d omega %.1f: fp64 %.2e fp32 %.2e" % (kind, om, ez, e32))
                assert ez < 1e-12 and e32 < 1e-4
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    u_d, it_d, _ = ora.pcg(f, tol=1e-10, precond=O.DAMPED, omega=0.5)
    for prec in (V.FP64, V.FP32):
        u, st = A.solve(f, tol=1e-10, cycle=V.CYCLE_DAMPED, precision=prec)
        say("skip level damped solve precision %d: %d iterations (oracle %d)" % (prec, st["iterations"], it_d))
        assert st["converged"] == 1 and rel(u, u_d) < 1e-8
        if prec == V.FP64:
            assert st["iterations"] == it_d


@pytest.mark.parametrize("seed", [0, 2, 5, 8, 12, 16])
def test_random_geometries_match_oracle(ke0, seed):
    """Random rough ellipsoid unions (16^3 and 23^3, every third with the skip level):
    the operator and the reference V-cycle agree with the oracle to rounding, the
    FP64 solves (reference cycle, and damped at the library's omega 0.5) reach the
    oracle's solution, FP32 reaches it to 1e-6, and iteration counts agree up to the
    rounding drift of long CG runs (exact below ~200 iterations in practice)."""
    n = 16 + (seed * 7) % 14
    vox, _ = problems.ellipsoid_union(n, 6 + seed % 7, 2.0, n / 3.0, 100 + seed)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    skip = 1 if seed % 3 == 0 else 0
    say("random geometry", seed, n, "skip", skip)
    A = V.assemble_system_matrix(vox, ke0, fixed, 3, skip_levels=skip)
    ora = O.OracleSystem(vox, ke0, fixed, 3, skip=skip)
    r = np.random.default_rng(seed).standard_normal(A.n)
    r[np.repeat(ora.fixed(0), 3)] = 0
    assert rel(A.matvec(r), ora.matvec(r)) < 1e-14
    assert rel(V.MultigridPreconditioner(A).solve(r), ora.vcycle(r, O.REF_ADDITIVE)) < 1e-12
    zs = np.nonzero((ids >= 0).any(axis=(1, 2)))[0]
    top = ids[zs.max()]
    f = np.zeros(A.n)
    f[3 * top[top >= 0] + 2] = -1.0 / np.count_nonzero(top >= 0)
    f[np.repeat(ora.fixed(0), 3)] = 0
    plans = [({}, dict(precond=O.REF_ADDITIVE)), (dict(cycle=V.CYCLE_DAMPED), dict(precond=O.DAMPED, omega=0.5))]
    for lib, orc in plans:
        u_ref, it_ref, _ = ora.pcg(f, tol=1e-10, **orc)
        u, st = A.solve(f, tol=1e-10, **lib)
        assert st["converged"] == 1 and rel(u, u_ref) < 1e-8
        assert abs(st["iterations"] - it_ref) <= max(1, it_ref // 50), (st["iterations"], it_ref)
        if lib:
            u32, s32 = A.solve(f, tol=1e-10, precision=V.FP32, **lib)
            assert s32["converged"] == 1 and rel(u32, u_ref) < 1e-6
    ora.close()
    A.close()


def test_in_cg_stencil_timing(systems):
    """Kind 8 of vfem_time_kernel times the level-0 stencil kernel launch by launch
    inside CG iterations; it rejects out-of-range arguments, and a solve after it
    is bitwise the solve before it (the timing leaves no state behind)."""
    A, ora, _, _ = systems["ell24"]
    f = np.zeros(A.n)
    f[2::3] = -1.0 / (A.n // 3)
    f[np.repeat(ora.fixed(0), 3)] = 0
    say("in-CG stencil timing")
    u1, s1 = A.solve(f, tol=1e-9, precision=V.FP32, cycle=V.CYCLE_DAMPED)
    t8 = A.time_kernel(8, V.FP32, reps=4)
    t7 = A.time_kernel(7, V.FP32, reps=4)
    assert np.isfinite(t8) and 0 < t8 < 100 and 0 < t7 < 100, (t8, t7)
    for which, reps in ((9, 4), (8, 0), (8, 4097)):
        with pytest.raises(V.VfemError):
            A.time_kernel(which, V.FP32, reps=reps)
    A.set_level0_kernel(V.L0_ELEMENT)  # no stencil launches to time
    try:
        with pytest.raises(V.VfemError) as e:
            A.time_kernel(8, V.FP32, reps=2)
        assert e.value.status == V.api.ERR_UNSUPPORTED
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    u2, s2 = A.solve(f, tol=1e-9, precision=V.FP32, cycle=V.CYCLE_DAMPED)
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(u1, u2)


def test_no_graph_path_matches_graph(systems):
    A, ora, _, _ = systems["ell24"]
    f = np.zeros(A.n)
    f[2::3] = -1.0 / (A.n // 3)
    f[np.repeat(ora.fixed(0), 3)] = 0
    say("graph vs stream")
    u1, s1 = A.solve(f, tol=1e-9, use_graph=1)
    u2, s2 = A.solve(f, tol=1e-9, use_graph=0)
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(u1, u2)  # deterministic reductions: bitwise identical


@pytest.mark.parametrize("prec", [V.FP64, V.FP32, V.MIXED])
def test_batched_solve_matches_single_solves(systems, prec):
    """Multi-RHS block: each column converges on its own tolerance and matches the
    single-RHS solve of that column -- bitwise in FP64 (same kernels per column,
    same reduction order), to the solver tolerance in FP32 (the block uses the
    multi-RHS matrix-core matvec).  Columns stop independently."""
    A, ora, vox, fixed = systems["ell24"]
    F = problems.patch_load_cases(vox, 5, seed=11, patch=4)
    F[:, np.repeat(ora.fixed(0), 3)] = 0
    F[3] *= 1e-3          # a column with a different scale converges on its own relative tolerance
    F[4] = 0              # a zero column stops at once (Eigen: x = 0)
    say("batched", prec)
    U, stats = A.solve_batched(F, tol=1e-9, precision=prec, cycle=V.CYCLE_DAMPED)
    assert stats[4]["iterations"] == 0 and not U[4].any()
    for k in range(4):
        u, st = A.solve(F[k], tol=1e-9, precision=prec, cycle=V.CYCLE_DAMPED)
        say("column", k, "iterations", stats[k]["iterations"], st["iterations"], "rel diff", rel(U[k], u))
        assert stats[k]["converged"] == 1 and stats[k]["rel_residual"] < 1e-9
        if prec == V.FP64:
            assert stats[k]["iterations"] == st["iterations"]
            assert np.array_equal(U[k], u)
        else:
            assert rel(U[k], u) < 1e-6


def test_batched_fp64_fused_smoothing_matches_single_solves(systems):
    """Regression: the FP64 damped cycle with the post-smoothing fused into the
    level-0 stencil (forced on this small case by L0_STENCIL) over several columns
    -- every column bitwise equal to its single solve, D^-1 shared by all columns."""
    A, ora, vox, fixed = systems["ell24"]
    F = problems.patch_load_cases(vox, 3, seed=13, patch=4)
    F[:, np.repeat(ora.fixed(0), 3)] = 0
    try:
        A.set_level0_kernel(V.L0_STENCIL)
        U, stats = A.solve_batched(F, tol=1e-9, precision=V.FP64, cycle=V.CYCLE_DAMPED)
        for k in range(3):
            u, st = A.solve(F[k], tol=1e-9, precision=V.FP64, cycle=V.CYCLE_DAMPED)
            say("fused column", k, stats[k]["iterations"], st["iterations"])
            assert stats[k]["converged"] == 1 and stats[k]["iterations"] == st["iterations"]
            assert np.array_equal(U[k], u)
    finally:
        A.set_level0_kernel(V.L0_AUTO)


@pytest.mark.parametrize("cfg", range(len(V.api.block_configs()) - 1 if V.api.os.path.exists(V.api.LIB_PATH) else 0))
def test_block_matvec_parity(systems, cfg):
    """Every tile shape of the multi-RHS matrix-core matvec against the oracle,
    column by column (20 columns: a full 16-column MFMA tile and a partial one)."""
    A, ora, _, _ = systems["ell33"]
    X = np.random.default_rng(5).standard_normal((20, A.n))
    A.set_block_config(cfg)
    say("block matvec config", cfg, V.api.block_configs()[cfg])
    A.set_level0_kernel(V.L0_ELEMENT)  # FP32 at 6 products goes to the stencil form otherwise
    try:
        Y32 = A.matvec_block(X, V.FP32)
        Y16 = A.matvec_block(X, V.BF16)
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    e32 = max(rel(Y32[k], ora.matvec(X[k])) for k in range(20))
    e16 = max(rel(Y16[k], ora.matvec(X[k])) for k in range(20))
    say("block matvec rel err fp32 %.2e bf16 %.2e" % (e32, e16))
    assert e32 < 1e-6 and e16 < 2e-2
    if cfg == 0:
        Y64 = A.matvec_block(X[:3], V.FP64)
        assert max(rel(Y64[k], ora.matvec(X[k])) for k in range(3)) < 1e-13
        Ys = A.matvec_block(X, V.FP32)  # stencil form, one launch per batch (blockIdx.y = column)
        assert max(rel(Ys[k], ora.matvec(X[k])) for k in range(20)) < 2e-6
    A.set_block_config(0)


def test_block_autotuner(systems):
    """The autotuner times every tile shape of the matrix-core kernel and, at FP32
    accuracy, the stencil form; it rejects none (the shapes reproduce shape 0
    bitwise, the stencil form to FP32 rounding) and keeps the fastest, which the
    FP32 batched solve then runs: its columns match their single solves."""
    A, ora, vox, _ = systems["ell33"]
    n = len(V.api.block_configs())
    best, table = A.autotune_block(nrhs=16, passes=6, reps=2)
    say("autotune", best, table)
    assert len(table) == n and table[n - 1][:3][0] == 0   # the stencil candidate
    assert all(v[3] > 0 for v in table.values())
    assert table[best][3] == min(v[3] for v in table.values())
    F = problems.patch_load_cases(vox, 3, seed=17, patch=4)
    F[:, np.repeat(ora.fixed(0), 3)] = 0
    for forced in (False, True):   # the tuned form, then the stencil form selected by hand
        if forced:
            A.set_block_config(n - 1)
        U, stats = A.solve_batched(F, tol=1e-9, precision=V.FP32, cycle=V.CYCLE_DAMPED)
        for k in range(3):
            u, st = A.solve(F[k], tol=1e-9, precision=V.FP32, cycle=V.CYCLE_DAMPED)
            assert stats[k]["converged"] == 1 and rel(U[k], u) < 1e-6, (forced, k)
    A.set_block_config(0)


def test_changed_geometry_warm_start(ke0):
    """Repeated solves on a changing geometry, as in a topology-optimisation loop:
    the solution on geometry A, moved node by node onto geometry B (A with a
    hole), starts B's solve; it reaches B's cold-start solution and the oracle's,
    in fewer iterations."""
    vox_a, _ = problems.ellipsoid_union(32, 16, 4.0, 12.0, 5)
    vox_b = vox_a.copy()
    vox_b[10:16, 12:20, 12:20] = 0
    assert vox_b.sum() < vox_a.sum()

    def setup(vox):
        ids = problems.compact_ids(vox)
        fixed = np.sort(ids[0][ids[0] >= 0])
        zs = np.nonzero((ids >= 0).any(axis=(1, 2)))[0]
        top = ids[zs.max()]
        f = np.zeros(3 * int(ids.max() + 1))
        f[3 * top[top >= 0] + 2] = -1.0 / np.count_nonzero(top >= 0)
        return fixed, f

    fa, f_a = setup(vox_a)
    fb, f_b = setup(vox_b)
    A = V.assemble_system_matrix(vox_a, ke0, fa, 3)
    B = V.assemble_system_matrix(vox_b, ke0, fb, 3)
    kw = dict(tol=1e-10, precision=V.FP64, cycle=V.CYCLE_DAMPED)
    A.set_rhs(f_a)
    sa = A.solve_resident(**kw)
    B.set_rhs(f_b)
    s_cold = B.solve_resident(**kw)
    u_cold = B.get_solution()
    B.map_solution_from(A)
    s_warm = B.solve_resident(warm_start=1, **kw)
    u_warm = B.get_solution()
    ora = O.OracleSystem(vox_b, ke0, fb, 3)
    u_ref, _, _ = ora.pcg(f_b, tol=1e-12, precond=O.DAMPED, omega=0.5)
    ora.close()
    say("changed geometry: A %d iterations; B cold %d, warm from A %d; warm err %.2e"
        % (sa["iterations"], s_cold["iterations"], s_warm["iterations"], rel(u_warm, u_ref)))
    assert s_warm["converged"] == 1 and s_warm["rel_residual"] < 1e-10
    assert rel(u_warm, u_ref) < 1e-8 and rel(u_cold, u_ref) < 1e-8
    assert s_warm["iterations"] < s_cold["iterations"]
    # fixed DOFs of B start (and stay) at zero
    assert not u_warm[np.repeat(np.isin(np.arange(B.n // 3), fb), 3)].any()
    with pytest.raises(V.VfemError):
        B.map_solution_from(V.assemble_system_matrix(vox_a[:, :, :16], ke0, setup(vox_a[:, :, :16])[0], 2))
    A.close()
    B.close()


@pytest.mark.parametrize("name", ["ell24", "ell33"])
def test_bf16_cycle_solve(systems, name):
    """VFEM_BF16: the damped cycle with one-product BF16 level-0 matvecs under the
    FP64 outer CG (flexible, true residual) reaches relative residual 1e-10 and
    the FP64 oracle's displacement to 1e-6, in at most twice the FP32 plan's
    iterations."""
    A, ora, vox, fixed = systems[name]
    ids = problems.compact_ids(vox)
    top = ids[np.nonzero((ids >= 0).any(axis=(1, 2)))[0].max()]
    f = np.zeros(A.n)
    f[3 * top[top >= 0] + 2] = -1.0 / np.count_nonzero(top >= 0)
    u_ref, _, _ = ora.pcg(f, tol=1e-13, precond=O.DAMPED)
    u16, s16 = A.solve(f, tol=1e-10, precision=V.BF16, cycle=V.CYCLE_DAMPED)
    _, s32 = A.solve(f, tol=1e-10, precision=V.FP32, cycle=V.CYCLE_DAMPED)
    say("bf16 cycle %s: %d iterations (fp32 %d), displacement err %.2e"
        % (name, s16["iterations"], s32["iterations"], rel(u16, u_ref)))
    assert s16["converged"] == 1 and s16["rel_residual"] < 1e-10
    assert rel(u16, u_ref) < 1e-6
    assert s16["iterations"] <= 2 * s32["iterations"]


@pytest.mark.parametrize("nx", [61, 62, 63, 64, 65, 124, 127, 129])
def test_row_kernel_tile_boundaries(ke0, nx):
    """The FP32 row-wave stencil kernel on grids whose x extent puts tile and
    64-node word boundaries everywhere (62-node tiles, lane 0 at x = 62 t - 1):
    bit for bit the plane-staged kernel, and the oracle's operator to 2e-6."""
    # a 6 x 7 x nx slab with interior voxel holes (surface nodes throughout),
    # fixed at z = 0
    z, y, xg = np.meshgrid(np.arange(6), np.arange(7), np.arange(nx), indexing="ij")
    vox = (~((xg % 3 == 1) & (y % 3 == 1) & (z % 2 == 1))).astype(np.int32)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[0][ids[0] >= 0])
    A = V.assemble_system_matrix(vox, ke0, fixed, 2)
    ora = O.OracleSystem(vox, ke0, fixed, 2)
    x = np.random.default_rng(nx).standard_normal(A.n)
    ref = ora.matvec(x)
    ora.close()
    try:
        A.set_level0_kernel(V.L0_STENCIL)
        y_rows = A.matvec(x, V.FP32)
        A.set_level0_kernel(V.L0_STENCIL_PLANES)
        y_planes = A.matvec(x, V.FP32)
    finally:
        A.set_level0_kernel(V.L0_AUTO)
    say("row kernel, %d x-nodes: rel err %.2e" % (nx + 1, rel(y_rows, ref)))
    assert np.array_equal(y_rows, y_planes)
    assert rel(y_rows, ref) < 2e-6
    A.close()
