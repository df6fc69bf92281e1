"""Benchmark: time to relative residual 1e-8 of the matrix-free MGPCG on one
MI355X, mixed precision vs plain FP64, plus level-0 matvec GDOF/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ellipsoids256]

A step is one complete solve (x0 = 0 to FP64 true relative residual 1e-8) of the
workload's right-hand side, resident in HBM; the timed region covers K steps
bracketed by device synchronisation.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import vulkanfem_amd as V  # noqa: E402
from vulkanfem_amd import problems  # noqa: E402

PEAK_HBM_GBS = 8000.0
# Damped Jacobi weight of the V(1,1) cycle.  The cycle is a positive definite
# preconditioner only while omega * lambda_max(D^-1 A) < 2 on every level; on the
# ellipsoid workloads lambda_max is about 3.8 (surface nodes), so 0.5 stays inside
# the bound, while 0.6 exceeds it and makes the iteration counts erratic.
OMEGA = 0.5
# The measured plans smooth the coarse levels >= 2 with two Chebyshev steps
# (interval [lmax / 60, lmax] of D^-1 A) before and after the coarse correction
# instead of one damped Jacobi step: 46 -> 36 iterations at 256^3 for about 8%
# more time per iteration (levels >= 2 are small).  Level 0 and 1 keep damped
# Jacobi.  CHEB_LEVEL = 0 gives damped Jacobi everywhere.
CHEB_LEVEL = 2
PLAN = dict(cycle=V.CYCLE_DAMPED, omega=OMEGA, cheb_level=CHEB_LEVEL)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def matvec_bytes(n_nodes, n_vox, s):
    # x read once, y written once (3 components of s bytes), bit-packed occupancy and fixed mask
    return 6 * s * n_nodes + math.ceil(n_vox / 8) + math.ceil(n_nodes / 8)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter
    summary (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by the
    same counter's reading for a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
        return rec[workload][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def learned_solve(A, f, nlev, level=3):
    """Damped-cycle solves with the trained U-Net as the level-3 coarse correction
    (falling back to the plain cycle when a polling interval stalls), under the
    FP64 cycle and under the mixed plan (FP32 cycle, FP64 outer CG).  Returns the
    figures and the two solutions."""
    path = os.path.join(ROOT, "vulkanfem_amd", "data", "unet_level3.npy")
    if not os.path.exists(path) or nlev <= level + 1:
        return {}, {}
    net = V.UNet(0)
    net.set_weights(np.load(path))
    A.attach_unet(net, level)
    A.set_rhs(f)
    out, sols = {}, {}
    for name, prec in (("fp64", V.FP64), ("mixed", V.FP32)):
        kw = dict(tol=1e-8, precision=prec, learned=1, **PLAN)
        A.solve_resident(**kw)
        st = A.solve_resident(**kw)
        sols[name] = A.get_solution()
        log("learned", name, st)
        out.update({f"learned_{name}_time_s": round(st["solve_ms"] / 1e3, 5),
                    f"learned_{name}_iterations": st["iterations"],
                    f"learned_{name}_iterations_with_network": st["learned_iters"],
                    f"learned_{name}_fell_back": st["fell_back"],
                    f"learned_{name}_rel_residual": st["rel_residual"]})
    A.attach_unet(None, 0)
    return out, sols


def batched_solves():
    """batched128x16: 16 patch load cases on one 128^3 geometry as one block
    (tile shape from the autotuner), FP32 cycle and FP64 cycle, FP64 block CG."""
    p = problems.workload("batched128x16")
    A = V.assemble_system_matrix(p.voxels, problems.ke0(1.0, 0.3), p.fixed, problems.WORKLOAD_LEVELS["batched128x16"])
    best, table = A.autotune_block(nrhs=p.loads.shape[0], passes=6, reps=3)
    out = {"batched_workload": p.name, "batched_nrhs": int(p.loads.shape[0]),
           "batched_autotuned_tile": list(table[best][:3]), "batched_tile_ms": round(table[best][3], 4)}
    # the single-RHS figure on the same geometry: column 0 alone, FP32 cycle
    A.set_rhs(p.loads[0])
    kw1 = dict(tol=1e-8, precision=V.FP32, **PLAN)
    A.solve_resident(**kw1)
    st1 = A.solve_resident(**kw1)
    single = A.n * st1["iterations"] / (st1["solve_ms"] / 1e3) / 1e9
    out.update({"single_rhs_128_fp32_iterations": st1["iterations"],
                "single_rhs_128_fp32_time_s": round(st1["solve_ms"] / 1e3, 5),
                "single_rhs_128_fp32_gdof_iter_per_s": round(single, 2)})
    A.set_rhs_block(p.loads)
    for name, prec in (("fp64", V.FP64), ("fp32", V.FP32)):
        kw = dict(tol=1e-8, precision=prec, **PLAN)
        A.solve_resident_block(**kw)
        t0 = time.perf_counter()
        st = A.solve_resident_block(**kw)
        t = time.perf_counter() - t0
        its = [s["iterations"] for s in st]
        log(f"batched {name}: {t:.3f} s, iterations {its}")
        out.update({f"batched_{name}_time_s": round(t, 4), f"batched_{name}_max_iterations": max(its),
                    f"batched_{name}_all_converged": all(s["converged"] == 1 for s in st),
                    f"batched_{name}_gdof_iter_per_s": round(A.n * sum(its) / t / 1e9, 2)})
    out["batched_fp32_vs_single_rhs"] = round(out["batched_fp32_gdof_iter_per_s"] / single, 2)
    A.close()
    return out


def skip_level_solves(p, ke):
    """The reference's canonical plan shape (-n levels --skip 1, .vscode/launch.json:12):
    the same workload with the x4 first coarsening, damped cycle, FP32 and FP64."""
    A = V.assemble_system_matrix(p.voxels, ke, p.fixed, problems.WORKLOAD_LEVELS["ellipsoids256"], skip_levels=1)
    A.set_rhs(p.loads[0])
    out = {}
    for name, prec in (("fp32", V.FP32), ("fp64", V.FP64)):
        kw = dict(tol=1e-8, precision=prec, **PLAN)
        A.solve_resident(**kw)
        st = A.solve_resident(**kw)
        log(f"skip level {name}", st)
        out.update({f"skip1_{name}_time_s": round(st["solve_ms"] / 1e3, 5),
                    f"skip1_{name}_iterations": st["iterations"]})
    A.close()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """The host cores this process may use (its CPU affinity, capped by
    OMP_NUM_THREADS when that is set)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(p, nlev, threads, lmax):
    """The oracle (C restatement of the reference CPU path, FP64) on the host
    cores, on the same workload as the GPU solves:
      - a complete timed solve with the measured plans' damped V(1,1) cycle
        (Chebyshev smoothing on levels >= CHEB_LEVEL with the library's
        lambda_max bounds `lmax`, the same as the GPU's FP64 plan) to relative
        residual 1e-8 from x = 0 -- the reported baseline;
      - continued from there to 1e-10 (Eigen's solveWithGuess): the FP64 CPU
        solution every GPU displacement error is measured against;
      - one iteration on one host thread, scaled to the full solve's iteration
        count (an extrapolation: the reference's Eigen path is single-threaded)."""
    from oracle import oracle as O

    O.set_threads(threads)
    ke = O.ke0_closed_form(1.0, 0.3)
    t0 = time.perf_counter()
    s = O.OracleSystem(p.voxels, ke, p.fixed, nlev)
    s.set_chebyshev(CHEB_LEVEL, 2, 60.0, lmax)
    t_setup = time.perf_counter() - t0
    f = p.loads[0]
    t0 = time.perf_counter()
    u8, it8, rr8 = s.pcg(f, tol=1e-8, precond=O.DAMPED, omega=OMEGA)
    t_solve = time.perf_counter() - t0
    log(f"cpu oracle: setup {t_setup:.1f} s, damped FP64 solve {t_solve:.1f} s, {it8} iterations, relres {rr8:.2e}")
    t0 = time.perf_counter()
    u_ref, it_more, rr = s.pcg(f, tol=1e-10, precond=O.DAMPED, omega=OMEGA, x0=u8)
    log(f"cpu oracle: continued to relres {rr:.2e} in {it_more} more iterations ({time.perf_counter() - t0:.1f} s)")
    O.set_threads(1)
    t0 = time.perf_counter()
    s.pcg(f, tol=1e-8, max_iter=1, precond=O.DAMPED, omega=OMEGA)
    t_it1 = (time.perf_counter() - t0) / 2  # the initial cycle + one iteration
    O.set_threads(threads)
    s.close()
    rec = {"value": round(t_solve, 3), "unit": "s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
           "iterations": it8,
           "sample": f"complete solve, not a sample: the oracle's FP64 PCG with the damped V(1,1) cycle (omega "
                     f"{OMEGA}, Chebyshev smoothing on levels >= {CHEB_LEVEL}: the GPU plans' algorithm) on the "
                     f"{p.name} workload to relative residual 1e-8 from x = 0, {it8} iterations, {threads} OpenMP "
                     f"threads; oracle setup {t_setup:.1f} s not counted"}
    extra = {"cpu_setup_s": round(t_setup, 2),
             "cpu_1thread_s_extrapolated": round(t_it1 * (it8 + 1), 1),
             "cpu_1thread_note": "one iteration (plus the initial cycle) on one thread, times the full solve's "
                                 "iteration count + 1"}
    return rec, extra, u_ref


def cpu_reference_cycle(threads):
    """The reference's own algorithm end to end: the oracle's FP64 PCG with the
    additive V-cycle of MultigridPreconditioner.hpp:158-193 on the 128^3
    workload (ellipsoids128, 5 levels), complete solve to 1e-8, next to the GPU's
    FP64 solve of the same cycle."""
    from oracle import oracle as O

    p = problems.workload("ellipsoids128")
    nl = problems.WORKLOAD_LEVELS["ellipsoids128"]
    O.set_threads(threads)
    s = O.OracleSystem(p.voxels, O.ke0_closed_form(1.0, 0.3), p.fixed, nl)
    t0 = time.perf_counter()
    _, it, rr = s.pcg(p.loads[0], tol=1e-8, precond=O.REF_ADDITIVE)
    t_cpu = time.perf_counter() - t0
    s.close()
    A = V.assemble_system_matrix(p.voxels, problems.ke0(1.0, 0.3), p.fixed, nl)
    A.set_rhs(p.loads[0])
    kw = dict(tol=1e-8, precision=V.FP64, cycle=V.CYCLE_REFERENCE)
    A.solve_resident(**kw)
    st = A.solve_resident(**kw)
    A.close()
    log(f"reference cycle at 128^3: cpu {t_cpu:.2f} s ({it} iterations), gpu {st['solve_ms']:.1f} ms "
        f"({st['iterations']} iterations)")
    return {"ref_cycle_128_cpu_s": round(t_cpu, 3), "ref_cycle_128_cpu_iterations": it,
            "ref_cycle_128_gpu_s": round(st["solve_ms"] / 1e3, 5), "ref_cycle_128_gpu_iterations": st["iterations"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ellipsoids256")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-correction solve")
    ap.add_argument("--no-batched", action="store_true", help="skip the 16-RHS batched128x16 block solve")
    ap.add_argument("--no-skip", action="store_true", help="skip the solves with the x4 skip level")
    ap.add_argument("--profile-kernels", type=int, default=0,
                    help="only launch the level-0 matvec variants and the copy reference this many times each "
                         "(for rocprofv3 counter runs), then exit")
    ap.add_argument("--profile-solve", action="store_true",
                    help="only run one mixed solve with its kernels launched one by one (no graph), for "
                         "rocprofv3 counter runs, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    nlev = problems.WORKLOAD_LEVELS[args.workload]
    p = problems.workload(args.workload)
    ke = problems.ke0(1.0, 0.3)
    log(f"workload {p.name}: ndof={p.ndof} levels={nlev} info={p.info}")
    # HIP runtime start-up (context, first allocations) is not part of the setup time
    q = problems.cantilever_box(8, 4, 4)
    V.assemble_system_matrix(q.voxels, ke, q.fixed, 2).close()
    t0 = time.perf_counter()
    A = V.assemble_system_matrix(p.voxels, ke, p.fixed, nlev)
    t_setup = time.perf_counter() - t0
    n_nodes = A.n // 3
    n_vox = int(np.prod(p.voxels.shape))
    log(f"setup {t_setup:.2f} s, levels {A.num_levels}")
    A.set_rhs(p.loads[0])
    if args.profile_kernels:
        for which, prec_k, form, name in (
                (8, V.FP32, V.L0_AUTO, "stencil kernel fp32 inside CG iterations"),
                (7, V.FP32, V.L0_AUTO, "stencil kernel fp32"), (0, V.FP32, V.L0_AUTO, "operator fp32 (stencil + surface)"),
                (7, V.FP32, V.L0_STENCIL_PLANES, "plane-staged sum-factorised stencil kernel fp32"),
                (7, V.FP64, V.L0_STENCIL_PLANES, "plane-staged sum-factorised stencil kernel fp64"),
                (7, V.FP32, V.L0_STENCIL_DIRECT, "structured (not sum-factorised) stencil kernel fp32"),
                (7, V.FP64, V.L0_AUTO, "stencil kernel fp64"), (0, V.FP64, V.L0_AUTO, "operator fp64 (stencil + surface)"),
                (0, V.BF16, V.L0_AUTO, "element form bf16 1-pass (matrix cores)"),
                (0, V.FP32, V.L0_ELEMENT, "element form fp32 6-pass (matrix cores)"),
                (0, V.FP64, V.L0_ELEMENT, "element form fp64 (matrix cores)"),
                (4, V.FP32, V.L0_AUTO, "copy fp32 vector"), (5, V.FP32, V.L0_AUTO, "copy fp64 vector")):
            A.set_level0_kernel(form)
            ms = A.time_kernel(which, prec_k, reps=args.profile_kernels)
            log(f"{name}: {ms * 1e3:.1f} us")
        A.set_level0_kernel(V.L0_AUTO)
        return
    prec = V.FP32 if args.precision == "fp32" else V.BF16
    mixed = dict(tol=1e-8, precision=prec, true_res_every=20, **PLAN)
    if args.profile_solve:
        log("profile solve", A.solve_resident(use_graph=0, **mixed))
        return

    for _ in range(args.warmup):
        st = A.solve_resident(**mixed)
        log("warmup", st)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(A.solve_resident(**mixed))
    t_mixed = (time.perf_counter() - t0) / args.steps
    u_mixed = A.get_solution()
    log("mixed", stats[-1], f"{t_mixed:.4f} s/solve")

    out_extra = {}
    sols = {"mixed": u_mixed}
    # the same plans with damped Jacobi on every level
    jac = dict(mixed, cheb_level=0)
    A.solve_resident(**jac)
    st_j = A.solve_resident(**jac)
    log("mixed, damped Jacobi everywhere", st_j)
    out_extra.update({"mixed_jacobi_time_s": round(st_j["solve_ms"] / 1e3, 5),
                      "mixed_jacobi_iterations": st_j["iterations"]})
    if not args.no_fp64:
        fp64 = dict(tol=1e-8, precision=V.FP64, **PLAN)
        A.solve_resident(**fp64)
        t0 = time.perf_counter()
        st64 = A.solve_resident(**fp64)
        t_fp64 = time.perf_counter() - t0
        sols["fp64"] = A.get_solution()
        log("fp64", st64, f"{t_fp64:.4f} s/solve")
        fp64j = dict(fp64, cheb_level=0)
        A.solve_resident(**fp64j)
        t0 = time.perf_counter()
        st64j = A.solve_resident(**fp64j)
        t_fp64j = time.perf_counter() - t0
        log("fp64, damped Jacobi everywhere", st64j, f"{t_fp64j:.4f} s/solve")
        out_extra.update({"fp64_jacobi_time_s": round(t_fp64j, 5), "fp64_jacobi_iterations": st64j["iterations"],
                          "speedup_vs_fp64_jacobi": round(t_fp64j / t_mixed, 3)})
        # the reference's own additive cycle in FP64
        st_ref = A.solve_resident(tol=1e-8, precision=V.FP64, cycle=V.CYCLE_REFERENCE)
        # FP64 vectors with the level-0 matvecs on the matrix cores (VFEM_MIXED)
        mx = dict(tol=1e-8, precision=V.MIXED, **PLAN)
        A.solve_resident(**mx)
        stmx = A.solve_resident(**mx)
        log("mixed-fp64-vectors", stmx)
        out_extra.update({"mixed_fp64_vectors_time_s": round(stmx["solve_ms"] / 1e3, 5),
                          "mixed_fp64_vectors_iterations": stmx["iterations"]})
        out_extra.update({"fp64_time_s": round(t_fp64, 5), "fp64_iterations": st64["iterations"],
                          "fp64_rel_residual": st64["rel_residual"],
                          "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
                          "ref_cycle_fp64_iterations": st_ref["iterations"],
                          "ref_cycle_fp64_time_s": round(st_ref["solve_ms"] / 1e3, 5)})
    # level-0 matvec: HIP-event timing on the library's stream
    # the roofline kernel: the FP32 level-0 stencil kernel timed launch by launch
    # inside 20 CG iterations of the FP32 plan (kind 8), and alone (kind 7, the
    # same launch repeated on one vector, which the last-level cache partly holds)
    t_st32_it = A.time_kernel(8, V.FP32, reps=20) / 1e3
    t_st32 = A.time_kernel(7, V.FP32, reps=20) / 1e3
    t_st64 = A.time_kernel(7, V.FP64, reps=20) / 1e3
    t_mv32 = A.time_kernel(0, V.FP32, reps=20) / 1e3  # operator: stencil kernel + surface pass
    t_mv64 = A.time_kernel(0, V.FP64, reps=20) / 1e3
    t_mv16 = A.time_kernel(0, V.BF16, reps=20) / 1e3
    A.set_level0_kernel(V.L0_ELEMENT)
    t_el32 = A.time_kernel(0, V.FP32, reps=20) / 1e3
    t_el64 = A.time_kernel(0, V.FP64, reps=20) / 1e3
    A.set_level0_kernel(V.L0_AUTO)
    t_copy = A.time_kernel(4, V.FP32, reps=20) / 1e3
    t_it = A.time_kernel(2, prec, reps=5) / 1e3
    b32 = matvec_bytes(n_nodes, n_vox, 4)
    achieved = b32 / t_st32_it / 1e9
    out_extra.update({"matvec_gdofs_fp32": round(A.n / t_mv32 / 1e9, 2), "matvec_gdofs_fp64": round(A.n / t_mv64 / 1e9, 2),
                      "matvec_gdofs_bf16_1pass": round(A.n / t_mv16 / 1e9, 2),
                      "matvec_ms_fp32": round(t_mv32 * 1e3, 4), "matvec_ms_fp64": round(t_mv64 * 1e3, 4),
                      "stencil_kernel_ms_fp32_in_cg": round(t_st32_it * 1e3, 4),
                      "stencil_kernel_ms_fp32": round(t_st32 * 1e3, 4), "stencil_kernel_ms_fp64": round(t_st64 * 1e3, 4),
                      "matvec_ms_bf16_1pass": round(t_mv16 * 1e3, 4),
                      "element_form_ms_fp32_6pass": round(t_el32 * 1e3, 4), "element_form_ms_fp64": round(t_el64 * 1e3, 4),
                      "copy_fp32_vector_gbs": round(2 * 4 * A.n / t_copy / 1e9, 1),
                      "iteration_ms": round(t_it * 1e3, 4), "iterations": stats[-1]["iterations"],
                      "rel_residual": stats[-1]["rel_residual"], "gdof_iter_per_s":
                          round(A.n * stats[-1]["iterations"] / t_mixed / 1e9, 2),
                      "setup_s": round(t_setup, 3), "ndof": A.n})
    if not args.no_learned:
        rec_l, sol_l = learned_solve(A, p.loads[0], nlev)
        out_extra.update(rec_l)
        sols.update({f"learned_{k}": v for k, v in sol_l.items()})
    if not args.no_batched:
        out_extra.update(batched_solves())
    if not args.no_skip:
        out_extra.update(skip_level_solves(p, ke))
    kernel = "k_matvec_rows<float,1,12,false>"
    traffic = measured_traffic(kernel, p.name)
    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        cpu, extra, u_ref = cpu_baseline(p, nlev, threads, [A.level_lmax(l) for l in range(A.num_levels)])
        out_extra.update(extra)
        # every GPU solution against the FP64 CPU oracle's (solved to 1e-10)
        nref = np.linalg.norm(u_ref)
        out_extra["disp_rel_err_vs_cpu_oracle"] = {k: float(np.linalg.norm(v - u_ref) / nref) for k, v in sols.items()}
        log("displacement errors vs the CPU oracle", out_extra["disp_rel_err_vs_cpu_oracle"])
        out_extra.update(cpu_reference_cycle(threads))
    rec = {
        "metric": "time to rel-res 1e-8 (s) + matvec GDOF/s, 256^3 voxels, mixed vs plain FP64",
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if prec == V.FP32 else "f64+bf16",
        "data": "synthetic",
        "config": {"workload": p.name, "voxels": list(p.voxels.shape[::-1]), "ndof": A.n, "levels": A.num_levels,
                   "fill": round(p.info.get("fill", 1.0), 4), "components": p.info.get("components", 1),
                   "kept_components": p.info.get("kept_components", 1),
                   "cycle": f"damped V(1,1), omega {OMEGA} on levels 0-{CHEB_LEVEL - 1}, 2 Chebyshev steps on "
                            f"levels >= {CHEB_LEVEL}" if CHEB_LEVEL else f"damped V(1,1), omega {OMEGA}",
                   "cg": "FP64 flexible PCG (x, r, dots, alpha, beta in FP64)", "tol": 1e-8,
                   "true_res_every": 20},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": kernel + " (level-0 sum-factorised stencil matvec of the CG operator and the cycle, "
                               "one wavefront per node row; its surface pass k_stencil_fix runs after it)",
                     "duration_us": round(t_st32_it * 1e6, 2),
                     "timing": "mean of the kernel's launches inside 20 CG iterations of the FP32 plan, each "
                               "between its own HIP events on the library stream",
                     "algorithmic_bytes_per_launch": b32},
        "cpu_baseline": cpu,
    }
    rec.update(out_extra)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
