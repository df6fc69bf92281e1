"""The command-line drop-in (vulkanfem_amd/vfem_cli) for the reference's solver
program (src/main.cpp:28-127) on models written in the reference's text formats
(src/inout.cpp:6-128): file parsing and error handling on the CPU, and on the
GPU the printed compliance f.u against the oracle's solution of the same model."""
import os
import re
import subprocess

import numpy as np
import pytest

from vulkanfem_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "vulkanfem_amd", "vfem_cli")


def write_model(d, name, voxels, ke, fixed, loaded, loads, write_ke=True):
    """<name>.dat, <name>_Ke0.dat, <name>_fixed.dat, <name>_loading.dat in the
    reference's formats; node ids are written 1-based."""
    nz, ny, nx = voxels.shape
    with open(os.path.join(d, name + ".dat"), "w") as fh:
        fh.write(f"{nx} {ny} {nz}\n")
        for z in range(nz):
            fh.write(",".join(str(int(v)) for v in voxels[z].reshape(-1)) + "\n")
    if write_ke:
        with open(os.path.join(d, name + "_Ke0.dat"), "w") as fh:
            for row in np.asarray(ke).reshape(24, 24):
                fh.write(",".join(repr(float(v)) for v in row) + "\n")
    with open(os.path.join(d, name + "_fixed.dat"), "w") as fh:
        fh.write(f"{len(fixed)}\n")
        for i in fixed:
            fh.write(f"{int(i) + 1}\n")
    with open(os.path.join(d, name + "_loading.dat"), "w") as fh:
        fh.write(f"{len(loaded)}\n")
        for i, (fx, fy, fz) in zip(loaded, loads):
            fh.write(f"{int(i) + 1}\t{float(fx)!r}\t{float(fy)!r}\t{float(fz)!r}\n")


def box_model(tmp_path, ke0, nx=12, ny=8, nz=8):
    vox = np.ones((nz, ny, nx), np.int32)
    ids = problems.compact_ids(vox)
    fixed = np.sort(ids[:, :, 0].reshape(-1))
    loaded = np.sort(ids[:, :, nx].reshape(-1))
    loads = np.tile([0.0, 0.0, -1.0 / loaded.size], (loaded.size, 1))
    write_model(str(tmp_path), "box", vox, ke0, fixed, loaded, loads)
    return vox, fixed, loaded, loads


def run_cli(tmp_path, *args):
    assert os.path.exists(CLI), "vfem_cli not built (python -c 'import __graft_entry__ as g; g.build()')"
    return subprocess.run([CLI, "box.dat", *args], cwd=str(tmp_path), capture_output=True, text=True, timeout=240)


def test_cli_reports_missing_and_malformed_files(tmp_path, ke0):
    out = run_cli(tmp_path)
    assert out.returncode != 0 and "cannot open box.dat" in out.stderr
    box_model(tmp_path, ke0)
    os.remove(os.path.join(str(tmp_path), "box_loading.dat"))
    out = run_cli(tmp_path)
    assert out.returncode != 0 and "box_loading.dat" in out.stderr
    box_model(tmp_path, ke0)
    with open(os.path.join(str(tmp_path), "box.dat"), "w") as fh:
        fh.write("12 8 8\n0,1,0\n")
    out = run_cli(tmp_path)
    assert out.returncode != 0 and "short voxel layer" in out.stderr
    box_model(tmp_path, ke0)
    out = run_cli(tmp_path, "--precision", "fp16")
    assert out.returncode != 0 and "unknown precision" in out.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("plan", [(), ("--skip", "1"), ("--precision", "fp32", "--cycle", "damped"),
                                  ("--skip", "1", "--cycle", "damped"),
                                  ("--precision", "fp32", "--cycle", "damped", "--cheb-level", "1")])
def test_cli_solves_reference_format_model(tmp_path, ke0, plan):
    from oracle import oracle as O

    vox, fixed, loaded, loads = box_model(tmp_path, ke0)
    out = run_cli(tmp_path, "-n", "3", "--tol", "1e-10", *plan)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    m = re.search(r"Compliance:\s+(\S+)", out.stdout)
    assert m, out.stdout
    ndof = 3 * int(problems.compact_ids(vox).max() + 1)
    f = problems.apply_boundary_conditions(ndof, loaded, loads)
    ora = O.OracleSystem(vox, ke0, fixed, 3)
    u_ref, _, _ = ora.pcg(f, tol=1e-12, precond=O.DAMPED)
    ora.close()
    c_ref = float(f @ u_ref)
    assert abs(float(m.group(1)) - c_ref) <= 1e-5 * abs(c_ref), (m.group(1), c_ref)
    with open(os.path.join(str(tmp_path), "box_out.csv")) as fh:
        u = np.array([float(v) for v in fh.read().strip().split(",")])
    assert u.size == ndof
    # the reference's writer keeps six decimals (std::to_string)
    assert np.max(np.abs(u - u_ref)) <= 1e-6


@pytest.mark.gpu
def test_cli_default_tolerance_and_non_binary_voxels(tmp_path, ke0):
    """Without --tol the CLI solves to the reference's compiled TOLERANCE 1e-6
    (include/FEMDefines.hpp:7), and a voxel value other than 1 is void, as in the
    reference's test voxelModel[...] == 1 (src/FEM.cpp:75)."""
    from oracle import oracle as O

    marked = np.ones((8, 8, 12), np.int32)
    marked[2:6, 3:5, 4:9] = 2          # a hole written as 2s: void for the reference
    solid = (marked == 1).astype(np.int32)
    ids = problems.compact_ids(solid)  # node ids of the model with the hole
    fixed = np.sort(ids[:, :, 0].reshape(-1))
    loaded = np.sort(ids[:, :, 12].reshape(-1))
    loads = np.tile([0.0, 0.0, -1.0 / loaded.size], (loaded.size, 1))
    write_model(str(tmp_path), "box", marked, ke0, fixed, loaded, loads)
    out = run_cli(tmp_path, "-n", "3")
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    assert f"Elements: {int(solid.sum())}" in out.stdout
    ndof = 3 * int(problems.compact_ids(solid).max() + 1)
    assert f"Degrees of freedom: {ndof}" in out.stdout
    m = re.search(r"true relative residual (\S+)\)", out.stdout)
    assert m and 1e-10 < float(m.group(1)) < 1e-6, out.stdout
    f = problems.apply_boundary_conditions(ndof, loaded, loads)
    ora = O.OracleSystem(solid, ke0, fixed, 3)
    u_ref, _, _ = ora.pcg(f, tol=1e-12, precond=O.DAMPED)
    ora.close()
    c_ref = float(f @ u_ref)
    c = float(re.search(r"Compliance:\s+(\S+)", out.stdout).group(1))
    assert abs(c - c_ref) <= 1e-4 * abs(c_ref), (c, c_ref)
