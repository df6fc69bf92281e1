"""The hierarchy vfem_assemble builds on the device against the host builder of
vfem_inspect_hierarchy (vfem_setup.cpp), level by level and bit for bit: node
and fixed counts, the fixed-node masks, the inverse diagonals and the free
coarsest DOFs; at 256^3 also the setup time.  The host builder is itself
pinned to the oracle (tests/test_host.py)."""
import time

import numpy as np
import pytest

import vulkanfem_amd as V
from vulkanfem_amd import problems

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _cases():
    vox24, _ = problems.ellipsoid_union(24, 10, 3.0, 9.0, 7)
    vox31, _ = problems.ellipsoid_union(31, 14, 2.0, 10.0, 5)   # odd extents: partial coarse elements
    box = np.ones((4, 4, 8), np.int32)
    hole = np.ones((12, 10, 20), np.int32)
    hole[3:9, 2:8, 4:15] = 2                                       # non-1 voxels are void
    out = []
    for name, vox, nl, skip in (("ell24", vox24, 4, 0), ("ell31", vox31, 4, 0), ("ell24skip", vox24, 3, 1),
                                ("box1", box, 1, 0), ("hole", hole, 3, 0), ("ell31skip", vox31, 4, 1)):
        ids = problems.compact_ids((vox == 1).astype(np.int32))
        out.append((name, vox, np.sort(ids[0][ids[0] >= 0]), nl, skip))
    return out


def _compare(A, vox, ke, fixed, nl, skip):
    levels = A.num_levels
    for l in range(levels):
        h = V.inspect_hierarchy(vox, ke, fixed, nl, skip, l)
        d = A.level_data(l)
        info = A.level_info(l)
        assert h["levels"] == levels
        assert info["nodes"] == h["nodes"] and info["fixed"] == h["fixed_count"], (l, info, h["nodes"])
        assert np.array_equal(d["fixed"], h["fixed"]), l
        assert np.array_equal(d["invdiag"].view(np.uint64), h["invdiag"].view(np.uint64)), l
        assert d["free_dofs"] == h["free_dofs"]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_device_hierarchy_equals_host_builder(ke0, case):
    name, vox, fixed, nl, skip = case
    say("device setup", name, vox.shape, nl, skip)
    A = V.assemble_system_matrix(vox, ke0, fixed, nl, skip_levels=skip)
    _compare(A, vox, ke0, fixed, nl, skip)
    A.close()


def test_device_setup_errors(ke0):
    with pytest.raises(V.VfemError) as e:
        V.assemble_system_matrix(np.zeros((4, 4, 4), np.int32), ke0, np.zeros(0, np.int64), 2)
    assert e.value.status == V.api.ERR_EMPTY
    box = np.ones((4, 4, 4), np.int32)
    with pytest.raises(V.VfemError) as e:
        V.assemble_system_matrix(box, ke0, np.array([125], np.int64), 2)   # 125 nodes: ids 0..124
    assert e.value.status == V.api.ERR_ARGUMENT
    with pytest.raises(V.VfemError) as e:                                  # no fixed node: Kc singular
        V.assemble_system_matrix(box, ke0, np.zeros(0, np.int64), 2)
    assert e.value.status == V.api.ERR_SINGULAR


def test_device_setup_at_256(ke0):
    p = problems.workload("ellipsoids256")
    nl = problems.WORKLOAD_LEVELS["ellipsoids256"]
    V.assemble_system_matrix(p.voxels, ke0, p.fixed, nl).close()   # first allocations on the device
    t0 = time.perf_counter()
    A = V.assemble_system_matrix(p.voxels, ke0, p.fixed, nl)
    t = time.perf_counter() - t0
    say("256^3 setup %.3f s (library %.1f ms)" % (t, A.solve(p.loads[0], tol=1e-2)[1]["setup_ms"]))
    _compare(A, p.voxels, ke0, p.fixed, nl, 0)
    A.close()
