"""Python host interface over the C-ABI of include/vfem.h (libvfem.so).

Mirrors the reference's solver interface for the GPU path:

    reference (hirademirkol/VulkanFEM)                       here
    assembleSystemMatrix<double>(...)  FEM.hpp:33-34         assemble_system_matrix(...)
    MatrixFreeSparse<double>           MatrixFreeSparse.hpp  MatrixFreeSparse
    applyBoundaryConditions(...)       FEM.hpp:63-64         apply_boundary_conditions(...)
    solveWithKompute<double>(A, f, u)  FEMGPU.hpp:11-12      solve_with_hip(A, f, u)
    MultigridPreconditioner            MultigridPreconditioner.hpp  MultigridPreconditioner

The library has no CPU fallback: importing works anywhere, but every call that
needs the device raises VfemError when the HIP library is missing or fails.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvfem.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "vfem.h")

FP64, FP32, BF16, MIXED = 0, 1, 2, 3
CYCLE_REFERENCE, CYCLE_DAMPED = 0, 1
L0_AUTO, L0_ELEMENT, L0_STENCIL, L0_STENCIL_DIRECT, L0_STENCIL_PLANES = 0, 1, 2, 3, 4

OK, ERR_ARGUMENT, ERR_EMPTY, ERR_DEVICE, ERR_SINGULAR, ERR_NOT_CONVERGED, ERR_BREAKDOWN, ERR_UNSUPPORTED = range(8)


class VfemError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"vfem status {status}: {message}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("cycle", ctypes.c_int32), ("omega", ctypes.c_double), ("true_res_every", ctypes.c_int32),
                ("check_every", ctypes.c_int32), ("use_graph", ctypes.c_int32), ("warm_start", ctypes.c_int32),
                ("op_passes", ctypes.c_int32), ("cycle_passes", ctypes.c_int32), ("learned", ctypes.c_int32),
                ("stall_ratio", ctypes.c_double), ("op_fp64", ctypes.c_int32),
                ("coarse_fp64", ctypes.c_int32), ("cheb_level", ctypes.c_int32), ("cheb_degree", ctypes.c_int32),
                ("cheb_ratio", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("rel_residual", ctypes.c_double),
                ("rec_residual", ctypes.c_double), ("solve_ms", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("true_res_count", ctypes.c_int32), ("status", ctypes.c_int32), ("learned_iters", ctypes.c_int32),
                ("fell_back", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def header_functions():
    """Names of every function declared in include/vfem.h."""
    with open(HEADER) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vfem_\w+)\s*\(", text)))


def lib():
    """Load libvfem.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VfemError(ERR_DEVICE, f"{LIB_PATH} is not built; run `python -m vulkanfem_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.vfem_default_options.argtypes = [ctypes.POINTER(Options)]
    L.vfem_default_options.restype = None
    L.vfem_assemble.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f64p, _i64p, ctypes.c_int64,
                                ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(vp)]
    L.vfem_destroy.argtypes = [vp]
    L.vfem_destroy.restype = None
    L.vfem_num_dofs.argtypes = [vp]
    L.vfem_num_dofs.restype = ctypes.c_int64
    L.vfem_num_levels.argtypes = [vp]
    L.vfem_num_levels.restype = ctypes.c_int32
    L.vfem_level_info.argtypes = [vp, ctypes.c_int32] + [ctypes.POINTER(ctypes.c_int64)] * 3
    L.vfem_level_lmax.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    L.vfem_apply_boundary_conditions.argtypes = [_f64p, ctypes.c_int64, _i64p, _f64p, ctypes.c_int64]
    L.vfem_solve.argtypes = [vp, _f64p, _f64p, ctypes.POINTER(Options), ctypes.POINTER(Stats)]
    L.vfem_solve_batched.argtypes = [vp, _f64p, _f64p, ctypes.c_int32, ctypes.POINTER(Options),
                                     ctypes.POINTER(Stats)]
    L.vfem_set_rhs.argtypes = [vp, _f64p]
    L.vfem_set_rhs_block.argtypes = [vp, _f64p, ctypes.c_int32]
    L.vfem_solve_resident_block.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Options), ctypes.POINTER(Stats)]
    L.vfem_get_solution_block.argtypes = [vp, _f64p, ctypes.c_int32]
    L.vfem_solve_resident.argtypes = [vp, ctypes.POINTER(Options), ctypes.POINTER(Stats)]
    L.vfem_get_solution.argtypes = [vp, _f64p]
    L.vfem_matvec.argtypes = [vp, _f64p, _f64p, ctypes.c_int32]
    L.vfem_matvec_block.argtypes = [vp, _f64p, _f64p, ctypes.c_int32, ctypes.c_int32]
    L.vfem_set_level0_kernel.argtypes = [vp, ctypes.c_int32]
    L.vfem_level0_kernel.argtypes = [vp]
    L.vfem_level0_kernel.restype = ctypes.c_int32
    L.vfem_level_operator_form.argtypes = [vp, ctypes.c_int32]
    L.vfem_level_operator_form.restype = ctypes.c_int32
    L.vfem_block_config_count.restype = ctypes.c_int32
    L.vfem_block_config.argtypes = [ctypes.c_int32] + [ctypes.POINTER(ctypes.c_int32)] * 3
    L.vfem_set_block_config.argtypes = [vp, ctypes.c_int32]
    L.vfem_autotune_block.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f64p,
                                      ctypes.POINTER(ctypes.c_int32)]
    L.vfem_precondition.argtypes = [vp, _f64p, _f64p, ctypes.POINTER(Options)]
    L.vfem_dot.argtypes = [vp, _f64p, _f64p, ctypes.POINTER(ctypes.c_double)]
    L.vfem_map_solution.argtypes = [vp, vp]
    L.vfem_level_data.argtypes = [vp, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    L.vfem_dot_noalloc.argtypes = [vp, vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    L.vfem_time_kernel.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_double)]
    L.vfem_inspect_hierarchy.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f64p, _i64p,
                                         ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int64)]
    _f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.vfem_unet_num_params.restype = ctypes.c_int64
    L.vfem_unet_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(vp)]
    L.vfem_unet_destroy.argtypes = [vp]
    L.vfem_unet_destroy.restype = None
    L.vfem_unet_get_weights.argtypes = [vp, _f32p]
    L.vfem_unet_set_weights.argtypes = [vp, _f32p]
    L.vfem_unet_forward.argtypes = [vp, _f32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f32p]
    L.vfem_unet_backward.argtypes = [vp, _f32p, _f32p, ctypes.c_void_p]
    L.vfem_unet_train.argtypes = [vp, _f32p, _f32p, _f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                                  _f32p]
    L.vfem_collect_coarse_pairs.argtypes = [vp, _f64p, ctypes.c_int32, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p,
                                            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.vfem_attach_unet.argtypes = [vp, vp, ctypes.c_int32]
    L.vfem_level_extent.argtypes = [vp, ctypes.c_int32] + [ctypes.POINTER(ctypes.c_int32)] * 3
    L.vfem_status_string.argtypes = [ctypes.c_int]
    L.vfem_status_string.restype = ctypes.c_char_p
    L.vfem_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def block_configs():
    """Tile shapes (tex, tey, mz) of the multi-RHS kernel, by index."""
    L = lib()
    out = []
    for i in range(L.vfem_block_config_count()):
        t, u, m = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        L.vfem_block_config(i, ctypes.byref(t), ctypes.byref(u), ctypes.byref(m))
        out.append((t.value, u.value, m.value))
    return out


def inspect_level_dims(mat, level):
    """Node grid (nz, ny, nx) of `level` of an assembled system."""
    nx, ny, nz = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(mat._lib.vfem_level_extent(mat._h, int(level), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
    return (nz.value, ny.value, nx.value)


class UNet:
    """The learned coarse-grid correction (include/vfem.h, vfem_unet_*)."""

    def __init__(self, seed=0):
        self._lib = lib()
        h = ctypes.c_void_p()
        _check(self._lib.vfem_unet_create(int(seed), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.vfem_unet_destroy(self._h)
            self._h = None

    @staticmethod
    def num_params():
        return int(lib().vfem_unet_num_params())

    def get_weights(self):
        w = np.zeros(self.num_params(), np.float32)
        _check(self._lib.vfem_unet_get_weights(self._h, w))
        return w

    def set_weights(self, w):
        _check(self._lib.vfem_unet_set_weights(self._h, np.ascontiguousarray(w, np.float32)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        B, C, D, H, W = x.shape
        out = np.zeros((B, 3, D, H, W), np.float32)
        _check(self._lib.vfem_unet_forward(self._h, x.reshape(-1), B, D, H, W, out.reshape(-1)))
        return out

    def backward(self, gout, input_grad=True):
        """Weight gradients (flat, layer by layer) and the input gradient of the
        last forward batch for upstream gradient gout."""
        gout = np.ascontiguousarray(gout, np.float32)
        gw = np.zeros(self.num_params(), np.float32)
        B, _, D, H, W = gout.shape
        gx = np.zeros((B, 4, D, H, W), np.float32) if input_grad else None
        _check(self._lib.vfem_unet_backward(self._h, gout.reshape(-1), gw,
                                            gx.ctypes.data_as(ctypes.c_void_p) if gx is not None else None))
        return gw, gx

    def train(self, x, y, mask, epochs=10, batch=32, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        mask = np.ascontiguousarray(mask, np.float32)
        n, _, D, H, W = x.shape
        losses = np.zeros(epochs, np.float32)
        _check(self._lib.vfem_unet_train(self._h, x.reshape(-1), y.reshape(-1), mask.reshape(-1), n, D, H, W,
                                         int(epochs), int(batch), float(lr), int(seed), losses))
        return losses


def _check(rc):
    if rc != OK:
        raise VfemError(rc, lib().vfem_last_error().decode())
    return rc


def options(**kw):
    o = Options()
    lib().vfem_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def _prep(voxels, ke, fixed):
    vox = np.ascontiguousarray(voxels, dtype=np.int32)
    if vox.ndim != 3:
        raise ValueError("voxels must be a (nz, ny, nx) array, x fastest")
    ke = np.ascontiguousarray(ke, dtype=np.float64).reshape(576)
    fixed = np.ascontiguousarray(np.asarray(sorted(set(int(v) for v in np.asarray(fixed).reshape(-1)))),
                                 dtype=np.int64).reshape(-1)
    return vox, ke, fixed


class MatrixFreeSparse:
    """The assembled system: MatrixFreeSparse<double> plus its multigrid data
    (MatrixFreeSparse.hpp:31-94), resident on the GPU."""

    def __init__(self, voxels, ke, fixed_nodes, num_levels=3, skip_levels=0):
        vox, ke, fixed = _prep(voxels, ke, fixed_nodes)
        nz, ny, nx = vox.shape
        L = lib()
        h = ctypes.c_void_p()
        _check(L.vfem_assemble(vox.reshape(-1), nx, ny, nz, ke, fixed, fixed.size, int(num_levels),
                               int(skip_levels), ctypes.byref(h)))
        self._h = h
        self._lib = L
        self.num_levels = int(L.vfem_num_levels(h))
        self.skip_levels = int(skip_levels)
        self.n = int(L.vfem_num_dofs(h))

    def rows(self):
        return self.n

    def cols(self):
        return self.n

    @property
    def numElements(self):  # the reference's name for the DOF count (FEM.cpp:186,221)
        return self.n

    def level_data(self, level):
        """The device-built hierarchy of `level`: inverse diagonal, fixed mask, free coarsest DOFs."""
        nn = self.level_info(level)["nodes"]
        inv = np.zeros(3 * nn)
        mask = np.zeros(nn, np.uint8)
        nfree = ctypes.c_int64()
        _check(self._lib.vfem_level_data(self._h, int(level), inv.ctypes.data_as(ctypes.c_void_p),
                                         mask.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nfree)))
        return {"invdiag": inv, "fixed": mask.astype(bool), "free_dofs": nfree.value}

    def level_lmax(self, level):
        """Upper end of the Chebyshev smoother's interval on `level` (vfem_level_lmax)."""
        out = ctypes.c_double()
        _check(self._lib.vfem_level_lmax(self._h, int(level), ctypes.byref(out)))
        return out.value

    def level_info(self, level):
        ne, nn, nf = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.vfem_level_info(self._h, int(level), ctypes.byref(ne), ctypes.byref(nn), ctypes.byref(nf)))
        return {"elements": ne.value, "nodes": nn.value, "fixed": nf.value}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vfem_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __matmul__(self, x):
        return self.matvec(x)

    def matvec(self, x, precision=FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.n,):
            raise ValueError("x has the wrong size")
        y = np.zeros(self.n)
        _check(self._lib.vfem_matvec(self._h, x, y, int(precision)))
        return y

    def dot(self, a, b):
        out = ctypes.c_double()
        _check(self._lib.vfem_dot(self._h, np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64),
                                  ctypes.byref(out)))
        return out.value

    def solve(self, f, u=None, **opts):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if u is None:
            u = np.zeros(self.n)
        st = Stats()
        o = options(**opts)
        rc = self._lib.vfem_solve(self._h, f, u, ctypes.byref(o), ctypes.byref(st))
        if rc not in (OK, ERR_NOT_CONVERGED):
            _check(rc)
        return u, st.as_dict()

    def solve_batched(self, F, **opts):
        F = np.ascontiguousarray(F, dtype=np.float64)
        nrhs = F.shape[0]
        U = np.zeros_like(F)
        st = (Stats * nrhs)()
        o = options(**opts)
        rc = self._lib.vfem_solve_batched(self._h, F.reshape(-1), U.reshape(-1), nrhs, ctypes.byref(o), st)
        if rc not in (OK, ERR_NOT_CONVERGED):
            _check(rc)
        return U, [s.as_dict() for s in st]

    # resident split, for timing without host transfers
    def set_rhs(self, f):
        _check(self._lib.vfem_set_rhs(self._h, np.ascontiguousarray(f, np.float64)))

    def solve_resident(self, **opts):
        st = Stats()
        o = options(**opts)
        rc = self._lib.vfem_solve_resident(self._h, ctypes.byref(o), ctypes.byref(st))
        if rc not in (OK, ERR_NOT_CONVERGED):
            _check(rc)
        return st.as_dict()

    def map_solution_from(self, other):
        """Start vector of the next warm-started resident solve: the resident solution
        of `other` (same voxel grid extents) moved node by node onto this geometry
        (vfem_map_solution)."""
        _check(self._lib.vfem_map_solution(other._h, self._h))

    def get_solution(self):
        u = np.zeros(self.n)
        _check(self._lib.vfem_get_solution(self._h, u))
        return u

    # resident split of the multi-RHS block (1..256 columns)
    def set_rhs_block(self, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        self._nrhs = F.shape[0]
        _check(self._lib.vfem_set_rhs_block(self._h, F.reshape(-1), self._nrhs))

    def solve_resident_block(self, nrhs=None, **opts):
        nrhs = self._nrhs if nrhs is None else nrhs
        st = (Stats * nrhs)()
        o = options(**opts)
        rc = self._lib.vfem_solve_resident_block(self._h, nrhs, ctypes.byref(o), st)
        if rc not in (OK, ERR_NOT_CONVERGED):
            _check(rc)
        return [s.as_dict() for s in st]

    def get_solution_block(self, nrhs=None):
        nrhs = self._nrhs if nrhs is None else nrhs
        U = np.zeros((nrhs, self.n))
        _check(self._lib.vfem_get_solution_block(self._h, U.reshape(-1), nrhs))
        return U

    def set_level0_kernel(self, kind):
        """Operator form: L0_AUTO, L0_ELEMENT, L0_STENCIL or L0_STENCIL_DIRECT (vfem_set_level0_kernel)."""
        _check(self._lib.vfem_set_level0_kernel(self._h, int(kind)))

    @property
    def level0_kernel(self):
        """Level-0 form in effect: L0_ELEMENT, L0_STENCIL (sum-factorised) or L0_STENCIL_DIRECT."""
        return int(self._lib.vfem_level0_kernel(self._h))

    def level_operator_form(self, level):
        """Operator form in effect on `level`: L0_ELEMENT, L0_STENCIL or L0_STENCIL_DIRECT."""
        return int(self._lib.vfem_level_operator_form(self._h, int(level)))

    def matvec_block(self, X, precision=FP64):
        """Y[k] = A X[k] for every row of X (multi-RHS kernel for FP32/BF16)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.zeros_like(X)
        _check(self._lib.vfem_matvec_block(self._h, X.reshape(-1), Y.reshape(-1), X.shape[0], int(precision)))
        return Y

    def set_block_config(self, index):
        _check(self._lib.vfem_set_block_config(self._h, int(index)))

    def autotune_block(self, nrhs=16, passes=6, reps=5):
        """Time every tile shape of the multi-RHS kernel, keep the fastest; returns
        (best index, {index: (tex, tey, mz, ms)})."""
        n = self._lib.vfem_block_config_count()
        ms = np.zeros(n)
        best = ctypes.c_int32()
        _check(self._lib.vfem_autotune_block(self._h, int(nrhs), int(passes), int(reps), ms, ctypes.byref(best)))
        return best.value, {i: block_configs()[i] + (float(ms[i]),) for i in range(n)}

    def collect_coarse_pairs(self, f, level, n_iter):
        """(x, y, yvc, mask) training pairs on `level`'s node grid from up to n_iter
        FP64 PCG iterations on right-hand side f (vfem_collect_coarse_pairs)."""
        L = self._lib
        hl = inspect_level_dims(self, level)
        V = int(np.prod(hl))
        x = np.zeros((n_iter, 4) + hl, np.float32)
        y = np.zeros((n_iter, 3) + hl, np.float32)
        yvc = np.zeros((n_iter, 3) + hl, np.float32)
        m = np.zeros((n_iter,) + hl, np.float32)
        cnt, grid = ctypes.c_int32(), (ctypes.c_int32 * 3)()
        _check(L.vfem_collect_coarse_pairs(self._h, np.ascontiguousarray(f, np.float64), int(level), int(n_iter),
                                           x.reshape(-1), y.reshape(-1), yvc.reshape(-1), m.reshape(-1),
                                           ctypes.byref(cnt), grid))
        assert tuple(grid) == hl and V > 0
        k = cnt.value
        return x[:k], y[:k], yvc[:k], m[:k]

    def attach_unet(self, net, level):
        """Use `net` (UNet) as the learned coarse correction at `level`; None detaches."""
        self._unet = net
        _check(self._lib.vfem_attach_unet(self._h, net._h if net is not None else None, int(level)))

    def time_kernel(self, which, precision=FP64, reps=10):
        out = ctypes.c_double()
        _check(self._lib.vfem_time_kernel(self._h, int(which), int(precision), int(reps), ctypes.byref(out)))
        return out.value


class MultigridPreconditioner:
    """Eigen preconditioner concept over the device V-cycle
    (MultigridPreconditioner.hpp:24-58,195-203)."""

    def __init__(self, mat=None, **opts):
        self.matrix = None
        self.opts = opts
        if mat is not None:
            self.compute(mat)

    def analyzePattern(self, mat):
        return self

    def factorize(self, mat):
        self.matrix = mat
        return self

    def compute(self, mat):
        return self.factorize(mat)

    def info(self):
        return 0  # Eigen::Success

    def solve(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        o = options(**self.opts)
        _check(self.matrix._lib.vfem_precondition(self.matrix._h, r, z, ctypes.byref(o)))
        return z


def assemble_system_matrix(voxels, ke, fixed_nodes, num_levels=3, skip_levels=0):
    """assembleSystemMatrix<double>(voxelModel, dims, Ke, fixedNodes, numLevels, skipLevels)."""
    return MatrixFreeSparse(voxels, ke, fixed_nodes, num_levels, skip_levels)


def apply_boundary_conditions(f, loaded_nodes, fixed_nodes=None):
    """applyBoundaryConditions(f, loadedNodes, fixedNodes) (FEM.cpp:568-584):
    loaded_nodes maps compact node id -> (x, y, z) load; f is modified in place."""
    items = sorted(loaded_nodes.items())
    ids = np.ascontiguousarray([k for k, _ in items], dtype=np.int64)
    loads = np.ascontiguousarray([v for _, v in items], dtype=np.float64).reshape(-1)
    if not f.flags["C_CONTIGUOUS"] or f.dtype != np.float64:
        raise ValueError("f must be a contiguous float64 array")
    _check(lib().vfem_apply_boundary_conditions(f, f.size, ids, loads, ids.size))
    return f


def solve_with_hip(A, f, u, **opts):
    """solveWithKompute<double>(systemMatrix, f, u): u is overwritten with the
    solution (FEMGPU.cpp:551).  Returns the solver statistics."""
    sol, st = A.solve(f, **opts)
    u[:] = sol
    return st


def dot_noalloc(a_ptr, b_ptr, n):
    """a . b over n doubles at device addresses a_ptr, b_ptr (vfem_dot_noalloc:
    no device allocation, fixed per-device scratch)."""
    out = ctypes.c_double()
    _check(lib().vfem_dot_noalloc(ctypes.c_void_p(a_ptr), ctypes.c_void_p(b_ptr), int(n), ctypes.byref(out)))
    return out.value


def inspect_hierarchy(voxels, ke, fixed_nodes, num_levels, skip_levels, level):
    """Host-only view of the hierarchy the device uses (no GPU call)."""
    vox, ke, fixed = _prep(voxels, ke, fixed_nodes)
    nz, ny, nx = vox.shape
    L = lib()
    nl, nn, nf, nfree = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(L.vfem_inspect_hierarchy(vox.reshape(-1), nx, ny, nz, ke, fixed, fixed.size, num_levels, skip_levels,
                                    level, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(nf), None, None,
                                    ctypes.byref(nfree)))
    if level >= nl.value:
        return {"levels": nl.value}
    inv = np.zeros(3 * nn.value)
    mask = np.zeros(nn.value, np.uint8)
    _check(L.vfem_inspect_hierarchy(vox.reshape(-1), nx, ny, nz, ke, fixed, fixed.size, num_levels, skip_levels,
                                    level, ctypes.byref(nl), ctypes.byref(nn), ctypes.byref(nf),
                                    inv.ctypes.data_as(ctypes.c_void_p), mask.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.byref(nfree)))
    return {"levels": nl.value, "nodes": nn.value, "fixed_count": nf.value, "invdiag": inv,
            "fixed": mask.astype(bool), "free_dofs": nfree.value}
