"""ctypes binding of the CPU FP64 oracle (oracle/vfem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to compare results with.  The product package
(vulkanfem_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with make (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    L.vo_build.argtypes = [_i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _i64p, ctypes.c_int64,
                           ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.vo_build.restype = ctypes.c_int
    L.vo_free.argtypes = [vp]
    L.vo_ke0_closed_form.argtypes = [ctypes.c_double, ctypes.c_double, _f64p]
    L.vo_transfer_weights.argtypes = [ctypes.c_int, _f64p]
    L.vo_child_lattice_map.argtypes = [ctypes.c_int, _i32p]
    L.vo_matvec.argtypes = [vp, _f64p, _f64p]
    L.vo_level_matvec.argtypes = [vp, ctypes.c_int, _f64p, _f64p]
    L.vo_vcycle.argtypes = [vp, _f64p, _f64p, ctypes.c_int, ctypes.c_double]
    L.vo_restrict.argtypes = [vp, ctypes.c_int, _f64p, _f64p]
    L.vo_interpolate.argtypes = [vp, ctypes.c_int, _f64p, _f64p]
    L.vo_pcg.argtypes = [vp, _f64p, _f64p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                         ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.vo_pcg.restype = ctypes.c_int
    L.vo_num_levels.argtypes = [vp]
    L.vo_num_levels.restype = ctypes.c_int
    for name in ("vo_level_nodes", "vo_level_elems", "vo_level_nfixed"):
        getattr(L, name).argtypes = [vp, ctypes.c_int]
        getattr(L, name).restype = ctypes.c_int64
    L.vo_level_extent.argtypes = [vp, ctypes.c_int, _i32p]
    L.vo_get_ecoord.argtypes = [vp, ctypes.c_int, _i32p]
    L.vo_get_e2n.argtypes = [vp, ctypes.c_int, _i32p]
    L.vo_get_invdiag.argtypes = [vp, ctypes.c_int, _f64p]
    L.vo_get_rcoef.argtypes = [vp, ctypes.c_int, _f64p]
    L.vo_get_rcoef.restype = ctypes.c_int
    L.vo_get_fixed.argtypes = [vp, ctypes.c_int, _u8p]
    L.vo_get_rmap.argtypes = [vp, ctypes.c_int, _i32p]
    L.vo_get_rmap.restype = ctypes.c_int
    L.vo_level_nb.argtypes = [vp, ctypes.c_int]
    L.vo_level_nb.restype = ctypes.c_int
    L.vo_get_element_matrix.argtypes = [vp, ctypes.c_int, ctypes.c_int64, _f64p]
    L.vo_nfree.argtypes = [vp]
    L.vo_nfree.restype = ctypes.c_int64
    L.vo_get_freedof.argtypes = [vp, _i32p]
    L.vo_get_kc.argtypes = [vp, ctypes.c_int, _f64p]
    L.vo_set_chebyshev.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _f64p]
    L.vo_set_chebyshev.restype = ctypes.c_int
    L.vo_set_threads.argtypes = [ctypes.c_int]
    L.vo_get_threads.restype = ctypes.c_int
    _lib = L
    return L


def ke0_closed_form(E=1.0, nu=0.3):
    out = np.zeros(576)
    lib().vo_ke0_closed_form(E, nu, out)
    return out.reshape(24, 24)


def transfer_weights(div):
    nb = div + 1
    W = np.zeros(nb ** 3 * 8)
    lib().vo_transfer_weights(div, W)
    return W.reshape(nb ** 3, 8)


def child_lattice_map(div):
    out = np.zeros(div ** 3 * 8, np.int32)
    lib().vo_child_lattice_map(div, out)
    return out.reshape(div ** 3, 8)


def set_threads(n):
    lib().vo_set_threads(int(n))


def get_threads():
    return lib().vo_get_threads()


# preconditioner modes of vo_vcycle / vo_pcg
IDENTITY, REF_ADDITIVE, REF_ADDITIVE_FP32, DAMPED = -1, 0, 1, 2


class OracleSystem:
    """assembleSystemMatrix<double> (FEM.cpp:157-517) restated; owns a vo_sys."""

    def __init__(self, voxels, ke, fixed, num_levels=3, skip=0):
        vox = np.ascontiguousarray(voxels, dtype=np.int32)
        if vox.ndim != 3:
            raise ValueError("voxels must be a (nz, ny, nx) array")
        nz, ny, nx = vox.shape
        ke = np.ascontiguousarray(ke, dtype=np.float64).reshape(576)
        fixed = np.ascontiguousarray(fixed, dtype=np.int64).reshape(-1)
        self._lib = lib()
        h = ctypes.c_void_p()
        rc = self._lib.vo_build(vox.reshape(-1), nx, ny, nz, ke, fixed, fixed.size, int(num_levels), int(skip),
                                ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle build failed with status {rc}")
        self._h = h
        self.num_levels = self._lib.vo_num_levels(h)
        self.ndof = 3 * int(self._lib.vo_level_nodes(h, 0))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vo_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- level data -------------------------------------------------------
    def nodes(self, l=0):
        return int(self._lib.vo_level_nodes(self._h, l))

    def elems(self, l=0):
        return int(self._lib.vo_level_elems(self._h, l))

    def nfixed(self, l=0):
        return int(self._lib.vo_level_nfixed(self._h, l))

    def extent(self, l):
        e = np.zeros(3, np.int32)
        self._lib.vo_level_extent(self._h, l, e)
        return tuple(int(v) for v in e)

    def ecoord(self, l):
        out = np.zeros(3 * self.elems(l), np.int32)
        self._lib.vo_get_ecoord(self._h, l, out)
        return out.reshape(-1, 3)

    def e2n(self, l):
        out = np.zeros(4 * self.elems(l), np.int32)
        self._lib.vo_get_e2n(self._h, l, out)
        return out.reshape(-1, 4)

    def invdiag(self, l):
        out = np.zeros(3 * self.nodes(l))
        self._lib.vo_get_invdiag(self._h, l, out)
        return out

    def rcoef(self, l):
        out = np.zeros(3 * self.nodes(l))
        if self._lib.vo_get_rcoef(self._h, l, out) != 0:
            return None
        return out

    def fixed(self, l):
        out = np.zeros(self.nodes(l), np.uint8)
        self._lib.vo_get_fixed(self._h, l, out)
        return out.astype(bool)

    def rmap(self, l):
        nb = int(self._lib.vo_level_nb(self._h, l))
        out = np.zeros(nb ** 3 * self.elems(l), np.int32)
        self._lib.vo_get_rmap(self._h, l, out)
        return out.reshape(-1, nb ** 3)

    def element_matrix(self, l, e):
        out = np.zeros(576)
        self._lib.vo_get_element_matrix(self._h, l, int(e), out)
        return out.reshape(24, 24)

    def free_dofs(self):
        n = int(self._lib.vo_nfree(self._h))
        out = np.zeros(max(n, 1), np.int32)
        self._lib.vo_get_freedof(self._h, out)
        return out[:n]

    def kc(self, which=0):
        n = int(self._lib.vo_nfree(self._h))
        out = np.zeros(max(n * n, 1))
        self._lib.vo_get_kc(self._h, which, out)
        return out[: n * n].reshape(n, n)

    # ---- operators --------------------------------------------------------
    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ndof)
        self._lib.vo_matvec(self._h, x, y)
        return y

    def level_matvec(self, l, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(3 * self.nodes(l))
        self._lib.vo_level_matvec(self._h, l, x, y)
        return y

    def restrict(self, l, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(3 * self.nodes(l + 1))
        self._lib.vo_restrict(self._h, l, r, out)
        return out

    def interpolate(self, l, xc):
        xc = np.ascontiguousarray(xc, dtype=np.float64)
        out = np.zeros(3 * self.nodes(l))
        self._lib.vo_interpolate(self._h, l, xc, out)
        return out

    def set_chebyshev(self, level, degree=2, ratio=60.0, lmax=None):
        """Damped cycle: Chebyshev smoothing (`degree` steps before and after) on
        levels >= level, fitted to [lmax[l] / ratio, lmax[l]]; level 0 = damped
        Jacobi everywhere."""
        lm = np.zeros(self.num_levels)
        if lmax is not None:
            lm[: len(lmax)] = lmax[: self.num_levels]
        if self._lib.vo_set_chebyshev(self._h, int(level), int(degree), float(ratio), lm) != 0:
            raise ValueError("invalid Chebyshev smoother settings")

    def vcycle(self, r, mode=REF_ADDITIVE, omega=0.6):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.ndof)
        self._lib.vo_vcycle(self._h, r, z, int(mode), float(omega))
        return z

    def pcg(self, b, tol=1e-8, max_iter=30000, precond=REF_ADDITIVE, omega=0.6, x0=None):
        """Returns (x, iterations, final recursive relative residual); x0 = start
        vector (Eigen's solveWithGuess), zero by default."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.ndof) if x0 is None else np.array(x0, dtype=np.float64)
        rr = ctypes.c_double()
        it = self._lib.vo_pcg(self._h, b, x, float(tol), int(max_iter), int(precond), float(omega),
                              ctypes.byref(rr), 0 if x0 is None else 1)
        return x, int(it), float(rr.value)
