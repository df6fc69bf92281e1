/*
 * vfem.hpp -- header-only C++ layer over vfem.h that re-creates the host
 * interface of the reference's GPU solver path, so a caller of
 *
 *     MatrixFreeSparse<double> K = assembleSystemMatrix<double>(voxels, dims, Ke, fixedNodes,
 *                                                               numLevels, skipLevels);
 *     applyBoundaryConditions(f, loadedNodes, fixedNodes);
 *     solveWithKompute(K, f, u);
 *
 * (src/main.cpp:95-107) compiles against this header instead of FEM.hpp /
 * FEMGPU.hpp.  No Eigen: the matrix type is an opaque handle on the device-side
 * system built by vfem_assemble.  Errors from the C-ABI become vfem::Error
 * exceptions here (the reference lets Vulkan errors propagate as exceptions,
 * src/FEMGPU.cpp:32-554).
 */
#ifndef VFEM_HPP
#define VFEM_HPP

#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "vfem.h"

namespace vfem {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int status)
{
    if (status != VFEM_OK && status != VFEM_ERR_NOT_CONVERGED)
        throw Error(status, std::string(vfem_status_string(status)) + ": " + vfem_last_error());
}

/* include/utils.hpp Vec3 / Vec3i: three components, x first. */
template <typename T> struct Vec3 {
    T x, y, z;
};
using Vec3i = Vec3<int>;

/* MatrixFreeSparse<double> (include/MatrixFreeSparse.hpp:31-94): the operator and
 * its multigrid hierarchy.  rows() == cols() == number of DOFs (the reference
 * stores that count in numElements, MatrixFreeSparse.hpp:57-61). */
template <typename scalar> class MatrixFreeSparse;

template <> class MatrixFreeSparse<double> {
  public:
    explicit MatrixFreeSparse(vfem_system *s) : sys_(s, vfem_destroy) {}
    int64_t rows() const { return vfem_num_dofs(sys_.get()); }
    int64_t cols() const { return rows(); }
    int numLevels() const { return vfem_num_levels(sys_.get()); }
    vfem_system *handle() const { return sys_.get(); }

    /* y = A x with fixed rows zeroed (MatrixFreeSparse.hpp:191-215). */
    std::vector<double> operator*(const std::vector<double> &x) const
    {
        std::vector<double> y(x.size());
        check(vfem_matvec(sys_.get(), x.data(), y.data(), VFEM_FP64));
        return y;
    }

    /* Options used by solveWithKompute / solveWithHIP for this matrix: the C-ABI
     * defaults with the reference's compiled tolerance, TOLERANCE 1e-6
     * (include/FEMDefines.hpp:7). */
    vfem_options options() const { return opts_; }
    void setOptions(const vfem_options &o) { opts_ = o; }

    /* Statistics of the last solve. */
    const vfem_stats &lastStats() const { return stats_; }
    vfem_stats &lastStats() { return stats_; }

  private:
    std::shared_ptr<vfem_system> sys_;
    vfem_options opts_ = [] {
        vfem_options o;
        vfem_default_options(&o);
        o.tol = 1e-6;
        return o;
    }();
    vfem_stats stats_{};
};

/* MultigridPreconditioner (include/MultigridPreconditioner.hpp:12-211): z = M r. */
class MultigridPreconditioner {
  public:
    MultigridPreconditioner() { vfem_default_options(&opts_); }
    MultigridPreconditioner &compute(const MatrixFreeSparse<double> &A)
    {
        A_ = &A;
        return *this;
    }
    MultigridPreconditioner &analyzePattern(const MatrixFreeSparse<double> &) { return *this; }
    MultigridPreconditioner &factorize(const MatrixFreeSparse<double> &A) { return compute(A); }
    bool info() const { return A_ != nullptr; }
    void setOptions(const vfem_options &o) { opts_ = o; }
    std::vector<double> solve(const std::vector<double> &r) const
    {
        if (!A_) throw Error(VFEM_ERR_ARGUMENT, "MultigridPreconditioner used before compute()");
        std::vector<double> z(r.size());
        check(vfem_precondition(A_->handle(), r.data(), z.data(), &opts_));
        return z;
    }

  private:
    const MatrixFreeSparse<double> *A_ = nullptr;
    vfem_options opts_;
};

/* assembleSystemMatrix<double> (include/FEM.hpp:33-34, src/FEM.cpp:157-517).
 * voxelModel: x fastest, 1 = solid; fixedNodes: compact node ids. */
template <typename scalar>
MatrixFreeSparse<scalar> assembleSystemMatrix(int *voxelModel, Vec3i dims, double elementStiffness[24][24],
                                              const std::set<uint64_t> &fixedNodes, int numLevels, int skipLevels)
{
    std::vector<int64_t> fixed(fixedNodes.begin(), fixedNodes.end());
    vfem_system *s = nullptr;
    check(vfem_assemble(voxelModel, dims.x, dims.y, dims.z, &elementStiffness[0][0], fixed.data(),
                        (int64_t)fixed.size(), numLevels, skipLevels, &s));
    return MatrixFreeSparse<scalar>(s);
}

/* applyBoundaryConditions (include/FEM.hpp:63-64, src/FEM.cpp:568-584, MATRIX_FREE
 * branch): f[3 node + c] = load.  fixedNodes is unused there as well. */
template <typename scalar>
void applyBoundaryConditions(std::vector<scalar> &f, std::map<uint64_t, Vec3<scalar>> &loadedNodes,
                             const std::set<uint64_t> &)
{
    std::vector<int64_t> nodes;
    std::vector<double> loads;
    for (const auto &kv : loadedNodes) {
        nodes.push_back((int64_t)kv.first);
        loads.push_back(kv.second.x);
        loads.push_back(kv.second.y);
        loads.push_back(kv.second.z);
    }
    check(vfem_apply_boundary_conditions(f.data(), (int64_t)f.size(), nodes.data(), loads.data(),
                                         (int64_t)nodes.size()));
}

/* solveWithKompute<double>(systemMatrix, f, u) (include/FEMGPU.hpp:11-12): u is
 * resized and overwritten (FEMGPU.cpp:551).  The matrix's options select the
 * precision and cycle; the defaults reproduce the reference cycle in FP64. */
template <typename scalar>
void solveWithKompute(const MatrixFreeSparse<scalar> &systemMatrix, const std::vector<scalar> &f,
                      std::vector<scalar> &u)
{
    u.assign(f.size(), 0.0);
    vfem_options o = systemMatrix.options();
    vfem_stats st{};
    int rc = vfem_solve(systemMatrix.handle(), f.data(), u.data(), &o, &st);
    const_cast<MatrixFreeSparse<scalar> &>(systemMatrix).lastStats() = st;
    check(rc);
}

/* The same call under the name of this backend. */
template <typename scalar>
void solveWithHIP(const MatrixFreeSparse<scalar> &systemMatrix, const std::vector<scalar> &f, std::vector<scalar> &u)
{
    solveWithKompute(systemMatrix, f, u);
}

}  // namespace vfem

#endif /* VFEM_HPP */
