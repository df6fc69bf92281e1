// Internal data structures shared by the host-side hierarchy setup
// (vfem_setup.cpp) and the device solver (vfem_device.hip).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace vfem {

constexpr int kMaxLevels = 16;
// largest level-0 node count: device kernels pack a node id with two flag bits
// into an int32 and index DOFs as 3 * id
constexpr int64_t kMaxNodes = (int64_t)1 << 29;

// One multigrid level, stored as a grid: every node row (y, z) of the level's
// node grid is split into 64-node segments.  A segment holds a 64-bit mask of
// the used nodes and the compact id of its first used node, so the compact id
// of node (x, y, z) is nbase[seg] + popcount(nbits[seg] & bits below x): the
// reference's compact order (FEM.cpp:179-184) without any per-node index array.
struct HostLevel {
    int ex = 0, ey = 0, ez = 0;      // element extents
    int nx = 0, ny = 0, nz = 0;      // node extents (element extents + 1)
    int rdx = 0, rdy = 0, rdz = 0;   // reference level dims (FEM.cpp:256-259), stop rule only
    int S = 0, SE = 0;               // 64-wide segments per node row / per element row
    int div = 1;                     // coarsening factor from the previous level
    int64_t ne = 0, nn = 0, nfixed = 0;
    double fullscale = 1.0;          // Ke multiplier of full elements (FEM.cpp:463)
    double partscale = 1.0;          // multiplier of the Galerkin sums of partial elements (FEM.cpp:381-383)
    std::vector<uint64_t> nbits;     // [nz*ny*S] used nodes
    std::vector<uint64_t> fbits;     // [nz*ny*S] fixed nodes
    std::vector<int32_t> nbase;      // [nz*ny*S] compact id of the segment's first used node
    std::vector<uint64_t> ebits;     // [ez*ey*SE] present elements
    std::vector<uint8_t> epat;       // [ez*ey*ex] child pattern (div 2 levels), 0 = absent, 255 = full
    std::vector<double> invdiag;     // [3*nn]
};

struct Hierarchy {
    int nlev = 0, skip = 0;
    double ke[576];                  // symmetrised from the upper triangle (FEM.cpp:200-210)
    std::vector<HostLevel> L;
    int64_t nfree = 0;               // free DOFs of the coarsest level
    std::vector<int32_t> freedof;    // coarsest DOF index of each free DOF
    std::vector<double> kcinv_ref;   // inverse of the reference Kc (FEM.cpp:498-513)
    std::vector<double> kcinv_gal;   // inverse of Kc assembled from the level's element matrices
    // element matrices of each coarse level's operator: per child pattern (256 x
    // 576) on a div-2 level; on a div-4 level row 0 is zero (absent element) and
    // row r + 1 belongs to the r-th distinct child mask (DevLevelArrays::eidx);
    // empty at level 0 (uniform Ke)
    std::vector<std::vector<double>> ktab;
};

// Device arrays of one level, built and owned by build_hierarchy_device (the
// pointers are also listed in its `allocs`): row masks as in HostLevel, the
// inverse diagonal in FP64 and FP32, and compact-order bit masks of the fixed
// nodes (fixc) and of the nodes whose D^-1 triple equals dcon bit for bit
// (dflat); dcon is the D^-1 triple of a node with eight full elements.
struct DevLevelArrays {
    uint64_t *nbits = nullptr, *fbits = nullptr, *ebits = nullptr;
    int32_t *nbase = nullptr;
    uint8_t *epat = nullptr;
    int32_t *eidx = nullptr;  // div-4 coarse level: row of each element in H.ktab[l] (0 = absent)
    double *dinv64 = nullptr;
    float *dinv32 = nullptr;
    uint32_t *fixc = nullptr, *dflat = nullptr;
    double dcon[3] = {0, 0, 0};
};

struct DevHierarchy {
    DevLevelArrays L[kMaxLevels];
    double *kinv_ref = nullptr, *kinv_gal = nullptr;  // inverses of the two coarsest matrices (nfree^2)
};

// Runs f(i) for every i in [0, count) on up to 16 host threads.  Indices are
// handed out in ascending order from a shared counter, so callers put the
// costliest rows first; every entry is computed by one fixed sequence of
// operations, so the result does not depend on the thread count.
template <class F> void parallel_for(int64_t count, const F &f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (count < 32) nt = 1;
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (int64_t i; (i = next.fetch_add(1, std::memory_order_relaxed)) < count;) f(i);
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
}

// Pieces shared by the host and the device builders.
int check_hierarchy_arguments(const int32_t *voxels, int vx, int vy, int vz, const double *ke_in, const int64_t *fixed,
                              int64_t nfixed, int num_levels, int skip, std::string &err);
void symmetrise_ke(const double *ke_in, double *ke);   // from the upper triangle (FEM.cpp:200-210)
// extents, coarsening factor and element scales of coarse level l from level l-1 (FEM.cpp:253-272)
void level_dims(const HostLevel &F, HostLevel &C, int l, int skip);
// free DOFs of the coarsest level (H.nfree, H.freedof) and its two matrices on
// them: the reference Kc (FEM.cpp:473-513) and the Galerkin-consistent one of the
// damped cycle (cmask: child masks of the coarsest elements, null for one level)
void coarsest_matrices(Hierarchy &H, const std::vector<uint64_t> *cmask, std::vector<double> &kref,
                       std::vector<double> &kgal);

// Builds the hierarchy on the device (stream: a hipStream_t) into DH, with the
// small host-side parts (extents, counts, masks, element tables, free DOFs) in H;
// H.L[l].invdiag and H.kcinv_* stay empty.  The same results as build_hierarchy,
// bit for bit.  Device allocations are appended to allocs.
int build_hierarchy_device(const int32_t *voxels, int nx, int ny, int nz, const double *ke, const int64_t *fixed,
                           int64_t nfixed, int num_levels, int skip, void *stream, Hierarchy &H, DevHierarchy &DH,
                           std::vector<void *> &allocs, std::string &err);

// Exclusive scan of n int32 counts on the device (stream: a hipStream_t) into
// base[0..n], base[n] = *total; fails when the total exceeds the int32 range.
int device_exclusive_scan(const int32_t *cnt, int64_t n, int32_t *base, int64_t *total, void *stream,
                          std::string &err);

// Builds the hierarchy on the host.  Returns 0 or a vfem_status code; *err gets a message.
int build_hierarchy(const int32_t *voxels, int nx, int ny, int nz, const double *ke, const int64_t *fixed,
                    int64_t nfixed, int num_levels, int skip, Hierarchy &H, std::string &err);

// Galerkin element matrix of a coarse element with the given child mask: sum over
// present children c of P_c^T Ke P_c, P_c the trilinear map from the coarse
// element's corners to child c's corners (div 2: 8 children, div 4: 64).
void galerkin_element(const double *ke, int div, uint64_t mask, double *K);

// Dense operator of a div-2 coarse level l >= 1 on its free DOFs (the damped
// cycle's level operator, from the element tables); freedof[i] = level DOF of
// row i.  Returns 0, or -1 when the level has no element tables.
int assemble_level_dense(const Hierarchy &H, int l, std::vector<double> &A, std::vector<int32_t> &freedof);
bool cholesky_factor(std::vector<double> &A, int64_t n);          // in place, lower, row-major
void cholesky_solve(const std::vector<double> &Lf, int64_t n, double *b);
int node_id(const HostLevel &L, int x, int y, int z);               // compact id or -1
bool node_fixed(const HostLevel &L, int x, int y, int z);

}  // namespace vfem
