// HIP kernels of the MGPCG solver (gfx950).
//
// Work decomposition: every level's node grid is cut into rows (y, z) of
// 64-node segments; one wavefront owns one segment, lane i <-> x = 64*s + i.
// Compact ids come from the segment's bit mask and base (mbcnt of the mask),
// so loads of one wavefront touch consecutive compact ids and every index
// computation is wave-uniform.  Blocks of 256 threads (4 wavefronts) take
// contiguous ranges of segments; block ids are remapped so that each of the
// 8 XCDs works on one contiguous slab of the grid and keeps its halo in its
// own L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vfem {

struct DevLevel {
    int nx, ny, nz, S;     // node extents, 64-node segments per row
    int ex, ey, ez, SE;    // element extents, 64-element segments per row
    int div;               // coarsening factor from the previous level
    int64_t nn;            // used nodes
    const uint64_t *nbits; // [nz*ny*S]
    const int32_t *nbase;  // [nz*ny*S]
    const uint64_t *fbits; // [nz*ny*S]
    const uint64_t *ebits; // [ez*ey*SE]
    const uint8_t *epat;   // [ez*ey*ex] child pattern (div-2 coarse levels)
    const int32_t *eidx;   // [ez*ey*ex] element-table row (div-4 coarse level), 0 = absent
};

// Scalars of the CG recurrence, resident on the device.
struct CGState {
    double rhs2;       // |f|^2
    double threshold;  // max(tol^2 |f|^2, DBL_MIN)  (Eigen / FEMGPU.cpp:416)
    double absNew;     // r.z
    double absOld;
    double pAp;
    double alpha;
    double beta;
    double rr;         // r.r
    double result;     // scratch reduction result (vfem_dot)
    double rzx;        // r_{k+1}.z_k, reduced by the residual update (flexible beta without z_old)
    double alpha_prev; // alpha of the deferred x term (x_lag bit 1)
    double tol;
    int iter;          // completed iterations
    int maxit;
    int done;          // 0 running, 1 converged, 2 iteration cap, 3 breakdown
    int last_true;     // 1 if the last convergence test used the FP64 true residual
    int no_stop;       // kernel timing: finalizers never end the iteration
    int x_lag;         // terms x still lacks: bit 0 alpha p (the residual update left it to the p
                       // update), bit 1 alpha_prev p_prev (a p update deferred it to the next)
    int xmode;         // what the next p update does with x (XM_*), set by the finalizer before it
    int xpair;         // 1: p updates may defer their x term and apply two at once
};
static_assert(sizeof(CGState) == 128, "one 128-byte CG state per column");

// x work of a p update (CGState::xmode)
enum XMode {
    XM_NONE = 0,    // x is complete
    XM_SINGLE = 1,  // x += alpha p
    XM_STASH = 2,   // p_prev = p, x untouched (alpha p deferred)
    XM_PAIR = 3,    // x += alpha_prev p_prev + alpha p
};

enum Finalize {
    FIN_STORE = 0,    // state->result = v
    FIN_PAP = 1,      // pAp = v; alpha = absNew / pAp
    FIN_RR = 2,       // rr = v; converged if rr < threshold
    FIN_RZ_INIT = 3,  // absNew = v (first direction)
    FIN_RZ = 4,       // absOld = absNew; absNew = v; beta; iter++
    FIN_RHS = 5,      // rhs2 = v; threshold; zero right-hand side -> done
    FIN_RR_TRUE = 6,  // rr = v (true residual); converged test, else continue
    FIN_RZ_FLEX = 7,  // v = r.z, v2 = r.z_old: Polak-Ribiere beta (flexible CG); iter++
    FIN_RZX = 8,      // rzx = v (r.z_k of a replaced residual)
    FIN_RZ_XFLEX = 9, // v = r.z: Polak-Ribiere beta with r.z_old = rzx; iter++
    FIN_RR_LAZY = 10  // FIN_RR after an update that left x += alpha p to k_cg_pupdate
};

constexpr int kMaxPartials = 4096;
constexpr int kMaxCols = 256;  // right-hand sides of one batched solve

struct Reduction {
    double *partials;   // [2 * cap]: block partials of up to two sums
    unsigned *counter;  // zero between launches (the last block resets it)
    int cap;            // >= the grid of any reducing launch
};

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

}  // namespace vfem
