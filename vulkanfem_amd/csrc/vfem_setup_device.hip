// Device-side multigrid hierarchy setup (gfx950).
//
// Builds on the GPU what vfem_setup.cpp's build_hierarchy builds on the host --
// the grid form of assembleSystemMatrix<double> (src/FEM.cpp:157-517) -- and
// keeps it resident: per level the row bit masks of present elements, used and
// fixed nodes, the segment bases of the compact numbering, the child patterns
// of coarse elements, the inverse diagonal (FP64 and FP32) and the
// compact-order fixed / shared-D^-1 flag masks.  Every value equals the host
// builder's bit for bit: the same per-node summation order for D^-1, the same
// element tables.  Work per level is a few grid passes, one wavefront per
// 64-node (or 64-element) row segment, the masks formed with ballots:
//   k_pack_voxels      voxels == 1 -> present elements (FEM.cpp:65-87)
//   k_used_nodes       corners of present elements -> used nodes (FEM.cpp:88-99)
//   scan               segment popcounts -> compact id bases (FEM.cpp:179-184)
//   k_mark_fixed       compact fixed ids -> fixed bits (level 0)
//   k_coarse_elements  child masks of coarse elements (FEM.cpp:266-283,364-395)
//   k_coarse_fixed     fixed propagation (FEM.cpp:402-460): coarse node N is
//                      fixed when a fixed fine node lies strictly inside its
//                      trilinear support, |f - div N|_inf < div
//   k_invdiag          D^-1 (FEM.cpp:104-151,463) and the flag masks
// The host keeps what is small: level extents, the Galerkin element tables, the
// coarsest-level matrices (FEM.cpp:473-513) -- whose inverses are formed on the
// device by blocked Gauss-Jordan elimination -- and a copy of the level masks
// (a few MB) for host-side tile lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "vfem_internal.h"
#include "vfem_kernels.h"

namespace vfem {

namespace {

constexpr int kSetupWaves = kBlock / 64;

// ---- kernels -----------------------------------------------------------------

// one wavefront per element-row segment word: bit b = (voxel == 1) at x = 64 s + b
__global__ __launch_bounds__(kBlock) void k_pack_voxels(const int32_t *__restrict__ vox, int ex, int64_t rows, int SE,
                                                        uint64_t *__restrict__ ebits)
{
    const int64_t wid = (int64_t)blockIdx.x * kSetupWaves + (threadIdx.x >> 6);
    if (wid >= rows * SE) return;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int64_t row = wid / SE;
    const int x = 64 * (int)(wid % SE) + lane;
    const bool on = x < ex && vox[row * ex + x] == 1;
    const uint64_t w = __ballot(on);
    if (lane == 0) ebits[wid] = w;
}

__device__ __forceinline__ bool elem_bit(const uint64_t *ebits, int ex, int ey, int ez, int SE, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= ex || y >= ey || z >= ez) return false;
    return (ebits[((int64_t)z * ey + y) * SE + (x >> 6)] >> (x & 63)) & 1;
}

// used nodes: corners of present elements; cnt = popcount per node segment word
__global__ __launch_bounds__(kBlock) void k_used_nodes(const uint64_t *__restrict__ ebits, int ex, int ey, int ez,
                                                       int SE, int nx, int ny, int nz, int S,
                                                       uint64_t *__restrict__ nbits, int32_t *__restrict__ cnt)
{
    const int64_t wid = (int64_t)blockIdx.x * kSetupWaves + (threadIdx.x >> 6);
    if (wid >= (int64_t)ny * nz * S) return;
    const int lane = threadIdx.x & 63;
    const int64_t row = wid / S;
    const int x = 64 * (int)(wid % S) + lane, y = (int)(row % ny), z = (int)(row / ny);
    bool used = false;
    if (x < nx)
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) used |= elem_bit(ebits, ex, ey, ez, SE, x - 1 + i, y - 1 + j, z - 1 + k);
    const uint64_t w = __ballot(used);
    if (lane == 0) {
        nbits[wid] = w;
        cnt[wid] = __popcll(w);
    }
}

// Exclusive scan of n int32 counts into int32 bases (the caller checks the total
// against the int32 range): each of kScanThreads threads owns a contiguous chunk.
constexpr int kScanThreads = 16384;
__global__ __launch_bounds__(kBlock) void k_scan_chunks(const int32_t *__restrict__ in, int64_t n, int64_t m,
                                                        int64_t *__restrict__ sums)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= kScanThreads) return;
    int64_t s = 0;
    const int64_t b = t * m, e = std::min(b + m, n);
    for (int64_t i = b; i < e; i++) s += in[i];
    sums[t] = s;
}

// one block: sums[0..kScanThreads) -> exclusive prefix, sums[kScanThreads] = total
__global__ __launch_bounds__(kBlock) void k_scan_sums(int64_t *__restrict__ sums)
{
    constexpr int PER = kScanThreads / kBlock;
    __shared__ int64_t part[kBlock];
    const int t = threadIdx.x;
    int64_t loc = 0;
    for (int i = 0; i < PER; i++) loc += sums[t * PER + i];
    part[t] = loc;
    __syncthreads();
    if (t == 0) {
        int64_t acc = 0;
        for (int i = 0; i < kBlock; i++) {
            const int64_t v = part[i];
            part[i] = acc;
            acc += v;
        }
        sums[kScanThreads] = acc;
    }
    __syncthreads();
    int64_t acc = part[t];
    for (int i = 0; i < PER; i++) {
        const int64_t v = sums[t * PER + i];
        sums[t * PER + i] = acc;
        acc += v;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_write(const int32_t *__restrict__ in, int64_t n, int64_t m,
                                                       const int64_t *__restrict__ sums, int32_t *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= kScanThreads) return;
    int64_t acc = sums[t];
    const int64_t b = t * m, e = std::min(b + m, n);
    for (int64_t i = b; i < e; i++) {
        out[i] = (int32_t)acc;
        acc += in[i];
    }
}

// fixed compact ids (validated by the host: 0 <= id < nn) -> fixed bits: the
// segment holding id is the last one whose base is <= id
__global__ __launch_bounds__(kBlock) void k_mark_fixed(const int64_t *__restrict__ fixed, int64_t nf,
                                                       const int32_t *__restrict__ nbase,
                                                       const uint64_t *__restrict__ nbits, int64_t nseg,
                                                       uint64_t *__restrict__ fbits)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= nf) return;
    const int64_t id = fixed[k];
    int64_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (nbase[mid] <= id) lo = mid;
        else hi = mid - 1;
    }
    uint64_t w = nbits[lo];
    for (int64_t r = id - nbase[lo]; r > 0; r--) w &= w - 1;
    const int bit = __ffsll((unsigned long long)w) - 1;
    atomicOr((unsigned long long *)&fbits[lo], 1ull << bit);
}

// child masks of the coarse elements (bit dx + div dy + div^2 dz for the present
// child (div cx + dx, div cy + dy, div cz + dz)); one wavefront per coarse
// element-row segment word
__global__ __launch_bounds__(kBlock) void k_coarse_elements(const uint64_t *__restrict__ febits, int fex, int fey,
                                                            int fez, int fSE, int div, int cex, int cey, int cez,
                                                            int cSE, uint64_t *__restrict__ cmask,
                                                            uint8_t *__restrict__ epat, uint64_t *__restrict__ cebits)
{
    const int64_t wid = (int64_t)blockIdx.x * kSetupWaves + (threadIdx.x >> 6);
    if (wid >= (int64_t)cey * cez * cSE) return;
    const int lane = threadIdx.x & 63;
    const int64_t row = wid / cSE;
    const int cx = 64 * (int)(wid % cSE) + lane, cy = (int)(row % cey), cz = (int)(row / cey);
    uint64_t m = 0;
    if (cx < cex) {
        const int fx = div * cx;
        for (int dz = 0; dz < div; dz++)
            for (int dy = 0; dy < div; dy++) {
                const int fy = div * cy + dy, fz = div * cz + dz;
                if (fy >= fey || fz >= fez) continue;
                // div divides 64: the div bits from fx on lie in one word (bits past fex are 0)
                const uint64_t w = febits[((int64_t)fz * fey + fy) * fSE + (fx >> 6)] >> (fx & 63);
                for (int dx = 0; dx < div; dx++)
                    if ((w >> dx) & 1) m |= 1ull << (dx + div * dy + div * div * dz);
            }
        const int64_t e = ((int64_t)cz * cey + cy) * cex + cx;
        cmask[e] = m;
        if (epat) epat[e] = (uint8_t)m;
    }
    const uint64_t w = __ballot(m != 0);
    if (lane == 0) cebits[wid] = w;
}

// fixed propagation onto a coarse level: a used coarse node N is fixed when a
// fixed fine node f has |f_d - div N_d| < div on every axis
__global__ __launch_bounds__(kBlock) void k_coarse_fixed(const uint64_t *__restrict__ ffbits, int fnx, int fny, int fnz,
                                                         int fS, int div, const uint64_t *__restrict__ cnbits, int cnx,
                                                         int cny, int cnz, int cS, uint64_t *__restrict__ cfbits)
{
    const int64_t wid = (int64_t)blockIdx.x * kSetupWaves + (threadIdx.x >> 6);
    if (wid >= (int64_t)cny * cnz * cS) return;
    const int lane = threadIdx.x & 63;
    const int64_t row = wid / cS;
    const int X = 64 * (int)(wid % cS) + lane, Y = (int)(row % cny), Z = (int)(row / cny);
    bool fx = false;
    if (X < cnx && ((cnbits[wid] >> lane) & 1)) {
        const int xa = std::max(div * X - div + 1, 0), xb = std::min(div * X + div - 1, fnx - 1);
        for (int fz = std::max(div * Z - div + 1, 0); fz <= std::min(div * Z + div - 1, fnz - 1) && !fx; fz++)
            for (int fy = std::max(div * Y - div + 1, 0); fy <= std::min(div * Y + div - 1, fny - 1) && !fx; fy++) {
                const int64_t rb = ((int64_t)fz * fny + fy) * fS;
                for (int f = xa; f <= xb; f++)
                    if ((ffbits[rb + (f >> 6)] >> (f & 63)) & 1) {
                        fx = true;
                        break;
                    }
            }
    }
    const uint64_t w = __ballot(fx);
    if (lane == 0) cfbits[wid] = w;
}

struct InvdiagArgs {
    int ex, ey, ez, SE, nx, ny, nz, S;
    const uint64_t *ebits, *nbits, *fbits;
    const int32_t *nbase;
    const uint8_t *epat;    // div-2 coarse level: the table row is the child pattern (255 = full)
    const int32_t *eidx;    // div-4 coarse level: the table row of each element (-1 absent)
    const double *dtab;     // rows of 24 element-matrix diagonals
    int full_row;           // level 0: every present element uses this row
    double dc[3];           // the D^-1 triple of a node with eight full elements
    double *d64;
    float *d32;
    uint32_t *dflat, *fixc; // compact-order flag masks (zeroed)
};

// D^-1 of every used node: the diagonals of its present elements summed in
// ascending element order (z, then y, then x), as an element-by-element scatter
// adds them (FEM.cpp:104-151); one wavefront per node segment word
__global__ __launch_bounds__(kBlock) void k_invdiag(InvdiagArgs a)
{
    const int64_t wid = (int64_t)blockIdx.x * kSetupWaves + (threadIdx.x >> 6);
    if (wid >= (int64_t)a.ny * a.nz * a.S) return;
    const int lane = threadIdx.x & 63;
    const int64_t row = wid / a.S;
    const int x = 64 * (int)(wid % a.S) + lane, y = (int)(row % a.ny), z = (int)(row / a.ny);
    const uint64_t w = a.nbits[wid];
    if (!((w >> lane) & 1)) return;
    const int64_t id = a.nbase[wid] + __popcll(w & ((1ull << lane) - 1));
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                const int ex = x - 1 + i, ey = y - 1 + j, ez = z - 1 + k;
                if (ex < 0 || ey < 0 || ez < 0 || ex >= a.ex || ey >= a.ey || ez >= a.ez) continue;
                const int64_t e = ((int64_t)ez * a.ey + ey) * a.ex + ex;
                int rowi;
                if (a.epat) {
                    rowi = a.epat[e];
                    if (!rowi) continue;
                } else if (a.eidx) {
                    rowi = a.eidx[e];
                    if (rowi < 0) continue;
                } else {
                    if (!((a.ebits[((int64_t)ez * a.ey + ey) * a.SE + (ex >> 6)] >> (ex & 63)) & 1)) continue;
                    rowi = a.full_row;
                }
                const int ln = (1 - i) + 2 * (1 - j) + 4 * (1 - k);
                const double *dv = a.dtab + (int64_t)rowi * 24 + 3 * ln;
                for (int c = 0; c < 3; c++) s[c] += dv[c];
            }
    bool flat = true;
    for (int c = 0; c < 3; c++) {
        const double v = 1.0 / s[c];
        a.d64[3 * id + c] = v;
        a.d32[3 * id + c] = (float)v;
        flat &= v == a.dc[c];
    }
    if (flat) atomicOr(&a.dflat[id >> 5], 1u << (id & 31));
    if ((a.fbits[wid] >> lane) & 1) atomicOr(&a.fixc[id >> 5], 1u << (id & 31));
}

// ---- blocked Gauss-Jordan inverse of an SPD matrix ---------------------------
// A (n x n, row-major, FP64) is replaced by A^-1, block column by block column
// (GB = 64): with D^-1 the inverse of the current diagonal block, the block row
// becomes D^-1 A_k*, every other row i gets A_ij -= A_ik (D^-1 A_kj), the block
// column becomes -A_ik D^-1, and the diagonal block D^-1.  No pivoting: the
// diagonal blocks of an SPD matrix (its successive Schur complements) are SPD.
constexpr int GB = 64;

// inverts the diagonal block (nb <= GB) in LDS by unblocked Gauss-Jordan; a
// non-positive or non-finite pivot sets *bad
__global__ __launch_bounds__(kBlock) void k_gj_diag(const double *__restrict__ A, int n, int kb, int nb,
                                                    double *__restrict__ Dinv, int *bad)
{
    __shared__ double D[GB][GB + 1];
    __shared__ double piv;
    for (int q = threadIdx.x; q < nb * nb; q += kBlock) D[q / nb][q % nb] = A[(int64_t)(kb + q / nb) * n + kb + q % nb];
    __syncthreads();
    for (int t = 0; t < nb; t++) {
        if (threadIdx.x == 0) {
            const double d = D[t][t];
            if (!(d > 0.0) || !isfinite(d)) *bad = 1;
            piv = 1.0 / d;
        }
        __syncthreads();
        const double p = piv;
        // every (i, j) with i != t, j != t from the old row t and column t
        double upd[(GB * GB + kBlock - 1) / kBlock];
        int cntq = 0;
        for (int q = threadIdx.x; q < nb * nb; q += kBlock, cntq++) {
            const int i = q / nb, j = q % nb;
            upd[cntq] = (i != t && j != t) ? D[i][j] - D[i][t] * (D[t][j] * p) : D[i][j];
        }
        __syncthreads();
        cntq = 0;
        for (int q = threadIdx.x; q < nb * nb; q += kBlock, cntq++) {
            const int i = q / nb, j = q % nb;
            if (i == t && j == t) D[i][j] = p;
            else if (i == t) D[i][j] = upd[cntq] * p;
            else if (j == t) D[i][j] = -upd[cntq] * p;
            else D[i][j] = upd[cntq];
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < nb * nb; q += kBlock) Dinv[q] = D[q / nb][q % nb];
}

// R = D^-1 A_k* (nb x n), one thread per entry
__global__ __launch_bounds__(kBlock) void k_gj_row(const double *__restrict__ A, int n, int kb, int nb,
                                                   const double *__restrict__ Dinv, double *__restrict__ R)
{
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nb * n) return;
    const int m = (int)(q / n), j = (int)(q % n);
    double s = 0;
    for (int t = 0; t < nb; t++) s += Dinv[m * nb + t] * A[(int64_t)(kb + t) * n + j];
    R[q] = s;
}

// A_ij -= sum_t A_i,kb+t R_tj for i, j outside the block; 64 x 64 output tiles
// (tile indices on the 64-grid, the block's own tile row and column skipped),
// 4 x 4 outputs per thread, both panels staged in LDS
__global__ __launch_bounds__(kBlock) void k_gj_update(double *__restrict__ A, int n, int kb, int nb,
                                                      const double *__restrict__ R)
{
    __shared__ double Ap[GB][GB];  // Ap[r][t] = A[i0 + r][kb + t]
    __shared__ double Rp[GB][GB];  // Rp[t][c] = R[t][j0 + c]
    const int i0 = blockIdx.y * GB, j0 = blockIdx.x * GB;
    if (i0 == kb || j0 == kb) return;  // block-uniform
    for (int q = threadIdx.x; q < GB * GB; q += kBlock) {
        const int r = q / GB, t = q % GB;
        Ap[r][t] = (i0 + r < n && t < nb) ? A[(int64_t)(i0 + r) * n + kb + t] : 0.0;
        const int tt = q / GB, c = q % GB;
        Rp[tt][c] = (tt < nb && j0 + c < n) ? R[(int64_t)tt * n + j0 + c] : 0.0;
    }
    __syncthreads();
    const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
    double acc[4][4] = {};
    for (int t = 0; t < nb; t++) {
        double a[4], b[4];
        for (int u = 0; u < 4; u++) {
            a[u] = Ap[tr + u][t];
            b[u] = Rp[t][tc + u];
        }
        for (int u = 0; u < 4; u++)
            for (int v = 0; v < 4; v++) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    for (int u = 0; u < 4; u++)
        for (int v = 0; v < 4; v++) {
            const int i = i0 + tr + u, j = j0 + tc + v;
            if (i < n && j < n) A[(int64_t)i * n + j] -= acc[u][v];
        }
}

// C_i* = -A_i,kb.. D^-1 for the rows outside the block (one thread per entry)
__global__ __launch_bounds__(kBlock) void k_gj_col(const double *__restrict__ A, int n, int kb, int nb,
                                                   const double *__restrict__ Dinv, double *__restrict__ C)
{
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)n * nb) return;
    const int i = (int)(q / nb), m = (int)(q % nb);
    if (i >= kb && i < kb + nb) return;
    double s = 0;
    for (int t = 0; t < nb; t++) s += A[(int64_t)i * n + kb + t] * Dinv[t * nb + m];
    C[q] = -s;
}

// writes the new block row (R), block column (C) and diagonal block (D^-1)
__global__ __launch_bounds__(kBlock) void k_gj_finish(double *__restrict__ A, int n, int kb, int nb,
                                                      const double *__restrict__ Dinv, const double *__restrict__ R,
                                                      const double *__restrict__ C)
{
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)n * nb) return;
    // row part: entry (kb + m, j) for q = m * n + j
    {
        const int m = (int)(q / n), j = (int)(q % n);
        if (j < kb || j >= kb + nb) A[(int64_t)(kb + m) * n + j] = R[q];
        else A[(int64_t)(kb + m) * n + j] = Dinv[m * nb + (j - kb)];
    }
    // column part: entry (i, kb + m) for q = i * nb + m
    {
        const int i = (int)(q / nb), m = (int)(q % nb);
        if (i < kb || i >= kb + nb) A[(int64_t)i * n + kb + m] = C[q];
    }
}

// ---- host helpers --------------------------------------------------------------

#define SETUP_TRY(expr)                                                                                        \
    do {                                                                                                       \
        hipError_t _e = (expr);                                                                                \
        if (_e != hipSuccess) {                                                                                \
            err = std::string(#expr) + ": " + hipGetErrorString(_e);                                           \
            return 3;                                                                                          \
        }                                                                                                      \
    } while (0)

struct Ctx {
    hipStream_t s;
    std::vector<void *> &allocs;   // kept: freed with the system
    std::vector<void *> temps;     // freed at the end of the build
    std::string &err;
    ~Ctx()
    {
        for (void *p : temps) (void)hipFree(p);
    }
};

template <typename T> int dalloc(Ctx &c, T **p, size_t n, bool keep, bool zero)
{
    std::string &err = c.err;
    void *q = nullptr;
    SETUP_TRY(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    (keep ? c.allocs : c.temps).push_back(q);
    if (zero) SETUP_TRY(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), c.s));
    *p = (T *)q;
    return 0;
}

int waves_grid(int64_t words) { return (int)((words + kSetupWaves - 1) / kSetupWaves); }

// exclusive scan of cnt[0..n) into base[0..n]; *total = base[n]
int scan_counts(Ctx &c, const int32_t *cnt, int64_t n, int32_t *base, int64_t *total)
{
    std::string &err = c.err;
    int64_t *sums;
    if (int rc = dalloc(c, &sums, kScanThreads + 1, false, false)) return rc;
    const int64_t m = (n + kScanThreads - 1) / kScanThreads;
    const int g = (kScanThreads + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_scan_chunks, dim3(g), dim3(kBlock), 0, c.s, cnt, n, m, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, c.s, sums);
    hipLaunchKernelGGL(k_scan_write, dim3(g), dim3(kBlock), 0, c.s, cnt, n, m, sums, base);
    SETUP_TRY(hipGetLastError());
    SETUP_TRY(hipMemcpyAsync(total, sums + kScanThreads, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    SETUP_TRY(hipStreamSynchronize(c.s));
    if (*total >= (int64_t)INT32_MAX) {
        err = "too many nodes on a level";
        return 1;
    }
    // base[n] = total (the sentinel past the last segment)
    const int32_t t32 = (int32_t)*total;
    SETUP_TRY(hipMemcpyAsync(base + n, &t32, sizeof(int32_t), hipMemcpyHostToDevice, c.s));
    SETUP_TRY(hipStreamSynchronize(c.s));
    return 0;
}

template <typename T> int download(Ctx &c, std::vector<T> &h, const T *d, size_t n)
{
    std::string &err = c.err;
    h.resize(n);
    if (n) SETUP_TRY(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, c.s));
    return 0;
}

// used nodes, bases and counts of level L from its element bits (device)
int build_nodes(Ctx &c, HostLevel &L, DevLevelArrays &D)
{
    L.S = (L.nx + 63) / 64;
    const int64_t nseg = (int64_t)L.ny * L.nz * L.S;
    int32_t *cnt;
    if (int rc = dalloc(c, &D.nbits, nseg, true, false)) return rc;
    if (int rc = dalloc(c, &D.nbase, nseg + 1, true, false)) return rc;
    if (int rc = dalloc(c, &D.fbits, nseg, true, true)) return rc;
    if (int rc = dalloc(c, &cnt, nseg, false, false)) return rc;
    hipLaunchKernelGGL(k_used_nodes, dim3(waves_grid(nseg)), dim3(kBlock), 0, c.s, D.ebits, L.ex, L.ey, L.ez, L.SE, L.nx,
                       L.ny, L.nz, L.S, D.nbits, cnt);
    int64_t total = 0;
    if (int rc = scan_counts(c, cnt, nseg, D.nbase, &total)) return rc;
    L.nn = total;
    return 0;
}

// inverse diagonal and flag masks of level L; dtab rows of 24 diagonals
int build_invdiag(Ctx &c, HostLevel &L, DevLevelArrays &D, const std::vector<double> &dtab, int full_row,
                  const int32_t *eidx)
{
    std::string &err = c.err;
    // D^-1 of a node with eight full elements, summed in the kernel's order
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                const int ln = (1 - i) + 2 * (1 - j) + 4 * (1 - k);
                for (int q = 0; q < 3; q++) s[q] += dtab[(size_t)full_row * 24 + 3 * ln + q];
            }
    for (int q = 0; q < 3; q++) D.dcon[q] = 1.0 / s[q];
    double *dt;
    if (int rc = dalloc(c, &dt, dtab.size(), false, false)) return rc;
    SETUP_TRY(hipMemcpyAsync(dt, dtab.data(), dtab.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
    const size_t nw = (size_t)(L.nn + 31) / 32 + 1;
    if (int rc = dalloc(c, &D.dinv64, 3 * (size_t)L.nn, true, false)) return rc;
    if (int rc = dalloc(c, &D.dinv32, 3 * (size_t)L.nn, true, false)) return rc;
    if (int rc = dalloc(c, &D.dflat, nw, true, true)) return rc;
    if (int rc = dalloc(c, &D.fixc, nw, true, true)) return rc;
    InvdiagArgs a;
    a.ex = L.ex; a.ey = L.ey; a.ez = L.ez; a.SE = L.SE;
    a.nx = L.nx; a.ny = L.ny; a.nz = L.nz; a.S = L.S;
    a.ebits = D.ebits; a.nbits = D.nbits; a.fbits = D.fbits; a.nbase = D.nbase;
    a.epat = D.epat; a.eidx = eidx; a.dtab = dt; a.full_row = full_row;
    for (int q = 0; q < 3; q++) a.dc[q] = D.dcon[q];
    a.d64 = D.dinv64; a.d32 = D.dinv32; a.dflat = D.dflat; a.fixc = D.fixc;
    const int64_t nseg = (int64_t)L.ny * L.nz * L.S;
    hipLaunchKernelGGL(k_invdiag, dim3(waves_grid(nseg)), dim3(kBlock), 0, c.s, a);
    SETUP_TRY(hipGetLastError());
    // the table must outlive the kernel: wait before the temporaries are freed
    SETUP_TRY(hipStreamSynchronize(c.s));
    return 0;
}

// host copies of the level masks (for tile lists and the coarse dense operators)
int download_masks(Ctx &c, HostLevel &L, const DevLevelArrays &D)
{
    std::string &err = c.err;
    const size_t nseg = (size_t)L.ny * L.nz * L.S, eseg = (size_t)L.ey * L.ez * L.SE;
    if (int rc = download(c, L.nbits, D.nbits, nseg)) return rc;
    if (int rc = download(c, L.nbase, D.nbase, nseg + 1)) return rc;
    if (int rc = download(c, L.fbits, D.fbits, nseg)) return rc;
    if (int rc = download(c, L.ebits, D.ebits, eseg)) return rc;
    if (D.epat)
        if (int rc = download(c, L.epat, (const uint8_t *)D.epat, (size_t)L.ex * L.ey * L.ez)) return rc;
    SETUP_TRY(hipStreamSynchronize(c.s));
    L.nfixed = 0;
    for (uint64_t w : L.fbits) L.nfixed += __builtin_popcountll(w);
    L.ne = 0;
    for (uint64_t w : L.ebits) L.ne += __builtin_popcountll(w);
    return 0;
}

// in-place inverse of the SPD matrix dA (n x n) on the device; 4 when not SPD
int gj_inverse(Ctx &c, double *dA, int64_t n)
{
    std::string &err = c.err;
    if (n == 0) return 0;
    double *Dinv, *R, *C;
    int *bad;
    if (int rc = dalloc(c, &Dinv, GB * GB, false, false)) return rc;
    if (int rc = dalloc(c, &R, (size_t)GB * n, false, false)) return rc;
    if (int rc = dalloc(c, &C, (size_t)GB * n, false, false)) return rc;
    if (int rc = dalloc(c, &bad, 1, false, true)) return rc;
    const int ni = (int)n;
    const int tiles = (ni + GB - 1) / GB;
    for (int kb = 0; kb < ni; kb += GB) {
        const int nb = std::min(GB, ni - kb);
        const int g = (int)(((int64_t)nb * n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_gj_diag, dim3(1), dim3(kBlock), 0, c.s, dA, ni, kb, nb, Dinv, bad);
        hipLaunchKernelGGL(k_gj_row, dim3(g), dim3(kBlock), 0, c.s, dA, ni, kb, nb, Dinv, R);
        hipLaunchKernelGGL(k_gj_update, dim3(tiles, tiles), dim3(kBlock), 0, c.s, dA, ni, kb, nb, R);
        hipLaunchKernelGGL(k_gj_col, dim3(g), dim3(kBlock), 0, c.s, dA, ni, kb, nb, Dinv, C);
        hipLaunchKernelGGL(k_gj_finish, dim3(g), dim3(kBlock), 0, c.s, dA, ni, kb, nb, Dinv, R, C);
    }
    SETUP_TRY(hipGetLastError());
    int hbad = 0;
    SETUP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c.s));
    SETUP_TRY(hipStreamSynchronize(c.s));
    if (hbad) {
        err = "the coarsest-level matrix is not positive definite (floating component or no fixed node?)";
        return 4;
    }
    return 0;
}

}  // namespace

int build_hierarchy_device(const int32_t *voxels, int vx, int vy, int vz, const double *ke_in, const int64_t *fixed,
                           int64_t nfixed, int num_levels, int skip, void *stream, Hierarchy &H, DevHierarchy &DH,
                           std::vector<void *> &allocs, std::string &err)
{
    if (int rc = check_hierarchy_arguments(voxels, vx, vy, vz, ke_in, fixed, nfixed, num_levels, skip, err)) return rc;
    Ctx c{(hipStream_t)stream, allocs, {}, err};
    H = Hierarchy();
    DH = DevHierarchy();
    H.skip = skip;
    symmetrise_ke(ke_in, H.ke);

    // ---- level 0 ------------------------------------------------------------
    H.L.emplace_back();
    {
        HostLevel &F = H.L[0];
        F.ex = vx; F.ey = vy; F.ez = vz;
        F.nx = vx + 1; F.ny = vy + 1; F.nz = vz + 1;
        F.rdx = vx; F.rdy = vy; F.rdz = vz;
        F.div = 1;
        F.fullscale = 1.0;
        F.SE = (F.ex + 63) / 64;
        DevLevelArrays &D = DH.L[0];
        const int64_t nvox = (int64_t)vx * vy * vz, erows = (int64_t)vy * vz;
        int32_t *dvox;
        if (int rc = dalloc(c, &dvox, nvox, false, false)) return rc;
        SETUP_TRY(hipMemcpyAsync(dvox, voxels, nvox * sizeof(int32_t), hipMemcpyHostToDevice, c.s));
        if (int rc = dalloc(c, &D.ebits, erows * F.SE, true, false)) return rc;
        hipLaunchKernelGGL(k_pack_voxels, dim3(waves_grid(erows * F.SE)), dim3(kBlock), 0, c.s, dvox, F.ex, erows, F.SE,
                           D.ebits);
        if (int rc = build_nodes(c, F, D)) return rc;
        if (F.nn == 0) {
            err = "the voxel model has no solid voxel";
            return 2;
        }
        // the device kernels pack a node id with two flag bits into an int32
        // (id << 2 | used << 1 | fixed) and index DOFs as 3 * id in 32 bits
        if (F.nn >= kMaxNodes) {
            err = "more than 2^29 - 1 used nodes on level 0 are not supported";
            return 1;
        }
        for (int64_t k = 0; k < nfixed; k++)
            if (fixed[k] < 0 || fixed[k] >= F.nn) {
                err = "fixed node id out of range";
                return 1;
            }
        if (nfixed > 0) {
            int64_t *dfix;
            if (int rc = dalloc(c, &dfix, nfixed, false, false)) return rc;
            SETUP_TRY(hipMemcpyAsync(dfix, fixed, nfixed * sizeof(int64_t), hipMemcpyHostToDevice, c.s));
            hipLaunchKernelGGL(k_mark_fixed, dim3((int)((nfixed + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.s, dfix,
                               nfixed, D.nbase, D.nbits, (int64_t)F.ny * F.nz * F.S, D.fbits);
            SETUP_TRY(hipGetLastError());
        }
        std::vector<double> dtab(24);
        for (int q = 0; q < 24; q++) dtab[q] = F.fullscale * H.ke[q * 24 + q];
        if (int rc = build_invdiag(c, F, D, dtab, 0, nullptr)) return rc;
        if (int rc = download_masks(c, F, D)) return rc;
    }

    // Galerkin element matrices of the 254 partial child patterns of a div-2
    // element (taken at scale 1; each level scales them by a power of two)
    std::vector<double> gal2;
    auto galerkin2 = [&]() {
        if (!gal2.empty()) return;
        gal2.assign(256 * 576, 0.0);
        parallel_for(254, [&](int64_t t) { galerkin_element(H.ke, 2, (uint64_t)(t + 1), &gal2[(size_t)(t + 1) * 576]); });
    };

    // ---- coarse levels --------------------------------------------------------
    int nlev = num_levels;
    uint64_t *cmask_dev = nullptr;  // child masks of the last level built
    for (int l = 1; l < nlev; l++) {
        H.L.emplace_back();
        HostLevel &F = H.L[l - 1];
        HostLevel &C = H.L[l];
        const DevLevelArrays &DF = DH.L[l - 1];
        DevLevelArrays &D = DH.L[l];
        level_dims(F, C, l, skip);
        const int div = C.div;
        C.SE = (C.ex + 63) / 64;
        const int64_t ne_all = (int64_t)C.ex * C.ey * C.ez, erows = (int64_t)C.ey * C.ez;
        if (int rc = dalloc(c, &cmask_dev, ne_all, false, false)) return rc;
        if (div == 2)
            if (int rc = dalloc(c, &D.epat, ne_all, true, false)) return rc;
        if (int rc = dalloc(c, &D.ebits, erows * C.SE, true, false)) return rc;
        hipLaunchKernelGGL(k_coarse_elements, dim3(waves_grid(erows * C.SE)), dim3(kBlock), 0, c.s, DF.ebits, F.ex, F.ey,
                           F.ez, F.SE, div, C.ex, C.ey, C.ez, C.SE, cmask_dev, D.epat, D.ebits);
        if (int rc = build_nodes(c, C, D)) return rc;
        const int64_t cseg = (int64_t)C.ny * C.nz * C.S;
        hipLaunchKernelGGL(k_coarse_fixed, dim3(waves_grid(cseg)), dim3(kBlock), 0, c.s, DF.fbits, F.nx, F.ny, F.nz, F.S,
                           div, D.nbits, C.nx, C.ny, C.nz, C.S, D.fbits);
        SETUP_TRY(hipGetLastError());
        std::vector<double> dtab;
        int32_t *eidx = nullptr;
        int full_row = 255;
        if (div == 2) {
            galerkin2();
            dtab.assign(256 * 24, 0.0);
            for (int p = 1; p < 255; p++)
                for (int q = 0; q < 24; q++) dtab[(size_t)p * 24 + q] = C.partscale * gal2[(size_t)p * 576 + q * 24 + q];
            for (int q = 0; q < 24; q++) dtab[255 * 24 + q] = C.fullscale * H.ke[q * 24 + q];
            // operator table of the level: patterns 0..255
            H.ktab.resize(l + 1);
            std::vector<double> &T = H.ktab[l];
            T.assign(256 * 576, 0.0);
            for (int p = 1; p < 255; p++)
                for (int q = 0; q < 576; q++) T[(size_t)p * 576 + q] = gal2[(size_t)p * 576 + q] * C.partscale;
            for (int q = 0; q < 576; q++) T[255 * 576 + q] = C.fullscale * H.ke[q];
        } else {
            // div 4: one table row per distinct child mask, in order of first appearance
            std::vector<uint64_t> cm;
            if (int rc = download(c, cm, (const uint64_t *)cmask_dev, ne_all)) return rc;
            SETUP_TRY(hipStreamSynchronize(c.s));
            std::unordered_map<uint64_t, int32_t> row_of;
            std::vector<uint64_t> masks;
            std::vector<int32_t> ei(ne_all, -1);
            for (int64_t e = 0; e < ne_all; e++) {
                if (!cm[e]) continue;
                auto it = row_of.emplace(cm[e], (int32_t)masks.size());
                if (it.second) masks.push_back(cm[e]);
                ei[e] = it.first->second;
            }
            dtab.assign(masks.size() * 24 + 24, 0.0);
            parallel_for((int64_t)masks.size(), [&](int64_t t) {
                if (masks[t] == ~0ull) {
                    for (int q = 0; q < 24; q++) dtab[(size_t)t * 24 + q] = C.fullscale * H.ke[q * 24 + q];
                    return;
                }
                std::vector<double> K(576);
                galerkin_element(H.ke, div, masks[t], K.data());
                for (int q = 0; q < 24; q++) dtab[(size_t)t * 24 + q] = C.partscale * K[q * 24 + q];
            });
            // the constant of a node with eight full elements: a full row past the masks
            full_row = (int)masks.size();
            for (int q = 0; q < 24; q++) dtab[(size_t)full_row * 24 + q] = C.fullscale * H.ke[q * 24 + q];
            if (int rc = dalloc(c, &eidx, ne_all, false, false)) return rc;
            SETUP_TRY(hipMemcpyAsync(eidx, ei.data(), ne_all * sizeof(int32_t), hipMemcpyHostToDevice, c.s));
            // the level operator: element matrices of the distinct masks (row 0: absent)
            H.ktab.resize(l + 1);
            std::vector<double> &T = H.ktab[l];
            T.assign((masks.size() + 1) * 576, 0.0);
            parallel_for((int64_t)masks.size(), [&](int64_t t) {
                double *K = &T[(size_t)(t + 1) * 576];
                if (masks[t] == ~0ull) {
                    for (int q = 0; q < 576; q++) K[q] = C.fullscale * H.ke[q];
                    return;
                }
                galerkin_element(H.ke, div, masks[t], K);
                for (int q = 0; q < 576; q++) K[q] *= C.partscale;
            });
            std::vector<int32_t> row1(ne_all);
            for (int64_t e = 0; e < ne_all; e++) row1[e] = ei[e] + 1;
            if (int rc = dalloc(c, &D.eidx, ne_all, true, false)) return rc;
            SETUP_TRY(hipMemcpyAsync(D.eidx, row1.data(), ne_all * sizeof(int32_t), hipMemcpyHostToDevice, c.s));
            SETUP_TRY(hipStreamSynchronize(c.s));
        }
        if (int rc = build_invdiag(c, C, D, dtab, full_row, eidx)) return rc;
        if (int rc = download_masks(c, C, D)) return rc;
        if (C.rdx <= 2 || C.rdy <= 2 || C.rdz <= 2) {  // FEM.cpp:465-470
            nlev = l + 1;
            break;
        }
    }
    H.nlev = nlev;
    H.ktab.resize(nlev);

    // ---- coarsest level: free DOFs, Kc (FEM.cpp:473-513) ------------------------
    HostLevel &Cc = H.L[nlev - 1];
    std::vector<uint64_t> cmask;
    if (nlev > 1) {
        if (int rc = download(c, cmask, (const uint64_t *)cmask_dev, (size_t)Cc.ex * Cc.ey * Cc.ez)) return rc;
        SETUP_TRY(hipStreamSynchronize(c.s));
    }
    std::vector<double> kref, kgal;
    coarsest_matrices(H, nlev > 1 ? &cmask : nullptr, kref, kgal);
    const int64_t n = H.nfree;
    if (n > 0) {
        if (int rc = dalloc(c, &DH.kinv_ref, (size_t)n * n, true, false)) return rc;
        if (int rc = dalloc(c, &DH.kinv_gal, (size_t)n * n, true, false)) return rc;
        SETUP_TRY(hipMemcpyAsync(DH.kinv_ref, kref.data(), kref.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
        SETUP_TRY(hipMemcpyAsync(DH.kinv_gal, kgal.data(), kgal.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
        if (int rc = gj_inverse(c, DH.kinv_ref, n)) return rc;
        if (int rc = gj_inverse(c, DH.kinv_gal, n)) return rc;
    }
    SETUP_TRY(hipStreamSynchronize(c.s));
    return 0;
}

}  // namespace vfem

namespace vfem {

int device_exclusive_scan(const int32_t *cnt, int64_t n, int32_t *base, int64_t *total, void *stream,
                          std::string &err)
{
    std::vector<void *> none;
    Ctx c{(hipStream_t)stream, none, {}, err};
    return scan_counts(c, cnt, n, base, total);
}

}  // namespace vfem
