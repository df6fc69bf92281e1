// Host-side multigrid hierarchy setup.
//
// Computes, from the voxel grid, what assembleSystemMatrix<double> computes in
// the reference (src/FEM.cpp:157-517, MATRIX_FREE + MULTIGRID branch), but as
// grid data for the device kernels instead of element->node tables:
//   - compact node numbering (FEM.cpp:88-99,179-184) as per-row bit masks;
//   - coarse elements = floor(fine / div) (FEM.cpp:266-283) with their child masks;
//   - fixed nodes of a coarse level (FEM.cpp:402-460): the restriction of the
//     fixed indicator is non-zero at coarse node N exactly when a fixed fine
//     node f lies strictly inside N's trilinear support, |f - div*N|_inf < div;
//   - inverse diagonals (FEM.cpp:104-151,463) with the Galerkin element
//     matrices of partially filled coarse elements (FEM.cpp:364-395), which
//     are the sums over present children of P_c^T Ke P_c (P_c = the trilinear
//     map of the coarse corners onto child c's corners);
//   - the coarsest-level matrix Kc on free DOFs (FEM.cpp:473-513) and its
//     inverse for the device GEMV.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "vfem_internal.h"

namespace vfem {

namespace {

inline int64_t lin(int x, int y, int z, int nx, int ny) { return (int64_t)x + (int64_t)nx * (y + (int64_t)ny * z); }

// w(c, L, k): weight of coarse corner k at child c's local node L
double child_weight(int div, int c, int L, int k)
{
    int cx = c % div, cy = (c / div) % div, cz = c / (div * div);
    double w = 1.0;
    int lat[3] = {cx + (L & 1), cy + ((L >> 1) & 1), cz + ((L >> 2) & 1)};
    int kk[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
    for (int d = 0; d < 3; d++) w *= 1.0 - std::fabs((double)lat[d] / div - kk[d]);
    return w;
}

// Pack a dense used-node grid into row segments; returns the number of used nodes.
int64_t pack_nodes(HostLevel &L, const std::vector<uint8_t> &used)
{
    L.S = (L.nx + 63) / 64;
    int64_t rows = (int64_t)L.ny * L.nz;
    L.nbits.assign(rows * L.S, 0);
    L.nbase.assign(rows * L.S + 1, 0);
    L.fbits.assign(rows * L.S, 0);
    int64_t id = 0;
    for (int64_t r = 0; r < rows; r++) {
        for (int s = 0; s < L.S; s++) {
            uint64_t w = 0;
            for (int b = 0; b < 64; b++) {
                int x = 64 * s + b;
                if (x < L.nx && used[r * L.nx + x]) w |= 1ull << b;
            }
            L.nbits[r * L.S + s] = w;
            L.nbase[r * L.S + s] = (int32_t)id;
            id += __builtin_popcountll(w);
        }
    }
    L.nbase[rows * L.S] = (int32_t)id;
    return id;
}

void pack_elements(HostLevel &L, const std::vector<uint8_t> &present)
{
    L.SE = (L.ex + 63) / 64;
    int64_t rows = (int64_t)L.ey * L.ez;
    L.ebits.assign(rows * L.SE, 0);
    for (int64_t r = 0; r < rows; r++)
        for (int x = 0; x < L.ex; x++)
            if (present[r * L.ex + x]) L.ebits[r * L.SE + x / 64] |= 1ull << (x % 64);
}

// dense grid -> compact id (or -1)
std::vector<int32_t> node_ids(const HostLevel &L, const std::vector<uint8_t> &used)
{
    std::vector<int32_t> id(used.size(), -1);
    int32_t k = 0;
    for (size_t i = 0; i < used.size(); i++)
        if (used[i]) id[i] = k++;
    return id;
}

std::vector<uint8_t> used_from_elements(const HostLevel &L, const std::vector<uint8_t> &present)
{
    std::vector<uint8_t> used((size_t)L.nx * L.ny * L.nz, 0);
    for (int z = 0; z < L.ez; z++)
        for (int y = 0; y < L.ey; y++)
            for (int x = 0; x < L.ex; x++) {
                if (!present[lin(x, y, z, L.ex, L.ey)]) continue;
                for (int c = 0; c < 8; c++)
                    used[lin(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1), L.nx, L.ny)] = 1;
            }
    return used;
}

// a[0..m) . b[0..m) with four partial sums
inline double dot4(const double *a, const double *b, int64_t m)
{
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int64_t k = 0;
    for (; k + 4 <= m; k += 4) {
        s0 += a[k] * b[k];
        s1 += a[k + 1] * b[k + 1];
        s2 += a[k + 2] * b[k + 2];
        s3 += a[k + 3] * b[k + 3];
    }
    for (; k < m; k++) s0 += a[k] * b[k];
    return (s0 + s1) + (s2 + s3);
}

// Cholesky (lower, row-major) in place; returns false if not positive definite.
// Blocked right-looking form: each 64-column block is factored, the rows below
// it are solved against it, and the trailing lower triangle is updated with
// dot products of contiguous row pieces; the last two steps run on threads.
// Only the lower triangle is read or written.
bool cholesky(std::vector<double> &A, int64_t n)
{
    constexpr int64_t NB = 64;
    for (int64_t kb = 0; kb < n; kb += NB) {
        const int64_t ke = std::min(n, kb + NB);
        // the diagonal block holds A minus the products of the earlier blocks
        for (int64_t i = kb; i < ke; i++) {
            double *Ai = &A[i * n];
            for (int64_t j = kb; j <= i; j++) {
                const double *Aj = &A[j * n];
                const double s = Ai[j] - dot4(Ai + kb, Aj + kb, j - kb);
                if (j == i) {
                    if (!(s > 0)) return false;
                    Ai[i] = std::sqrt(s);
                } else {
                    Ai[j] = s / Aj[j];
                }
            }
        }
        if (ke == n) break;
        parallel_for(n - ke, [&](int64_t r) {
            double *Ai = &A[(ke + r) * n];
            for (int64_t j = kb; j < ke; j++) {
                const double *Aj = &A[j * n];
                Ai[j] = (Ai[j] - dot4(Ai + kb, Aj + kb, j - kb)) / Aj[j];
            }
        });
        parallel_for(n - ke, [&](int64_t r) {
            const int64_t i = n - 1 - r;  // longest rows first
            double *Ai = &A[i * n];
            for (int64_t j = ke; j <= i; j++) Ai[j] -= dot4(Ai + kb, &A[j * n] + kb, ke - kb);
        });
    }
    return true;
}

// Inverse of A = L L^T from the factor, as W^T W with W = L^-1.  Column j of W
// (nonzero from row j on) is stored as row j of WT, so both the triangular
// solves and the product are dot products of contiguous row pieces:
// (A^-1)_ij = sum over k >= max(i, j) of W_ki W_kj.
std::vector<double> spd_inverse(std::vector<double> Lf, int64_t n, bool &ok)
{
    std::vector<double> inv;
    ok = cholesky(Lf, n);
    if (!ok) return inv;
    std::vector<double> WT((size_t)n * n, 0.0);
    parallel_for(n, [&](int64_t j) {
        double *w = &WT[j * n];
        w[j] = 1.0 / Lf[j * n + j];
        for (int64_t i = j + 1; i < n; i++) {
            const double *Li = &Lf[i * n];
            w[i] = -dot4(Li + j, w + j, i - j) / Li[i];
        }
    });
    inv.assign((size_t)n * n, 0.0);
    parallel_for(n, [&](int64_t i) {
        const double *wi = &WT[i * n];
        for (int64_t j = i; j < n; j++) {
            const double v = dot4(wi + j, &WT[j * n] + j, n - j);
            inv[i * n + j] = v;
            inv[j * n + i] = v;
        }
    });
    return inv;
}
}  // namespace

bool cholesky_factor(std::vector<double> &A, int64_t n) { return cholesky(A, n); }

void cholesky_solve(const std::vector<double> &Lf, int64_t n, double *b)
{
    for (int64_t i = 0; i < n; i++) {
        double v = b[i];
        const double *Li = &Lf[i * n];
        for (int64_t k = 0; k < i; k++) v -= Li[k] * b[k];
        b[i] = v / Li[i];
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        double v = b[i];
        for (int64_t k = i + 1; k < n; k++) v -= Lf[k * n + i] * b[k];
        b[i] = v / Lf[i * n + i];
    }
}

int node_id(const HostLevel &L, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= L.nx || y >= L.ny || z >= L.nz) return -1;
    const size_t seg = ((size_t)z * L.ny + y) * L.S + (x >> 6);
    const uint64_t w = L.nbits[seg];
    if (!((w >> (x & 63)) & 1)) return -1;
    return L.nbase[seg] + __builtin_popcountll(w & ((1ull << (x & 63)) - 1));
}

bool node_fixed(const HostLevel &L, int x, int y, int z)
{
    const size_t seg = ((size_t)z * L.ny + y) * L.S + (x >> 6);
    return (L.fbits[seg] >> (x & 63)) & 1;
}

int assemble_level_dense(const Hierarchy &H, int l, std::vector<double> &A, std::vector<int32_t> &freedof)
{
    if (l < 1 || l >= H.nlev || H.ktab[l].empty()) return -1;
    const HostLevel &L = H.L[l];
    const int64_t n3 = 3 * L.nn;
    std::vector<int32_t> map(n3, -1);
    freedof.clear();
    for (int z = 0; z < L.nz; z++)
        for (int y = 0; y < L.ny; y++)
            for (int x = 0; x < L.nx; x++) {
                const int id = node_id(L, x, y, z);
                if (id < 0 || node_fixed(L, x, y, z)) continue;
                for (int c = 0; c < 3; c++) {
                    map[3 * id + c] = (int32_t)freedof.size();
                    freedof.push_back(3 * id + c);
                }
            }
    const int64_t nf = (int64_t)freedof.size();
    A.assign((size_t)nf * nf, 0.0);
    const std::vector<double> &T = H.ktab[l];
    for (int ez = 0; ez < L.ez; ez++)
        for (int ey = 0; ey < L.ey; ey++)
            for (int ex = 0; ex < L.ex; ex++) {
                const int p = L.epat[((size_t)ez * L.ey + ey) * L.ex + ex];
                if (!p) continue;
                int dof[24];
                for (int n = 0; n < 8; n++) {
                    const int id = node_id(L, ex + (n & 1), ey + ((n >> 1) & 1), ez + (n >> 2));
                    for (int c = 0; c < 3; c++) dof[3 * n + c] = id < 0 ? -1 : map[3 * id + c];
                }
                const double *K = &T[(size_t)p * 576];
                for (int i = 0; i < 24; i++) {
                    if (dof[i] < 0) continue;
                    for (int j = 0; j < 24; j++)
                        if (dof[j] >= 0) A[(size_t)dof[i] * nf + dof[j]] += K[i * 24 + j];
                }
            }
    return 0;
}

void galerkin_element(const double *ke, int div, uint64_t mask, double *K)
{
    int nch = div * div * div;
    std::memset(K, 0, sizeof(double) * 576);
    double P[8][8];
    double KP[24][24];
    for (int c = 0; c < nch; c++) {
        if (!((mask >> c) & 1)) continue;
        for (int L = 0; L < 8; L++)
            for (int k = 0; k < 8; k++) P[L][k] = child_weight(div, c, L, k);
        // KP = Ke * P_c   (P_c acts on each component separately)
        for (int r = 0; r < 24; r++)
            for (int k = 0; k < 8; k++)
                for (int j = 0; j < 3; j++) {
                    double s = 0;
                    for (int Lp = 0; Lp < 8; Lp++) s += ke[r * 24 + 3 * Lp + j] * P[Lp][k];
                    KP[r][3 * k + j] = s;
                }
        // K += P_c^T KP
        for (int k = 0; k < 8; k++)
            for (int i = 0; i < 3; i++)
                for (int col = 0; col < 24; col++) {
                    double s = 0;
                    for (int L = 0; L < 8; L++) s += P[L][k] * KP[3 * L + i][col];
                    K[(3 * k + i) * 24 + col] += s;
                }
    }
}

int check_hierarchy_arguments(const int32_t *voxels, int vx, int vy, int vz, const double *ke_in, const int64_t *fixed,
                              int64_t nfixed, int num_levels, int skip, std::string &err)
{
    if (!voxels || !ke_in || (nfixed > 0 && !fixed) || vx <= 0 || vy <= 0 || vz <= 0 || num_levels < 1 ||
        num_levels > kMaxLevels || skip < 0 || skip > 1) {
        err = "invalid argument to vfem_assemble";
        return 1;
    }
    if (vx + 1 > 1023 || vy + 1 > 1023 || vz + 1 > 1023) {
        err = "grid extents above 1022 voxels are not supported";
        return 1;
    }
    return 0;
}

void symmetrise_ke(const double *ke_in, double *ke)
{
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) ke[i * 24 + j] = (i <= j) ? ke_in[i * 24 + j] : ke_in[j * 24 + i];
}

void level_dims(const HostLevel &F, HostLevel &C, int l, int skip)
{
    C.rdx = F.rdx / 2 + F.rdx % 2; C.rdy = F.rdy / 2 + F.rdy % 2; C.rdz = F.rdz / 2 + F.rdz % 2;
    C.div = (l == 1 && skip > 0) ? (1 << (skip + 1)) : 2;
    C.ex = (F.ex - 1) / C.div + 1; C.ey = (F.ey - 1) / C.div + 1; C.ez = (F.ez - 1) / C.div + 1;
    C.nx = C.ex + 1; C.ny = C.ey + 1; C.nz = C.ez + 1;
    C.fullscale = std::ldexp(1.0, l + skip);
    C.partscale = std::ldexp(1.0, l - 1);
}

void coarsest_matrices(Hierarchy &H, const std::vector<uint64_t> *cmask, std::vector<double> &kref,
                       std::vector<double> &kgal)
{
    const int nlev = H.nlev, skip = H.skip;
    const HostLevel &Cc = H.L[nlev - 1];
    std::vector<int32_t> freeidx(Cc.nn, -1);
    {
        int32_t nf = 0;
        int64_t k = 0;
        for (int64_t r = 0; r < (int64_t)Cc.ny * Cc.nz; r++)
            for (int x = 0; x < Cc.nx; x++) {
                uint64_t w = Cc.nbits[r * Cc.S + x / 64];
                if (!((w >> (x % 64)) & 1)) continue;
                bool fx = (Cc.fbits[r * Cc.S + x / 64] >> (x % 64)) & 1;
                freeidx[k++] = fx ? -1 : nf++;
            }
        H.nfree = 3 * (int64_t)nf;
    }
    H.freedof.resize(H.nfree);
    for (int64_t k = 0; k < Cc.nn; k++)
        if (freeidx[k] >= 0)
            for (int c = 0; c < 3; c++) H.freedof[3 * freeidx[k] + c] = (int32_t)(3 * k + c);
    const int64_t n = H.nfree;
    kref.assign((size_t)n * n, 0.0);
    kgal.assign((size_t)n * n, 0.0);
    if (n == 0) return;
    const double kscale = std::ldexp(1.0, nlev + skip - 1);
    const uint64_t fullmask = (Cc.div == 4) ? ~0ull : 0xFFull;
    std::vector<double> G(576);
    for (int z = 0; z < Cc.ez; z++)
        for (int y = 0; y < Cc.ey; y++)
            for (int x = 0; x < Cc.ex; x++) {
                if (!((Cc.ebits[((size_t)z * Cc.ey + y) * Cc.SE + (x >> 6)] >> (x & 63)) & 1)) continue;
                const int64_t e = lin(x, y, z, Cc.ex, Cc.ey);
                int32_t fn[8];
                for (int L = 0; L < 8; L++)
                    fn[L] = freeidx[node_id(Cc, x + (L & 1), y + ((L >> 1) & 1), z + ((L >> 2) & 1))];
                const bool partial = cmask && (*cmask)[e] != fullmask;
                if (partial) {
                    galerkin_element(H.ke, Cc.div, (*cmask)[e], G.data());
                    for (int q = 0; q < 576; q++) G[q] *= Cc.partscale;
                } else {
                    for (int q = 0; q < 576; q++) G[q] = Cc.fullscale * H.ke[q];
                }
                for (int a = 0; a < 8; a++) {
                    if (fn[a] < 0) continue;
                    for (int b = 0; b < 8; b++) {
                        if (fn[b] < 0) continue;
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) {
                                size_t idx = (size_t)(3 * fn[a] + i) * n + 3 * fn[b] + j;
                                kref[idx] += kscale * H.ke[(3 * a + i) * 24 + 3 * b + j];
                                kgal[idx] += G[(3 * a + i) * 24 + 3 * b + j];
                            }
                    }
                }
            }
}

int build_hierarchy(const int32_t *voxels, int vx, int vy, int vz, const double *ke_in, const int64_t *fixed,
                    int64_t nfixed, int num_levels, int skip, Hierarchy &H, std::string &err)
{
    if (int rc = check_hierarchy_arguments(voxels, vx, vy, vz, ke_in, fixed, nfixed, num_levels, skip, err)) return rc;
    H = Hierarchy();
    H.skip = skip;
    symmetrise_ke(ke_in, H.ke);

    // ---- level 0 ----------------------------------------------------------
    H.L.emplace_back();
    HostLevel *F = &H.L[0];
    F->ex = vx; F->ey = vy; F->ez = vz;
    F->nx = vx + 1; F->ny = vy + 1; F->nz = vz + 1;
    F->rdx = vx; F->rdy = vy; F->rdz = vz;
    F->div = 1;
    F->fullscale = 1.0;
    std::vector<uint8_t> present((size_t)vx * vy * vz);
    int64_t ne = 0;
    for (size_t i = 0; i < present.size(); i++) {
        present[i] = voxels[i] == 1;
        ne += present[i];
    }
    if (ne == 0) {
        err = "the voxel model has no solid voxel";
        return 2;
    }
    F->ne = ne;
    pack_elements(*F, present);
    std::vector<uint8_t> used = used_from_elements(*F, present);
    F->nn = pack_nodes(*F, used);
    // the device kernels pack a node id with two flag bits into an int32
    // (id << 2 | used << 1 | fixed) and index DOFs as 3 * id in 32 bits
    if (F->nn >= kMaxNodes) {
        err = "more than 2^29 - 1 used nodes on level 0 are not supported";
        return 1;
    }
    std::vector<int32_t> ids = node_ids(*F, used);
    // fixed compact ids -> grid bits
    {
        std::vector<int64_t> coord(F->nn);
        for (size_t i = 0; i < ids.size(); i++)
            if (ids[i] >= 0) coord[ids[i]] = (int64_t)i;
        for (int64_t k = 0; k < nfixed; k++) {
            if (fixed[k] < 0 || fixed[k] >= F->nn) {
                err = "fixed node id out of range";
                return 1;
            }
            int64_t g = coord[fixed[k]];
            int x = (int)(g % F->nx);
            int64_t r = g / F->nx;
            uint64_t &w = F->fbits[r * F->S + x / 64];
            if (!((w >> (x % 64)) & 1)) F->nfixed++;
            w |= 1ull << (x % 64);
        }
    }
    std::vector<uint8_t> cur_fixed(used.size(), 0);
    for (int64_t r = 0; r < (int64_t)F->ny * F->nz; r++)
        for (int x = 0; x < F->nx; x++) cur_fixed[r * F->nx + x] = (F->fbits[r * F->S + x / 64] >> (x % 64)) & 1;

    // inverse diagonal with per-element diagonals, node by node: a used node sums
    // the diagonals of its present elements in ascending element order (the
    // order an element-by-element scatter adds them), so every node whose eight
    // elements are full gets the same value bit for bit
    auto inverse_diagonal = [&](HostLevel &L, const std::vector<uint8_t> &pres, const std::vector<uint64_t> *masks,
                                const std::vector<int32_t> &nid) {
        const uint64_t fullmask = (L.div == 4) ? ~0ull : 0xFFull;
        std::array<double, 24> fulld;
        for (int q = 0; q < 24; q++) fulld[q] = L.fullscale * H.ke[q * 24 + q];
        // diagonals of the partial elements' Galerkin matrices, one per child mask
        std::unordered_map<uint64_t, std::array<double, 24>> partial;
        if (masks) {
            std::vector<uint64_t> todo;
            for (size_t e = 0; e < pres.size(); e++) {
                const uint64_t m = (*masks)[e];
                if (pres[e] && m != fullmask && partial.emplace(m, std::array<double, 24>{}).second) todo.push_back(m);
            }
            // the map is not modified below: threads look up and fill distinct entries
            parallel_for((int64_t)todo.size(), [&](int64_t t) {
                std::vector<double> K(576);
                galerkin_element(H.ke, L.div, todo[t], K.data());
                std::array<double, 24> &d = partial.find(todo[t])->second;
                for (int q = 0; q < 24; q++) d[q] = L.partscale * K[q * 24 + q];
            });
        }
        L.invdiag.assign(3 * L.nn, 0.0);
        parallel_for(L.nz, [&](int64_t z) {
            for (int y = 0; y < L.ny; y++)
                for (int x = 0; x < L.nx; x++) {
                    const int32_t id = nid[lin(x, y, (int)z, L.nx, L.ny)];
                    if (id < 0) continue;
                    double s[3] = {0.0, 0.0, 0.0};
                    for (int k = 0; k < 2; k++)
                        for (int j = 0; j < 2; j++)
                            for (int i = 0; i < 2; i++) {
                                const int ex = x - 1 + i, ey = y - 1 + j, ez = (int)z - 1 + k;
                                if (ex < 0 || ey < 0 || ez < 0 || ex >= L.ex || ey >= L.ey || ez >= L.ez) continue;
                                const int64_t e = lin(ex, ey, ez, L.ex, L.ey);
                                if (!pres[e]) continue;
                                const double *dv = fulld.data();
                                if (masks && (*masks)[e] != fullmask) dv = partial.find((*masks)[e])->second.data();
                                const int ln = (1 - i) + 2 * (1 - j) + 4 * (1 - k);
                                for (int c = 0; c < 3; c++) s[c] += dv[3 * ln + c];
                            }
                    for (int c = 0; c < 3; c++) L.invdiag[3 * (int64_t)id + c] = 1.0 / s[c];
                }
        });
    };
    inverse_diagonal(*F, present, nullptr, ids);

    // ---- coarse levels ----------------------------------------------------
    int nlev = num_levels;
    std::vector<uint64_t> cmask;
    for (int l = 1; l < nlev; l++) {
        H.L.emplace_back();
        F = &H.L[l - 1];
        HostLevel &C = H.L[l];
        level_dims(*F, C, l, skip);
        const int div = C.div;
        cmask.assign((size_t)C.ex * C.ey * C.ez, 0);
        for (int z = 0; z < F->ez; z++)
            for (int y = 0; y < F->ey; y++)
                for (int x = 0; x < F->ex; x++) {
                    if (!present[lin(x, y, z, F->ex, F->ey)]) continue;
                    int cx = x / div, cy = y / div, cz = z / div;
                    int ch = (x - cx * div) + div * (y - cy * div) + div * div * (z - cz * div);
                    cmask[lin(cx, cy, cz, C.ex, C.ey)] |= 1ull << ch;
                }
        std::vector<uint8_t> cpres(cmask.size());
        C.ne = 0;
        for (size_t i = 0; i < cmask.size(); i++) {
            cpres[i] = cmask[i] != 0;
            C.ne += cpres[i];
        }
        if (div == 2) {
            C.epat.resize(cmask.size());
            for (size_t i = 0; i < cmask.size(); i++) C.epat[i] = (uint8_t)cmask[i];
        }
        pack_elements(C, cpres);
        std::vector<uint8_t> cused = used_from_elements(C, cpres);
        C.nn = pack_nodes(C, cused);
        std::vector<int32_t> cids = node_ids(C, cused);
        // fixed propagation
        std::vector<uint8_t> cfix(cused.size(), 0);
        for (int z = 0; z < F->nz; z++)
            for (int y = 0; y < F->ny; y++)
                for (int x = 0; x < F->nx; x++) {
                    if (!cur_fixed[lin(x, y, z, F->nx, F->ny)]) continue;
                    int lo[3], hi[3];
                    int f[3] = {x, y, z};
                    for (int d = 0; d < 3; d++) {
                        lo[d] = f[d] / div;
                        hi[d] = (f[d] % div) ? lo[d] + 1 : lo[d];
                    }
                    for (int Z = lo[2]; Z <= hi[2]; Z++)
                        for (int Y = lo[1]; Y <= hi[1]; Y++)
                            for (int X = lo[0]; X <= hi[0]; X++) {
                                if (X >= C.nx || Y >= C.ny || Z >= C.nz) continue;
                                int64_t g = lin(X, Y, Z, C.nx, C.ny);
                                if (cused[g]) cfix[g] = 1;
                            }
                }
        for (int64_t r = 0; r < (int64_t)C.ny * C.nz; r++)
            for (int x = 0; x < C.nx; x++)
                if (cfix[r * C.nx + x]) {
                    C.fbits[r * C.S + x / 64] |= 1ull << (x % 64);
                    C.nfixed++;
                }
        inverse_diagonal(C, cpres, &cmask, cids);
        // operator table of a div-2 level (patterns 0..255)
        if (div == 2) {
            H.ktab.resize(l + 1);
            std::vector<double> &T = H.ktab[l];
            T.assign(256 * 576, 0.0);
            for (int p = 1; p < 256; p++) {
                if (p == 255) {
                    for (int q = 0; q < 576; q++) T[255 * 576 + q] = C.fullscale * H.ke[q];
                } else {
                    galerkin_element(H.ke, 2, (uint64_t)p, &T[(size_t)p * 576]);
                    for (int q = 0; q < 576; q++) T[(size_t)p * 576 + q] *= C.partscale;
                }
            }
        }
        present.swap(cpres);
        cur_fixed.swap(cfix);
        ids.swap(cids);
        if (C.rdx <= 2 || C.rdy <= 2 || C.rdz <= 2) {  // FEM.cpp:465-470
            nlev = l + 1;
            break;
        }
    }
    H.nlev = nlev;
    H.ktab.resize(nlev);

    // ---- coarsest level: free DOFs, Kc (FEM.cpp:473-513) --------------------
    std::vector<double> kref, kgal;
    coarsest_matrices(H, nlev > 1 ? &cmask : nullptr, kref, kgal);
    const int64_t n = H.nfree;
    if (n > 0) {
        bool ok1, ok2;
        H.kcinv_ref = spd_inverse(kref, n, ok1);
        H.kcinv_gal = spd_inverse(kgal, n, ok2);
        if (!ok1 || !ok2) {
            err = "the coarsest-level matrix is not positive definite (floating component or no fixed node?)";
            return 4;
        }
    }
    return 0;
}

}  // namespace vfem
