/*
 * vfem_oracle.c -- CPU FP64 restatement of the reference solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library's results are compared with it: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may build, load
 * or call it.  The product library (vulkanfem_amd/csrc) never links or calls it.
 *
 * Reference = hirademirkol/VulkanFEM.  Each function below cites the reference
 * file:line it restates.  The reference itself cannot be built here (its Eigen
 * and Kompute submodules are empty), so this restatement is pinned instead by
 * closed forms, invariants and independently assembled scipy goldens under
 * tests/golden (see tests/test_oracle.py).
 *
 * Third-party arithmetic restated from published algorithms:
 *   - Eigen 3.4.0 ConjugateGradient (FEM.cpp:607): preconditioned CG, Saad,
 *     "Iterative Methods for Sparse Linear Systems", Alg. 9.1, with Eigen's
 *     stopping rule |r|^2 < max(tol^2 |b|^2, DBL_MIN) tested right after the
 *     residual update and the iteration counter excluding the converged step.
 *   - Eigen 3.4.0 SimplicialLDLT on Kc (MultigridPreconditioner.hpp:45,177):
 *     replaced by a dense Cholesky factorisation, which yields the same solution
 *     up to rounding (Golub & Van Loan, Alg. 4.2.1).
 *
 * Deliberate deviation (reference quirk): FEM.cpp:350 linearises a fine lattice
 * node with the fine level's reference dims, so on a level with an odd element
 * extent a lattice node one past the grid edge aliases to (0, y+1, z).  Here a
 * lattice node outside the fine node grid is treated as unused (-1).  No grid
 * whose level extents are all even is affected.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VO_MAXL 16

typedef struct {
    int ex, ey, ez;        /* element extents of this level (coarse coordinates) */
    int rdx, rdy, rdz;     /* reference levelDims (FEM.cpp:256-259), used for the stop rule */
    int64_t ne, nn;        /* number of elements / used nodes */
    int32_t *ecoord;       /* ne*3, elements in lexicographic (z,y,x) order */
    int32_t *e2n;          /* ne*4, compact id of the x-low corners, column y+2z (FEM.cpp:193-197) */
    int32_t *nodeid;       /* (ex+1)(ey+1)(ez+1) dense map grid node -> compact id or -1 */
    /* transfer from level l-1 into this level (l >= 1) */
    int div, nb;           /* coarsening factor (2, or 4 for the skipped level) and lattice width div+1 */
    int32_t *rmap;         /* ne * nb^3 : restrictionMapping (FEM.cpp:348-362) */
    uint64_t *childmask;   /* ne : which children exist, bit = Linearize(child, div) */
    int32_t *ukidx;        /* ne : index into uk for partially filled elements, -1 if full */
    double *uk;            /* unique element matrices, 576 each (FEM.cpp:364-395) */
    int nuk;
    double fullscale;      /* multiplier of Ke on full elements for invDiag (FEM.cpp:463) */
    /* data that lives on this level's nodes */
    double *rcoef;         /* 3*nn, 1/multiplicity towards level l+1 (FEM.cpp:354-356,400) */
    double *invdiag;       /* 3*nn (FEM.cpp:104-151) */
    uint8_t *fixed;        /* nn, 1 = fixed node */
    int64_t nfixed;
} vo_level;

typedef struct {
    int nlev, skip;
    int vx, vy, vz;
    double ke[576];        /* symmetrised from the upper triangle (FEM.cpp:200-210) */
    vo_level L[VO_MAXL];
    /* coarsest-level direct solve: reference Kc (FEM.cpp:473-513) */
    int64_t nfree;
    int32_t *freedof;      /* nfree : coarse DOF index of each free DOF */
    double *kc;            /* nfree*nfree dense, lower Cholesky factor after factorisation */
    /* coarsest-level direct solve for the damped multiplicative cycle: Kc assembled
       from the same per-element matrices as the coarse operator */
    double *kcg;
    /* FP32-rounded copies of rcoef / invdiag (the reference GPU path stores them as
       float, FEMGPU.cpp:226-242) */
    double *rcoef32[VO_MAXL];
    double *invdiag32[VO_MAXL];
    /* Chebyshev smoothing of the damped cycle on levels >= cheb_level (0 = off):
       degree cheb_degree, fitted to [lmax/ratio, lmax] of D^-1 A per level */
    int cheb_level, cheb_degree;
    double cheb_ratio;
    double cheb_lmax[VO_MAXL];
} vo_sys;

/* ------------------------------------------------------------------------- */
/* closed forms                                                              */
/* ------------------------------------------------------------------------- */

/* Trilinear hex8 stiffness on the unit cube, isotropic E, nu, 2x2x2 Gauss,
 * nodes lexicographic x-fastest, then every node's DOF triple rotated by
 * R = [[0,1,0],[-1,0,0],[0,0,1]] so that the result is in the DOF frame of
 * data/Ke0.dat (scripts/outputToC.m:13-14 documents the node reorder). */
void vo_ke0_closed_form(double E, double nu, double *out)
{
    double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    double mu = E / (2 * (1 + nu));
    double D[6][6] = {{0}};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) D[i][j] = lam;
        D[i][i] = lam + 2 * mu;
        D[3 + i][3 + i] = mu;
    }
    double K[24][24];
    memset(K, 0, sizeof(K));
    double g = 1.0 / sqrt(3.0);
    double gp[2] = {0.5 - 0.5 * g, 0.5 + 0.5 * g};
    for (int qz = 0; qz < 2; qz++)
        for (int qy = 0; qy < 2; qy++)
            for (int qx = 0; qx < 2; qx++) {
                double px = gp[qx], py = gp[qy], pz = gp[qz];
                double dN[8][3];
                for (int n = 0; n < 8; n++) {
                    int a = n & 1, b = (n >> 1) & 1, c = (n >> 2) & 1;
                    double fx = a ? px : 1 - px, fy = b ? py : 1 - py, fz = c ? pz : 1 - pz;
                    double sx = a ? 1 : -1, sy = b ? 1 : -1, sz = c ? 1 : -1;
                    dN[n][0] = sx * fy * fz;
                    dN[n][1] = fx * sy * fz;
                    dN[n][2] = fx * fy * sz;
                }
                double B[6][24];
                memset(B, 0, sizeof(B));
                for (int n = 0; n < 8; n++) {
                    B[0][3 * n + 0] = dN[n][0];
                    B[1][3 * n + 1] = dN[n][1];
                    B[2][3 * n + 2] = dN[n][2];
                    B[3][3 * n + 0] = dN[n][1];
                    B[3][3 * n + 1] = dN[n][0];
                    B[4][3 * n + 1] = dN[n][2];
                    B[4][3 * n + 2] = dN[n][1];
                    B[5][3 * n + 0] = dN[n][2];
                    B[5][3 * n + 2] = dN[n][0];
                }
                double DB[6][24];
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 24; j++) {
                        double s = 0;
                        for (int k = 0; k < 6; k++) s += D[i][k] * B[k][j];
                        DB[i][j] = s;
                    }
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++) {
                        double s = 0;
                        for (int k = 0; k < 6; k++) s += B[k][i] * DB[k][j];
                        K[i][j] += 0.125 * s; /* weight 1/8 per point on the unit cube */
                    }
            }
    /* rotate each node's DOF triple: u' = R u */
    static const double R[3][3] = {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}};
    for (int a = 0; a < 8; a++)
        for (int b = 0; b < 8; b++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0;
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) s += R[i][k] * K[3 * a + k][3 * b + l] * R[j][l];
                    out[(3 * a + i) * 24 + 3 * b + j] = s;
                }
}

/* Trilinear transfer weights, closed form.  div = 2 gives the 27x8 table of
 * RestrictionOperator.hpp:86-114, div = 4 the 125x8 table (:117-244):
 * W[m][k] = prod_d (1 - |xi_d(m) - k_d|), xi(m) = lattice position / div. */
void vo_transfer_weights(int div, double *W)
{
    int nb = div + 1;
    for (int c = 0; c < nb; c++)
        for (int b = 0; b < nb; b++)
            for (int a = 0; a < nb; a++) {
                int m = a + nb * b + nb * nb * c;
                for (int k = 0; k < 8; k++) {
                    double kx = k & 1, ky = (k >> 1) & 1, kz = (k >> 2) & 1;
                    W[m * 8 + k] = (1 - fabs((double)a / div - kx)) * (1 - fabs((double)b / div - ky)) *
                                   (1 - fabs((double)c / div - kz));
                }
            }
}

/* Child element -> lattice index of its local node (RestrictionOperator.hpp:6-84). */
void vo_child_lattice_map(int div, int32_t *out)
{
    int nb = div + 1;
    for (int cz = 0; cz < div; cz++)
        for (int cy = 0; cy < div; cy++)
            for (int cx = 0; cx < div; cx++) {
                int child = cx + div * cy + div * div * cz;
                for (int L = 0; L < 8; L++) {
                    int lx = L & 1, ly = (L >> 1) & 1, lz = (L >> 2) & 1;
                    out[child * 8 + L] = (cx + lx) + nb * (cy + ly) + nb * nb * (cz + lz);
                }
            }
}

/* ------------------------------------------------------------------------- */
/* helpers                                                                   */
/* ------------------------------------------------------------------------- */

static inline int64_t lin3(int x, int y, int z, int nx, int ny) { return (int64_t)x + (int64_t)nx * (y + (int64_t)ny * z); }

/* local DOF d of element e -> global DOF (MatrixFreeSparse.hpp:200-206) */
static inline int64_t gdof(const int32_t *e2n, int64_t e, int d) { return 3 * (int64_t)e2n[4 * e + d / 6] + d % 6; }

static void free_level(vo_level *L)
{
    free(L->ecoord); free(L->e2n); free(L->nodeid); free(L->rmap); free(L->childmask);
    free(L->ukidx); free(L->uk); free(L->rcoef); free(L->invdiag); free(L->fixed);
    memset(L, 0, sizeof(*L));
}

void vo_free(vo_sys *s)
{
    if (!s) return;
    for (int l = 0; l < VO_MAXL; l++) { free_level(&s->L[l]); free(s->rcoef32[l]); free(s->invdiag32[l]); }
    free(s->freedof); free(s->kc); free(s->kcg);
    free(s);
}

/* number nodes of a level: mark corners of every element, then rank in
 * lexicographic (z,y,x) order -- the std::map ordering of FEM.cpp:88-99,179-184
 * (level 0) and :298-319 (coarse levels). */
static int number_nodes(vo_level *L)
{
    int nx = L->ex + 1, ny = L->ey + 1, nz = L->ez + 1;
    int64_t ng = (int64_t)nx * ny * nz;
    L->nodeid = (int32_t *)malloc(sizeof(int32_t) * ng);
    if (!L->nodeid) return -1;
    for (int64_t i = 0; i < ng; i++) L->nodeid[i] = -1;
    for (int64_t e = 0; e < L->ne; e++) {
        int x = L->ecoord[3 * e], y = L->ecoord[3 * e + 1], z = L->ecoord[3 * e + 2];
        for (int c = 0; c < 8; c++) L->nodeid[lin3(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1), nx, ny)] = 0;
    }
    int64_t id = 0;
    for (int64_t i = 0; i < ng; i++)
        if (L->nodeid[i] == 0) L->nodeid[i] = (int32_t)id++;
    L->nn = id;
    L->e2n = (int32_t *)malloc(sizeof(int32_t) * 4 * (L->ne ? L->ne : 1));
    if (!L->e2n) return -1;
    for (int64_t e = 0; e < L->ne; e++) {
        int x = L->ecoord[3 * e], y = L->ecoord[3 * e + 1], z = L->ecoord[3 * e + 2];
        for (int zz = 0; zz < 2; zz++)
            for (int yy = 0; yy < 2; yy++) L->e2n[4 * e + yy + 2 * zz] = L->nodeid[lin3(x, y + yy, z + zz, nx, ny)];
    }
    return 0;
}

static inline int32_t node_at(const vo_level *L, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x > L->ex || y > L->ey || z > L->ez) return -1;
    return L->nodeid[lin3(x, y, z, L->ex + 1, L->ey + 1)];
}

/* ------------------------------------------------------------------------- */
/* restriction / interpolation (MultigridPreconditioner.hpp:67-155)           */
/* ------------------------------------------------------------------------- */

/* r_{l+1} = R_l (mu_l .* r_l); coarse = level l+1 (MultigridPreconditioner.hpp:67-108,
 * GPU twin CWiseMult.comp + Restrict.comp).  Coarse elements of one parity class
 * share no corner, so each class is scattered in parallel without races. */
static void restrict_level(const vo_sys *s, int l, const double *r, double *rc, const double *mu)
{
    const vo_level *F = &s->L[l], *C = &s->L[l + 1];
    int nb = C->nb, nl = nb * nb * nb;
    double W[125 * 8];
    vo_transfer_weights(C->div, W);
    memset(rc, 0, sizeof(double) * 3 * C->nn);
    for (int color = 0; color < 8; color++) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < C->ne; e++) {
            const int32_t *ec = C->ecoord + 3 * e;
            if (((ec[0] & 1) | ((ec[1] & 1) << 1) | ((ec[2] & 1) << 2)) != color) continue;
            const int32_t *map = C->rmap + (int64_t)nl * e;
            for (int k = 0; k < 8; k++) {
                int64_t node = C->e2n[4 * e + k / 2] + k % 2;
                for (int c = 0; c < 3; c++) {
                    double acc = 0;
                    for (int m = 0; m < nl; m++) {
                        if (map[m] < 0) continue;
                        acc += r[3 * (int64_t)map[m] + c] * mu[3 * (int64_t)map[m] + c] * W[m * 8 + k];
                    }
                    rc[3 * node + c] += acc;
                }
            }
        }
    }
    (void)F;
}

/* x_l = mu_l .* (sum_E W x_{l+1}) (MultigridPreconditioner.hpp:110-155). */
static void interpolate_level(const vo_sys *s, int l, const double *xc, double *xf, const double *mu)
{
    const vo_level *F = &s->L[l], *C = &s->L[l + 1];
    int nb = C->nb, nl = nb * nb * nb;
    double W[125 * 8];
    vo_transfer_weights(C->div, W);
    memset(xf, 0, sizeof(double) * 3 * F->nn);
    for (int color = 0; color < 8; color++) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < C->ne; e++) {
            const int32_t *ec = C->ecoord + 3 * e;
            if (((ec[0] & 1) | ((ec[1] & 1) << 1) | ((ec[2] & 1) << 2)) != color) continue;
            const int32_t *map = C->rmap + (int64_t)nl * e;
            double xv[8][3];
            for (int k = 0; k < 8; k++) {
                int64_t node = C->e2n[4 * e + k / 2] + k % 2;
                for (int c = 0; c < 3; c++) xv[k][c] = xc[3 * node + c];
            }
            for (int m = 0; m < nl; m++) {
                if (map[m] < 0) continue;
                for (int c = 0; c < 3; c++) {
                    double acc = 0;
                    for (int k = 0; k < 8; k++) acc += W[m * 8 + k] * xv[k][c];
                    xf[3 * (int64_t)map[m] + c] += acc;
                }
            }
        }
    }
    for (int64_t i = 0; i < 3 * F->nn; i++) xf[i] *= mu[i];
}

/* ------------------------------------------------------------------------- */
/* dense Cholesky (stands in for SimplicialLDLT)                              */
/* ------------------------------------------------------------------------- */

static int cholesky(double *A, int64_t n)
{
    for (int64_t j = 0; j < n; j++) {
        double d = A[j * n + j];
        for (int64_t k = 0; k < j; k++) d -= A[j * n + k] * A[j * n + k];
        if (!(d > 0)) return -1;
        d = sqrt(d);
        A[j * n + j] = d;
#pragma omp parallel for schedule(static)
        for (int64_t i = j + 1; i < n; i++) {
            double v = A[i * n + j];
            for (int64_t k = 0; k < j; k++) v -= A[i * n + k] * A[j * n + k];
            A[i * n + j] = v / d;
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) A[i * n + j] = 0;
    return 0;
}

static void chol_solve(const double *Lf, int64_t n, double *b)
{
    for (int64_t i = 0; i < n; i++) {
        double v = b[i];
        for (int64_t k = 0; k < i; k++) v -= Lf[i * n + k] * b[k];
        b[i] = v / Lf[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        double v = b[i];
        for (int64_t k = i + 1; k < n; k++) v -= Lf[k * n + i] * b[k];
        b[i] = v / Lf[i * n + i];
    }
}

/* ------------------------------------------------------------------------- */
/* setup: assembleSystemMatrix, MATRIX_FREE + MULTIGRID branch (FEM.cpp:157-517) */
/* ------------------------------------------------------------------------- */

/* unique element matrix of a partially filled coarse element
 * (FEM.cpp:364-395): K = R^T (sum_children scale*Ke placed on the lattice) R */
static void unique_k(const double *ke, int div, uint64_t mask, double scale, double *K)
{
    int nb = div + 1, nl = nb * nb * nb, nd = 3 * nl;
    double *T = (double *)calloc((size_t)nd * nd, sizeof(double));
    int32_t lat[64 * 8];
    vo_child_lattice_map(div, lat);
    double W[125 * 8];
    vo_transfer_weights(div, W);
    int nch = div * div * div;
    for (int ch = 0; ch < nch; ch++) {
        if (!((mask >> ch) & 1)) continue;
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        T[(size_t)(3 * lat[ch * 8 + a] + i) * nd + 3 * lat[ch * 8 + b] + j] +=
                            scale * ke[(3 * a + i) * 24 + 3 * b + j];
    }
    /* K[3k+c][3k'+c'] = sum_{m,m'} W[m][k] T[3m+c][3m'+c'] W[m'][k'] */
    double *TW = (double *)calloc((size_t)nd * 24, sizeof(double));
    for (int r = 0; r < nd; r++)
        for (int mp = 0; mp < nl; mp++)
            for (int cp = 0; cp < 3; cp++) {
                double t = T[(size_t)r * nd + 3 * mp + cp];
                if (t == 0) continue;
                for (int kp = 0; kp < 8; kp++) TW[(size_t)r * 24 + 3 * kp + cp] += t * W[mp * 8 + kp];
            }
    for (int k = 0; k < 8; k++)
        for (int c = 0; c < 3; c++)
            for (int col = 0; col < 24; col++) {
                double acc = 0;
                for (int m = 0; m < nl; m++) acc += W[m * 8 + k] * TW[(size_t)(3 * m + c) * 24 + col];
                K[(3 * k + c) * 24 + col] = acc;
            }
    free(T);
    free(TW);
}

/* invDiag from per-element matrices (FEM.cpp:104-151) */
static void inverse_diagonal(const vo_level *L, const double *ke, double fullscale, double *invdiag)
{
    memset(invdiag, 0, sizeof(double) * 3 * L->nn);
    for (int64_t e = 0; e < L->ne; e++) {
        const double *K = (L->ukidx && L->ukidx[e] >= 0) ? L->uk + 576 * (int64_t)L->ukidx[e] : NULL;
        for (int d = 0; d < 24; d++) {
            double v = K ? K[d * 24 + d] : fullscale * ke[d * 24 + d];
            invdiag[gdof(L->e2n, e, d)] += v;
        }
    }
    for (int64_t i = 0; i < 3 * L->nn; i++) invdiag[i] = 1.0 / invdiag[i];
}

typedef struct { uint64_t key; int32_t val; } hslot;
static inline uint64_t hslot_index(uint64_t m, int bits) { return (m * 0x9E3779B97F4A7C15ull) >> (64 - bits); }

int vo_build(const int32_t *vox, int vx, int vy, int vz, const double *ke_in, const int64_t *fixed, int64_t nfixed,
             int num_levels, int skip, vo_sys **out)
{
    *out = NULL;
    if (vx <= 0 || vy <= 0 || vz <= 0 || num_levels < 1 || num_levels > VO_MAXL || skip < 0 || skip > 1) return -1;
    vo_sys *s = (vo_sys *)calloc(1, sizeof(vo_sys));
    s->vx = vx; s->vy = vy; s->vz = vz; s->skip = skip;
    /* Ke symmetrised from the upper triangle (FEM.cpp:200-210) */
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) s->ke[i * 24 + j] = (i <= j) ? ke_in[i * 24 + j] : ke_in[j * 24 + i];

    /* level 0: EnlistUsedElements (FEM.cpp:65-100), elements in FOR3 (z,y,x) order */
    vo_level *L0 = &s->L[0];
    L0->ex = vx; L0->ey = vy; L0->ez = vz;
    L0->rdx = vx; L0->rdy = vy; L0->rdz = vz;
    int64_t ne = 0;
    for (int64_t i = 0; i < (int64_t)vx * vy * vz; i++) ne += (vox[i] == 1);
    L0->ne = ne;
    L0->ecoord = (int32_t *)malloc(sizeof(int32_t) * 3 * (ne ? ne : 1));
    ne = 0;
    for (int z = 0; z < vz; z++)
        for (int y = 0; y < vy; y++)
            for (int x = 0; x < vx; x++)
                if (vox[lin3(x, y, z, vx, vy)] == 1) {
                    L0->ecoord[3 * ne] = x; L0->ecoord[3 * ne + 1] = y; L0->ecoord[3 * ne + 2] = z;
                    ne++;
                }
    if (ne == 0) { vo_free(s); return -2; }
    if (number_nodes(L0)) { vo_free(s); return -3; }
    L0->fixed = (uint8_t *)calloc(L0->nn, 1);
    for (int64_t i = 0; i < nfixed; i++) {
        if (fixed[i] < 0 || fixed[i] >= L0->nn) { vo_free(s); return -4; }
        L0->fixed[fixed[i]] = 1;
    }
    for (int64_t i = 0; i < L0->nn; i++) L0->nfixed += L0->fixed[i];
    L0->fullscale = 1.0;
    L0->invdiag = (double *)malloc(sizeof(double) * 3 * L0->nn);
    inverse_diagonal(L0, s->ke, 1.0, L0->invdiag);

    int nlev = num_levels;
    for (int i = 1; i < nlev; i++) {
        vo_level *P = &s->L[i - 1], *C = &s->L[i];
        /* FEM.cpp:256-259 */
        C->rdx = P->rdx / 2 + P->rdx % 2; C->rdy = P->rdy / 2 + P->rdy % 2; C->rdz = P->rdz / 2 + P->rdz % 2;
        /* FEM.cpp:266-272 */
        int div = 2;
        if (i == 1 && skip > 0) div = 1 << (skip + 1);
        int nch = div * div * div;
        C->div = div; C->nb = div + 1;
        C->ex = (P->ex - 1) / div + 1; C->ey = (P->ey - 1) / div + 1; C->ez = (P->ez - 1) / div + 1;
        /* FEM.cpp:275-283: coarse element = floor(fine / div) */
        int64_t nce = (int64_t)C->ex * C->ey * C->ez;
        uint64_t *cm = (uint64_t *)calloc(nce, sizeof(uint64_t));
        for (int64_t e = 0; e < P->ne; e++) {
            int fx = P->ecoord[3 * e], fy = P->ecoord[3 * e + 1], fz = P->ecoord[3 * e + 2];
            int cx = fx / div, cy = fy / div, cz = fz / div;
            int ch = (fx - cx * div) + div * (fy - cy * div) + div * div * (fz - cz * div);
            cm[lin3(cx, cy, cz, C->ex, C->ey)] |= 1ull << ch;
        }
        int64_t cne = 0;
        for (int64_t g = 0; g < nce; g++) cne += (cm[g] != 0);
        C->ne = cne;
        C->ecoord = (int32_t *)malloc(sizeof(int32_t) * 3 * cne);
        C->childmask = (uint64_t *)malloc(sizeof(uint64_t) * cne);
        cne = 0;
        for (int z = 0; z < C->ez; z++)
            for (int y = 0; y < C->ey; y++)
                for (int x = 0; x < C->ex; x++) {
                    uint64_t m = cm[lin3(x, y, z, C->ex, C->ey)];
                    if (!m) continue;
                    C->ecoord[3 * cne] = x; C->ecoord[3 * cne + 1] = y; C->ecoord[3 * cne + 2] = z;
                    C->childmask[cne] = m;
                    cne++;
                }
        free(cm);
        if (number_nodes(C)) { vo_free(s); return -3; }
        /* restrictionMapping + multiplicity (FEM.cpp:340-362) */
        int nb = C->nb, nl = nb * nb * nb;
        C->rmap = (int32_t *)malloc(sizeof(int32_t) * nl * C->ne);
        double *count = (double *)calloc(3 * P->nn, sizeof(double));
        for (int64_t e = 0; e < C->ne; e++) {
            int bx = C->ecoord[3 * e] * div, by = C->ecoord[3 * e + 1] * div, bz = C->ecoord[3 * e + 2] * div;
            for (int c = 0; c < nb; c++)
                for (int b = 0; b < nb; b++)
                    for (int a = 0; a < nb; a++) {
                        int32_t id = node_at(P, bx + a, by + b, bz + c);
                        C->rmap[(int64_t)nl * e + a + nb * b + nb * nb * c] = id;
                        if (id >= 0) { count[3 * (int64_t)id]++; count[3 * (int64_t)id + 1]++; count[3 * (int64_t)id + 2]++; }
                    }
        }
        P->rcoef = (double *)malloc(sizeof(double) * 3 * P->nn);
        for (int64_t k = 0; k < 3 * P->nn; k++) P->rcoef[k] = 1.0 / count[k]; /* cwiseInverse, FEM.cpp:400 */
        free(count);
        /* unique Ks of partially filled coarse elements (FEM.cpp:364-395), cached per child mask */
        double scale = ldexp(1.0, i - 1);
        C->ukidx = (int32_t *)malloc(sizeof(int32_t) * C->ne);
        int hbits = 10, hcap = 1 << hbits;
        hslot *tab = (hslot *)malloc(sizeof(hslot) * hcap);
        for (int k = 0; k < hcap; k++) tab[k].val = -1;
        int ucap = 16;
        C->uk = (double *)malloc(sizeof(double) * 576 * ucap);
        C->nuk = 0;
        uint64_t full = (nch == 64) ? ~0ull : ((1ull << nch) - 1);
        for (int64_t e = 0; e < C->ne; e++) {
            uint64_t m = C->childmask[e];
            if (m == full) { C->ukidx[e] = -1; continue; }
            uint64_t h = hslot_index(m, hbits);
            while (tab[h].val >= 0 && tab[h].key != m) h = (h + 1) & (hcap - 1);
            if (tab[h].val < 0) {
                if (C->nuk == ucap) { ucap *= 2; C->uk = (double *)realloc(C->uk, sizeof(double) * 576 * ucap); }
                unique_k(s->ke, div, m, scale, C->uk + 576 * (int64_t)C->nuk);
                tab[h].key = m; tab[h].val = C->nuk++;
                if (C->nuk * 2 > hcap) { /* grow and rehash */
                    int nbits = hbits + 1, ncap = 1 << nbits;
                    hslot *nt = (hslot *)malloc(sizeof(hslot) * ncap);
                    for (int k = 0; k < ncap; k++) nt[k].val = -1;
                    for (int k = 0; k < hcap; k++)
                        if (tab[k].val >= 0) {
                            uint64_t hh = hslot_index(tab[k].key, nbits);
                            while (nt[hh].val >= 0) hh = (hh + 1) & (ncap - 1);
                            nt[hh] = tab[k];
                        }
                    free(tab); tab = nt; hcap = ncap; hbits = nbits;
                }
            }
            C->ukidx[e] = -1;
            for (uint64_t hh = hslot_index(m, hbits);; hh = (hh + 1) & (hcap - 1))
                if (tab[hh].key == m && tab[hh].val >= 0) { C->ukidx[e] = tab[hh].val; break; }
        }
        free(tab);
        /* fixed nodes on this level: restrict the fixed indicator (FEM.cpp:402-460) */
        {
            double *ind = (double *)calloc(3 * P->nn, sizeof(double));
            for (int64_t k = 0; k < P->nn; k++)
                if (P->fixed[k]) ind[3 * k] = ind[3 * k + 1] = ind[3 * k + 2] = 1.0;
            double *rc = (double *)malloc(sizeof(double) * 3 * C->nn);
            restrict_level(s, i - 1, ind, rc, P->rcoef);
            C->fixed = (uint8_t *)calloc(C->nn, 1);
            for (int64_t k = 0; k < 3 * C->nn; k++)
                if (rc[k] != 0.0) C->fixed[k / 3] = 1;
            for (int64_t k = 0; k < C->nn; k++) C->nfixed += C->fixed[k];
            free(ind); free(rc);
        }
        /* invDiag on this level (FEM.cpp:463): full elements use 2^(i+skip) Ke */
        C->fullscale = ldexp(1.0, i + skip);
        C->invdiag = (double *)malloc(sizeof(double) * 3 * C->nn);
        inverse_diagonal(C, s->ke, C->fullscale, C->invdiag);
        /* FEM.cpp:465-470 */
        if (C->rdx <= 2 || C->rdy <= 2 || C->rdz <= 2) { nlev = i + 1; break; }
    }
    s->nlev = nlev;

    /* coarsest level: free DOFs and Kc (FEM.cpp:473-513) */
    vo_level *Cc = &s->L[nlev - 1];
    int32_t *fidx = (int32_t *)malloc(sizeof(int32_t) * Cc->nn);
    int64_t nf = 0;
    for (int64_t k = 0; k < Cc->nn; k++) fidx[k] = Cc->fixed[k] ? -1 : (int32_t)nf++;
    s->nfree = 3 * nf;
    s->freedof = (int32_t *)malloc(sizeof(int32_t) * (s->nfree ? s->nfree : 1));
    for (int64_t k = 0; k < Cc->nn; k++)
        if (fidx[k] >= 0)
            for (int c = 0; c < 3; c++) s->freedof[3 * fidx[k] + c] = (int32_t)(3 * k + c);
    int64_t n = s->nfree;
    s->kc = (double *)calloc((size_t)(n ? n * n : 1), sizeof(double));
    s->kcg = (double *)calloc((size_t)(n ? n * n : 1), sizeof(double));
    double kscale = ldexp(1.0, nlev + skip - 1);
    for (int64_t e = 0; e < Cc->ne; e++) {
        int x = Cc->ecoord[3 * e], y = Cc->ecoord[3 * e + 1], z = Cc->ecoord[3 * e + 2];
        int32_t fn[8];
        for (int L = 0; L < 8; L++) {
            int32_t id = node_at(Cc, x + (L & 1), y + ((L >> 1) & 1), z + ((L >> 2) & 1));
            fn[L] = fidx[id];
        }
        const double *G = (Cc->ukidx && Cc->ukidx[e] >= 0) ? Cc->uk + 576 * (int64_t)Cc->ukidx[e] : NULL;
        for (int a = 0; a < 8; a++) {
            if (fn[a] < 0) continue;
            for (int b = 0; b < 8; b++) {
                if (fn[b] < 0) continue;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double kv = s->ke[(3 * a + i) * 24 + 3 * b + j];
                        s->kc[(3 * (int64_t)fn[a] + i) * n + 3 * (int64_t)fn[b] + j] += kscale * kv;
                        double gv = G ? G[(3 * a + i) * 24 + 3 * b + j] : Cc->fullscale * kv;
                        s->kcg[(3 * (int64_t)fn[a] + i) * n + 3 * (int64_t)fn[b] + j] += gv;
                    }
            }
        }
    }
    free(fidx);
    if (n > 0) {
        if (cholesky(s->kc, n)) { vo_free(s); return -5; }
        if (cholesky(s->kcg, n)) { vo_free(s); return -6; }
    }
    for (int l = 0; l < nlev; l++) {
        vo_level *L = &s->L[l];
        s->invdiag32[l] = (double *)malloc(sizeof(double) * 3 * L->nn);
        for (int64_t k = 0; k < 3 * L->nn; k++) s->invdiag32[l][k] = (double)(float)L->invdiag[k];
        if (L->rcoef) {
            s->rcoef32[l] = (double *)malloc(sizeof(double) * 3 * L->nn);
            for (int64_t k = 0; k < 3 * L->nn; k++) s->rcoef32[l][k] = (double)(float)L->rcoef[k];
        }
    }
    *out = s;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* operators                                                                 */
/* ------------------------------------------------------------------------- */

/* y = A x on level 0: element loop then zero fixed rows
 * (MatrixFreeSparse.hpp:198-214).  Elements of one parity class share no node. */
void vo_matvec(const vo_sys *s, const double *x, double *y)
{
    const vo_level *L = &s->L[0];
    memset(y, 0, sizeof(double) * 3 * L->nn);
    for (int color = 0; color < 8; color++) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < L->ne; e++) {
            const int32_t *ec = L->ecoord + 3 * e;
            if (((ec[0] & 1) | ((ec[1] & 1) << 1) | ((ec[2] & 1) << 2)) != color) continue;
            double xe[24];
            for (int d = 0; d < 24; d++) xe[d] = x[gdof(L->e2n, e, d)];
            for (int i = 0; i < 24; i++) {
                double acc = 0;
                for (int j = 0; j < 24; j++) acc += s->ke[i * 24 + j] * xe[j];
                y[gdof(L->e2n, e, i)] += acc;
            }
        }
    }
    for (int64_t k = 0; k < L->nn; k++)
        if (L->fixed[k]) y[3 * k] = y[3 * k + 1] = y[3 * k + 2] = 0;
}

/* coarse-level operator of the damped multiplicative cycle: per-element matrices
 * equal to the reference's invDiag elements (2^(l+skip) Ke on full elements, the
 * unique Galerkin K on partial ones), fixed rows zeroed. */
void vo_level_matvec(const vo_sys *s, int l, const double *x, double *y)
{
    const vo_level *L = &s->L[l];
    memset(y, 0, sizeof(double) * 3 * L->nn);
    for (int color = 0; color < 8; color++) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < L->ne; e++) {
            const int32_t *ec = L->ecoord + 3 * e;
            if (((ec[0] & 1) | ((ec[1] & 1) << 1) | ((ec[2] & 1) << 2)) != color) continue;
            const double *K = (L->ukidx && L->ukidx[e] >= 0) ? L->uk + 576 * (int64_t)L->ukidx[e] : NULL;
            double sc = K ? 1.0 : L->fullscale;
            if (!K) K = s->ke;
            double xe[24];
            for (int d = 0; d < 24; d++) xe[d] = x[gdof(L->e2n, e, d)];
            for (int i = 0; i < 24; i++) {
                double acc = 0;
                for (int j = 0; j < 24; j++) acc += K[i * 24 + j] * xe[j];
                y[gdof(L->e2n, e, i)] += sc * acc;
            }
        }
    }
    for (int64_t k = 0; k < L->nn; k++)
        if (L->fixed[k]) y[3 * k] = y[3 * k + 1] = y[3 * k + 2] = 0;
}

static void coarse_solve(const vo_sys *s, const double *Lf, const double *r, double *x)
{
    const vo_level *C = &s->L[s->nlev - 1];
    memset(x, 0, sizeof(double) * 3 * C->nn);
    int64_t n = s->nfree;
    if (n == 0) return;
    double *b = (double *)malloc(sizeof(double) * n);
    for (int64_t k = 0; k < n; k++) b[k] = r[s->freedof[k]];
    chol_solve(Lf, n, b);
    for (int64_t k = 0; k < n; k++) x[s->freedof[k]] = b[k];
    free(b);
}

/* Reference additive V-cycle (MultigridPreconditioner.hpp:158-193).
 * fp32 != 0 rounds mu and invDiag to float like the Vulkan path (FEMGPU.cpp:226-242). */
static void vcycle_ref(const vo_sys *s, const double *b, double *x, int fp32)
{
    int nl = s->nlev;
    double *rs[VO_MAXL], *xs[VO_MAXL];
    for (int l = 0; l < nl; l++) {
        rs[l] = (double *)malloc(sizeof(double) * 3 * s->L[l].nn);
        xs[l] = (double *)calloc(3 * s->L[l].nn, sizeof(double));
    }
    memcpy(rs[0], b, sizeof(double) * 3 * s->L[0].nn);
    for (int l = 0; l < nl - 1; l++) {
        const double *dinv = fp32 ? s->invdiag32[l] : s->L[l].invdiag;
        for (int64_t k = 0; k < 3 * s->L[l].nn; k++) xs[l][k] += rs[l][k] * dinv[k];
        restrict_level(s, l, rs[l], rs[l + 1], fp32 ? s->rcoef32[l] : s->L[l].rcoef);
    }
    coarse_solve(s, s->kc, rs[nl - 1], xs[nl - 1]);
    double *tmp = (double *)malloc(sizeof(double) * 3 * s->L[0].nn);
    for (int l = nl - 1; l > 0; l--) {
        interpolate_level(s, l - 1, xs[l], tmp, fp32 ? s->rcoef32[l - 1] : s->L[l - 1].rcoef);
        const double *dinv = fp32 ? s->invdiag32[l - 1] : s->L[l - 1].invdiag;
        for (int64_t k = 0; k < 3 * s->L[l - 1].nn; k++) xs[l - 1][k] += tmp[k];
        for (int64_t k = 0; k < 3 * s->L[l - 1].nn; k++) xs[l - 1][k] += rs[l - 1][k] * dinv[k];
    }
    free(tmp);
    memcpy(x, xs[0], sizeof(double) * 3 * s->L[0].nn);
    for (int64_t k = 0; k < s->L[0].nn; k++)
        if (s->L[0].fixed[k]) x[3 * k] = x[3 * k + 1] = x[3 * k + 2] = 0;
    for (int l = 0; l < nl; l++) { free(rs[l]); free(xs[l]); }
}

static void mask_fixed(const vo_level *L, double *v)
{
    for (int64_t k = 0; k < L->nn; k++)
        if (L->fixed[k]) v[3 * k] = v[3 * k + 1] = v[3 * k + 2] = 0;
}

/* Chebyshev smoothing of A x = r on level l (Saad, "Iterative Methods for Sparse
 * Linear Systems", Alg. 12.1, preconditioned by D^-1), `steps` steps from x, or
 * from x = 0 when zero != 0, on the interval [lmax/ratio, lmax]; t is scratch.
 * With x = 0 the first step needs no operator product. */
static void cheb_smooth(const vo_sys *s, int l, const double *r, double *x, double *t, int zero, int steps)
{
    const vo_level *L = &s->L[l];
    int64_t n3 = 3 * L->nn;
    double hi = s->cheb_lmax[l], lo = hi / s->cheb_ratio;
    double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta, rho = 1 / sigma;
    double *d = (double *)malloc(sizeof(double) * n3);
    for (int k = 0; k < steps; k++) {
        double a, b;
        if (k == 0) {
            a = 0;
            b = 1 / theta;
        } else {
            double rho1 = 1 / (2 * sigma - rho);
            a = rho1 * rho;
            b = 2 * rho1 / delta;
            rho = rho1;
        }
        if (k == 0 && zero) {
            for (int64_t i = 0; i < n3; i++) t[i] = 0;
        } else if (l == 0) vo_matvec(s, x, t); else vo_level_matvec(s, l, x, t);
        for (int64_t i = 0; i < n3; i++) {
            d[i] = (k ? a * d[i] : 0.0) + b * L->invdiag[i] * (r[i] - t[i]);
            x[i] = (k == 0 && zero ? 0.0 : x[i]) + d[i];
        }
        mask_fixed(L, x);
        for (int64_t i = 0; i < L->nn; i++)
            if (L->fixed[i]) d[3 * i] = d[3 * i + 1] = d[3 * i + 2] = 0;
    }
    free(d);
}

/* Damped multiplicative V-cycle (no reference counterpart: the configured
 * damped-Jacobi variant with one pre- and one post-smooth; on levels >=
 * cheb_level, Chebyshev smoothing of cheb_degree steps before and after).
 * Level operator: vo_level_matvec; transfers: the reference's
 * restrict/interpolate; coarsest: Kc assembled from the same element matrices.
 * Fixed DOFs stay zero. */
static void vcycle_damped_level(const vo_sys *s, int l, const double *r, double *x, double omega)
{
    const vo_level *L = &s->L[l];
    int64_t n3 = 3 * L->nn;
    if (l == s->nlev - 1) { coarse_solve(s, s->kcg, r, x); return; }
    const int cheb = s->cheb_level > 0 && l >= s->cheb_level;
    double *t = (double *)malloc(sizeof(double) * n3);
    double *res = (double *)malloc(sizeof(double) * n3);
    if (cheb) {
        cheb_smooth(s, l, r, x, t, 1, s->cheb_degree);
    } else {
        for (int64_t k = 0; k < n3; k++) x[k] = omega * L->invdiag[k] * r[k];
        mask_fixed(L, x);
    }
    if (l == 0) vo_matvec(s, x, t); else vo_level_matvec(s, l, x, t);
    for (int64_t k = 0; k < n3; k++) res[k] = r[k] - t[k];
    const vo_level *C = &s->L[l + 1];
    double *rc = (double *)malloc(sizeof(double) * 3 * C->nn);
    double *xc = (double *)malloc(sizeof(double) * 3 * C->nn);
    restrict_level(s, l, res, rc, L->rcoef);
    mask_fixed(C, rc);
    vcycle_damped_level(s, l + 1, rc, xc, omega);
    interpolate_level(s, l, xc, t, L->rcoef);
    mask_fixed(L, t);
    for (int64_t k = 0; k < n3; k++) x[k] += t[k];
    if (cheb) {
        cheb_smooth(s, l, r, x, t, 0, s->cheb_degree);
    } else {
        if (l == 0) vo_matvec(s, x, t); else vo_level_matvec(s, l, x, t);
        for (int64_t k = 0; k < n3; k++) x[k] += omega * L->invdiag[k] * (r[k] - t[k]);
        mask_fixed(L, x);
    }
    free(t); free(res); free(rc); free(xc);
}

/* Chebyshev smoothing of the damped cycle on levels >= level (0 = off), with
 * `degree` steps before and after the coarse correction, fitted to
 * [lmax[l] / ratio, lmax[l]] (lmax: one bound of lambda_max(D^-1 A) per level). */
int vo_set_chebyshev(vo_sys *s, int level, int degree, double ratio, const double *lmax)
{
    if (level < 0 || (level > 0 && (degree < 1 || !(ratio > 1)))) return 1;
    s->cheb_level = level;
    s->cheb_degree = degree;
    s->cheb_ratio = ratio;
    for (int l = 0; l < s->nlev; l++) s->cheb_lmax[l] = level > 0 ? lmax[l] : 0;
    if (level > 0)
        for (int l = level; l < s->nlev - 1; l++)
            if (!(s->cheb_lmax[l] > 0)) return 1;
    return 0;
}

/* mode 0: reference additive cycle, 1: reference additive cycle with FP32 mu/invDiag,
 * 2: damped multiplicative cycle */
void vo_vcycle(const vo_sys *s, const double *r, double *z, int mode, double omega)
{
    if (mode == 2) {
        double *rr = (double *)malloc(sizeof(double) * 3 * s->L[0].nn);
        memcpy(rr, r, sizeof(double) * 3 * s->L[0].nn);
        mask_fixed(&s->L[0], rr);
        vcycle_damped_level(s, 0, rr, z, omega);
        free(rr);
    } else {
        vcycle_ref(s, r, z, mode == 1);
    }
}

void vo_restrict(const vo_sys *s, int l, const double *r, double *rc) { restrict_level(s, l, r, rc, s->L[l].rcoef); }
void vo_interpolate(const vo_sys *s, int l, const double *xc, double *xf) { interpolate_level(s, l, xc, xf, s->L[l].rcoef); }

static double dot(const double *a, const double *b, int64_t n)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Eigen 3.4 ConjugateGradient semantics (FEM.cpp:607-628; loop as mirrored by
 * FEMGPU.cpp:416-541).  x starts at zero, or at the x passed in when warm != 0
 * (Eigen's solveWithGuess: r = b - A x0).  precond: 0/1/2 as vo_vcycle mode,
 * -1 = identity (IdentityPreconditioner, FEM.cpp:610).  Returns iterations. */
int vo_pcg(const vo_sys *s, const double *b, double *x, double tol, int maxit, int precond, double omega,
           double *relres_out, int warm)
{
    int64_t n = 3 * s->L[0].nn;
    if (!warm) memset(x, 0, sizeof(double) * n);
    double rhs2 = dot(b, b, n);
    if (rhs2 == 0) { memset(x, 0, sizeof(double) * n); if (relres_out) *relres_out = 0; return 0; }
    double *r = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n), *t = (double *)malloc(sizeof(double) * n);
    if (warm) {
        vo_matvec(s, x, t);
        for (int64_t k = 0; k < n; k++) r[k] = b[k] - t[k];
    } else {
        memcpy(r, b, sizeof(double) * n); /* r = b - A*0 */
    }
    double threshold = tol * tol * rhs2;
    if (threshold < DBL_MIN) threshold = DBL_MIN;
    double r2 = dot(r, r, n);
    int it = 0;
    if (r2 >= threshold) {
        if (precond < 0) memcpy(p, r, sizeof(double) * n); else vo_vcycle(s, r, p, precond, omega);
        double absNew = dot(r, p, n);
        while (it < maxit) {
            vo_matvec(s, p, t);
            double alpha = absNew / dot(p, t, n);
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] -= alpha * t[k]; }
            r2 = dot(r, r, n);
            if (r2 < threshold) break;
            if (precond < 0) memcpy(z, r, sizeof(double) * n); else vo_vcycle(s, r, z, precond, omega);
            double absOld = absNew;
            absNew = dot(r, z, n);
            double beta = absNew / absOld;
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
            it++;
        }
    }
    if (relres_out) *relres_out = sqrt(r2 / rhs2);
    free(r); free(p); free(z); free(t);
    return it;
}

/* ------------------------------------------------------------------------- */
/* accessors for the tests                                                   */
/* ------------------------------------------------------------------------- */
int vo_num_levels(const vo_sys *s) { return s->nlev; }
int64_t vo_level_nodes(const vo_sys *s, int l) { return s->L[l].nn; }
int64_t vo_level_elems(const vo_sys *s, int l) { return s->L[l].ne; }
int64_t vo_level_nfixed(const vo_sys *s, int l) { return s->L[l].nfixed; }
void vo_level_extent(const vo_sys *s, int l, int32_t *e3) { e3[0] = s->L[l].ex; e3[1] = s->L[l].ey; e3[2] = s->L[l].ez; }
void vo_get_ecoord(const vo_sys *s, int l, int32_t *out) { memcpy(out, s->L[l].ecoord, sizeof(int32_t) * 3 * s->L[l].ne); }
void vo_get_e2n(const vo_sys *s, int l, int32_t *out) { memcpy(out, s->L[l].e2n, sizeof(int32_t) * 4 * s->L[l].ne); }
void vo_get_invdiag(const vo_sys *s, int l, double *out) { memcpy(out, s->L[l].invdiag, sizeof(double) * 3 * s->L[l].nn); }
int vo_get_rcoef(const vo_sys *s, int l, double *out)
{
    if (!s->L[l].rcoef) return -1;
    memcpy(out, s->L[l].rcoef, sizeof(double) * 3 * s->L[l].nn);
    return 0;
}
void vo_get_fixed(const vo_sys *s, int l, uint8_t *out) { memcpy(out, s->L[l].fixed, s->L[l].nn); }
int vo_level_nb(const vo_sys *s, int l) { return s->L[l].nb; }
int vo_get_rmap(const vo_sys *s, int l, int32_t *out)
{
    if (l < 1 || !s->L[l].rmap) return -1;
    int nb = s->L[l].nb;
    memcpy(out, s->L[l].rmap, sizeof(int32_t) * nb * nb * nb * s->L[l].ne);
    return nb;
}
/* per-element matrix of level l (the reference's unique K or fullscale*Ke) */
void vo_get_element_matrix(const vo_sys *s, int l, int64_t e, double *K)
{
    const vo_level *L = &s->L[l];
    if (L->ukidx && L->ukidx[e] >= 0) memcpy(K, L->uk + 576 * (int64_t)L->ukidx[e], sizeof(double) * 576);
    else for (int k = 0; k < 576; k++) K[k] = L->fullscale * s->ke[k];
}
int64_t vo_nfree(const vo_sys *s) { return s->nfree; }
void vo_get_freedof(const vo_sys *s, int32_t *out) { memcpy(out, s->freedof, sizeof(int32_t) * s->nfree); }
/* Kc = L L^T reconstructed from the factor (which = 0: reference Kc, 1: damped-cycle Kc) */
void vo_get_kc(const vo_sys *s, int which, double *out)
{
    int64_t n = s->nfree;
    const double *Lf = which ? s->kcg : s->kc;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
            double acc = 0;
            int64_t m = i < j ? i : j;
            for (int64_t k = 0; k <= m; k++) acc += Lf[i * n + k] * Lf[j * n + k];
            out[i * n + j] = acc;
        }
}
void vo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int vo_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
