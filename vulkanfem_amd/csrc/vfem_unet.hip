// Learned coarse-grid correction: a 3-D conv U-Net with hand-written HIP
// training and inference kernels.  No counterpart in the reference (its coarse
// correction is the multigrid recursion of include/MultigridPreconditioner.hpp:158-193);
// the network maps a coarse level's residual (3 channels) and its occupancy to
// the error estimate that the damped cycle would otherwise get from the levels
// below.
//
// Tensors are [B][C][D][H][W] FP32 (x fastest).  Weights: FP32 master copies
// [cout][cin][k^3] for the optimiser, BF16 copies for the arithmetic in two
// layouts -- [cin][k^3][cout] (forward: the 8 output channels a thread owns are
// contiguous) and [cout][k^3][cin] (input gradient) -- FP32 accumulation.
// Layers (no biases, so N(a r) = a N(r) for a > 0):
//
//   e0a  x            4 -> 16  3^3 s1 ReLU      d3  d2b       32 -> 64 3^3 s2 ReLU
//   e0b  e0a         16 -> 16  3^3 s1 ReLU      u2  up(d3)|d2b 96 -> 32 3^3 s1 ReLU
//   d1a  e0b         16 -> 24  3^3 s2 ReLU      u1  up(u2)|d1b 56 -> 24 3^3 s1 ReLU
//   d1b  d1a         24 -> 24  3^3 s1 ReLU      u0  up(u1)|e0b 40 -> 16 3^3 s1 ReLU
//   d2a  d1b         24 -> 32  3^3 s2 ReLU      out u0         16 -> 3  1^3 s1
//   d2b  d2a         32 -> 32  3^3 s1 ReLU
//
// A stride-2 3^3 convolution (padding 1) maps n nodes to (n-1)/2 + 1; up() is
// node-aligned trilinear interpolation back to the skip tensor's size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "vfem.h"
#include "vfem_unet.h"

namespace vfem {
namespace unet {

constexpr int kT = 256;  // threads per block
constexpr int kCB = 8;   // output (forward) / input (backward) channels per thread

struct LayerDef {
    const char *name;
    int cin0, cin1, cout, k, s, relu;
};
const LayerDef kLayers[kNumLayers] = {
    {"e0a", 4, 0, 16, 3, 1, 1},  {"e0b", 16, 0, 16, 3, 1, 1}, {"d1a", 16, 0, 24, 3, 2, 1},
    {"d1b", 24, 0, 24, 3, 1, 1}, {"d2a", 24, 0, 32, 3, 2, 1}, {"d2b", 32, 0, 32, 3, 1, 1},
    {"d3", 32, 0, 64, 3, 2, 1},  {"u2", 64, 32, 32, 3, 1, 1}, {"u1", 32, 24, 24, 3, 1, 1},
    {"u0", 24, 16, 16, 3, 1, 1}, {"out", 16, 0, 3, 1, 1, 0},
};
enum { E0A, E0B, D1A, D1B, D2A, D2B, D3, U2, U1, U0, OUT };

int64_t layer_params(int l)
{
    const LayerDef &d = kLayers[l];
    return (int64_t)(d.cin0 + d.cin1) * d.cout * d.k * d.k * d.k;
}

inline int down(int n) { return (n - 1) / 2 + 1; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// out[b][o][v'] = act(sum_{c, t} w[c][t][o] in[b][c][S v' + t - p])
__global__ __launch_bounds__(kT) void k_conv_fwd(const float *__restrict__ in0, int C0, const float *__restrict__ in1,
                                                 int C1, int B, int D, int H, int W, const __bf16 *__restrict__ wf,
                                                 int Cout, int K, int S, int Do, int Ho, int Wo, int relu,
                                                 float *__restrict__ out)
{
    const int64_t Vo = (int64_t)Do * Ho * Wo, V = (int64_t)D * H * W;
    const int ncb = (Cout + kCB - 1) / kCB;
    const int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= (int64_t)B * ncb * Vo) return;
    const int64_t v = idx % Vo;
    const int cb = (int)((idx / Vo) % ncb), b = (int)(idx / (Vo * ncb));
    const int ox = (int)(v % Wo), oy = (int)((v / Wo) % Ho), oz = (int)(v / ((int64_t)Wo * Ho));
    const int p = K / 2, K3 = K * K * K, Cin = C0 + C1;
    const int nj = min(kCB, Cout - cb * kCB);
    float acc[kCB];
#pragma unroll
    for (int j = 0; j < kCB; j++) acc[j] = 0.0f;
    for (int ci = 0; ci < Cin; ci++) {
        const float *src = ci < C0 ? in0 + ((int64_t)b * C0 + ci) * V : in1 + ((int64_t)b * C1 + (ci - C0)) * V;
        for (int tz = 0; tz < K; tz++) {
            const int iz = oz * S + tz - p;
            if (iz < 0 || iz >= D) continue;
            for (int ty = 0; ty < K; ty++) {
                const int iy = oy * S + ty - p;
                if (iy < 0 || iy >= H) continue;
                for (int tx = 0; tx < K; tx++) {
                    const int ix = ox * S + tx - p;
                    if (ix < 0 || ix >= W) continue;
                    const float xv = src[((int64_t)iz * H + iy) * W + ix];
                    const __bf16 *wp = wf + ((int64_t)ci * K3 + (tz * K + ty) * K + tx) * Cout + cb * kCB;
#pragma unroll
                    for (int j = 0; j < kCB; j++)
                        if (j < nj) acc[j] += (float)wp[j] * xv;
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kCB; j++)
        if (j < nj) {
            const float o = relu ? fmaxf(acc[j], 0.0f) : acc[j];
            out[((int64_t)b * Cout + cb * kCB + j) * Vo + v] = o;
        }
}

// g *= (y > 0): gradient through a ReLU, in place
__global__ __launch_bounds__(kT) void k_relu_bwd(int64_t n, const float *__restrict__ y, float *__restrict__ g)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n && !(y[i] > 0.0f)) g[i] = 0.0f;
}

// gx[b][c][v] += sum_{o, t} w[o][t][c] gy[b][o][(v + p - t) / S]   (input gradient)
__global__ __launch_bounds__(kT) void k_conv_bwd_data(const float *__restrict__ gy, int Cout, int B, int Do, int Ho,
                                                      int Wo, const __bf16 *__restrict__ wb, int C0, int C1,
                                                      float *__restrict__ gx0, float *__restrict__ gx1, int D, int H,
                                                      int W, int K, int S)
{
    const int64_t Vo = (int64_t)Do * Ho * Wo, V = (int64_t)D * H * W;
    const int Cin = C0 + C1, ncb = (Cin + kCB - 1) / kCB;
    const int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= (int64_t)B * ncb * V) return;
    const int64_t v = idx % V;
    const int cb = (int)((idx / V) % ncb), b = (int)(idx / (V * ncb));
    const int ix = (int)(v % W), iy = (int)((v / W) % H), iz = (int)(v / ((int64_t)W * H));
    const int p = K / 2, K3 = K * K * K;
    const int nj = min(kCB, Cin - cb * kCB);
    float acc[kCB];
#pragma unroll
    for (int j = 0; j < kCB; j++) acc[j] = 0.0f;
    for (int co = 0; co < Cout; co++) {
        const float *g = gy + ((int64_t)b * Cout + co) * Vo;
        for (int tz = 0; tz < K; tz++) {
            const int nz = iz + p - tz;
            if (nz < 0 || nz % S) continue;
            const int oz = nz / S;
            if (oz >= Do) continue;
            for (int ty = 0; ty < K; ty++) {
                const int ny = iy + p - ty;
                if (ny < 0 || ny % S) continue;
                const int oy = ny / S;
                if (oy >= Ho) continue;
                for (int tx = 0; tx < K; tx++) {
                    const int nx = ix + p - tx;
                    if (nx < 0 || nx % S) continue;
                    const int ox = nx / S;
                    if (ox >= Wo) continue;
                    const float gv = g[((int64_t)oz * Ho + oy) * Wo + ox];
                    const __bf16 *wp = wb + ((int64_t)co * K3 + (tz * K + ty) * K + tx) * Cin + cb * kCB;
#pragma unroll
                    for (int j = 0; j < kCB; j++)
                        if (j < nj) acc[j] += (float)wp[j] * gv;
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kCB; j++) {
        if (j >= nj) continue;
        const int ci = cb * kCB + j;
        if (ci < C0) {
            if (gx0) gx0[((int64_t)b * C0 + ci) * V + v] += acc[j];
        } else if (gx1) {
            gx1[((int64_t)b * C1 + (ci - C0)) * V + v] += acc[j];
        }
    }
}

// gw[o][c][t] = sum_{b, v'} gy[b][o][v'] in[b][c][S v' + t - p]; one block per (o, c),
// fixed-order block reduction (deterministic)
__global__ __launch_bounds__(kT) void k_conv_bwd_weight(const float *__restrict__ gy, int Cout, int B, int Do, int Ho,
                                                        int Wo, const float *__restrict__ in0, int C0,
                                                        const float *__restrict__ in1, int C1, int D, int H, int W,
                                                        int K, int S, float *__restrict__ gw)
{
    const int Cin = C0 + C1;
    const int co = blockIdx.x / Cin, ci = blockIdx.x % Cin;
    const int64_t Vo = (int64_t)Do * Ho * Wo, V = (int64_t)D * H * W;
    const int p = K / 2, K3 = K * K * K;
    float acc[27];
#pragma unroll
    for (int t = 0; t < 27; t++) acc[t] = 0.0f;
    for (int64_t n = threadIdx.x; n < (int64_t)B * Vo; n += kT) {
        const int b = (int)(n / Vo);
        const int64_t o = n % Vo;
        const float gv = gy[((int64_t)b * Cout + co) * Vo + o];
        if (gv == 0.0f) continue;
        const int ox = (int)(o % Wo), oy = (int)((o / Wo) % Ho), oz = (int)(o / ((int64_t)Wo * Ho));
        const float *src = ci < C0 ? in0 + ((int64_t)b * C0 + ci) * V : in1 + ((int64_t)b * C1 + (ci - C0)) * V;
#pragma unroll
        for (int t = 0; t < 27; t++) {
            if (t >= K3) break;
            const int tz = t / (K * K), ty = (t / K) % K, tx = t % K;
            const int iz = oz * S + tz - p, iy = oy * S + ty - p, ix = ox * S + tx - p;
            if (iz >= 0 && iz < D && iy >= 0 && iy < H && ix >= 0 && ix < W)
                acc[t] += gv * src[((int64_t)iz * H + iy) * W + ix];
        }
    }
    __shared__ float sh[kT / 64][27];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < 27; t++) {
        float v = acc[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wave][t] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < K3) {
        float s = 0.0f;
        for (int w = 0; w < kT / 64; w++) s += sh[w][threadIdx.x];
        gw[((int64_t)co * Cin + ci) * K3 + threadIdx.x] = s;
    }
}

__device__ __forceinline__ void up_weights(int j, int m, int &i0, int &i1, float &w0, float &w1)
{
    i0 = min(j / 2, m - 1);
    if (j % 2 == 0) {
        i1 = i0;
        w0 = 1.0f;
        w1 = 0.0f;
    } else {
        i1 = min(j / 2 + 1, m - 1);
        w0 = w1 = 0.5f;
    }
}

// out[b][c][Z][Y][X] = node-aligned trilinear interpolation of in[b][c] (m -> n per axis)
__global__ __launch_bounds__(kT) void k_upsample(const float *__restrict__ in, int BC, int md, int mh, int mw,
                                                 float *__restrict__ out, int nd, int nh, int nw)
{
    const int64_t V = (int64_t)nd * nh * nw;
    const int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= BC * V) return;
    const int64_t v = idx % V, bc = idx / V;
    const int X = (int)(v % nw), Y = (int)((v / nw) % nh), Z = (int)(v / ((int64_t)nw * nh));
    int z0, z1, y0, y1, x0, x1;
    float wz0, wz1, wy0, wy1, wx0, wx1;
    up_weights(Z, md, z0, z1, wz0, wz1);
    up_weights(Y, mh, y0, y1, wy0, wy1);
    up_weights(X, mw, x0, x1, wx0, wx1);
    const float *s = in + bc * ((int64_t)md * mh * mw);
    auto at = [&](int z, int y, int x) { return s[((int64_t)z * mh + y) * mw + x]; };
    const float r = wz0 * (wy0 * (wx0 * at(z0, y0, x0) + wx1 * at(z0, y0, x1)) +
                           wy1 * (wx0 * at(z0, y1, x0) + wx1 * at(z0, y1, x1))) +
                    wz1 * (wy0 * (wx0 * at(z1, y0, x0) + wx1 * at(z1, y0, x1)) +
                           wy1 * (wx0 * at(z1, y1, x0) + wx1 * at(z1, y1, x1)));
    out[idx] = r;
}

// gin[b][c][coarse] += sum over fine nodes of (interpolation weight) * gout[fine]
__global__ __launch_bounds__(kT) void k_upsample_bwd(const float *__restrict__ gout, int BC, int nd, int nh, int nw,
                                                     float *__restrict__ gin, int md, int mh, int mw)
{
    const int64_t V = (int64_t)md * mh * mw;
    const int64_t idx = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (idx >= BC * V) return;
    const int64_t v = idx % V, bc = idx / V;
    const int ix = (int)(v % mw), iy = (int)((v / mw) % mh), iz = (int)(v / ((int64_t)mw * mh));
    const float *g = gout + bc * ((int64_t)nd * nh * nw);
    auto wt = [](int j, int m, int i) {
        int i0, i1;
        float w0, w1;
        up_weights(j, m, i0, i1, w0, w1);
        return (i0 == i ? w0 : 0.0f) + (i1 == i ? w1 : 0.0f);
    };
    float acc = 0.0f;
    for (int Z = max(2 * iz - 1, 0); Z <= min(2 * iz + 1, nd - 1); Z++) {
        const float a = wt(Z, md, iz);
        if (a == 0.0f) continue;
        for (int Y = max(2 * iy - 1, 0); Y <= min(2 * iy + 1, nh - 1); Y++) {
            const float bw = a * wt(Y, mh, iy);
            if (bw == 0.0f) continue;
            for (int X = max(2 * ix - 1, 0); X <= min(2 * ix + 1, nw - 1); X++) {
                const float cw = bw * wt(X, mw, ix);
                if (cw != 0.0f) acc += cw * g[((int64_t)Z * nh + Y) * nw + X];
            }
        }
    }
    gin[idx] += acc;
}

__global__ __launch_bounds__(kT) void k_fill(int64_t n, float *__restrict__ p, float v)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

// Masked relative squared error per sample, loss[b] = |m (o - t)|^2 / |m t|^2 and
// gradient g = 2 m (o - t) / (|m t|^2 B).  One block per sample.
__global__ __launch_bounds__(kT) void k_mse(const float *__restrict__ out, const float *__restrict__ tgt,
                                            const float *__restrict__ mask, int64_t V, int B,
                                            float *__restrict__ gout, float *__restrict__ loss)
{
    const int b = blockIdx.x;
    const float *o = out + (int64_t)b * 3 * V, *t = tgt + (int64_t)b * 3 * V, *m = mask + (int64_t)b * V;
    double num = 0, den = 0;
    for (int64_t i = threadIdx.x; i < 3 * V; i += kT) {
        const double mk = m[i % V], d = (double)o[i] - (double)t[i];
        num += mk * d * d;
        den += mk * (double)t[i] * (double)t[i];
    }
    __shared__ double sh[2][kT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        num += __shfl_xor(num, off, 64);
        den += __shfl_xor(den, off, 64);
    }
    if (lane == 0) {
        sh[0][wave] = num;
        sh[1][wave] = den;
    }
    __syncthreads();
    double N = 0, Dn = 0;
    for (int w = 0; w < kT / 64; w++) {
        N += sh[0][w];
        Dn += sh[1][w];
    }
    Dn = Dn > 1e-30 ? Dn : 1e-30;
    if (threadIdx.x == 0) loss[b] = (float)(N / Dn);
    const float sc = (float)(2.0 / (Dn * B));
    for (int64_t i = threadIdx.x; i < 3 * V; i += kT) gout[(int64_t)b * 3 * V + i] = sc * m[i % V] * (o[i] - t[i]);
}

// BF16 copies of a layer's master weights [o][c][t] in the two arithmetic layouts
__device__ __forceinline__ void pack_one(int64_t i, float w, int cout, int cin, int k3, __bf16 *wf, __bf16 *wb)
{
    const int t = (int)(i % k3), c = (int)((i / k3) % cin), o = (int)(i / ((int64_t)k3 * cin));
    const __bf16 h = (__bf16)w;
    wf[((int64_t)c * k3 + t) * cout + o] = h;
    wb[((int64_t)o * k3 + t) * cin + c] = h;
}

__global__ __launch_bounds__(kT) void k_pack(int64_t n, const float *__restrict__ w, int cout, int cin, int k3,
                                             __bf16 *__restrict__ wf, __bf16 *__restrict__ wb)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) pack_one(i, w[i], cout, cin, k3, wf, wb);
}

// Adam step on the FP32 master weights, then the BF16 copies
__global__ __launch_bounds__(kT) void k_adam(int64_t n, float *__restrict__ w, const float *__restrict__ g,
                                             float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                             float b2, float eps, float c1, float c2, int cout, int cin, int k3,
                                             __bf16 *__restrict__ wf, __bf16 *__restrict__ wb)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    pack_one(i, wi, cout, cin, k3, wf, wb);
}

// batch[k] = data[idx[k]] (sample = n floats)
__global__ __launch_bounds__(kT) void k_gather(const float *__restrict__ data, const int32_t *__restrict__ idx,
                                               int64_t n, int count, float *__restrict__ batch)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= n * count) return;
    batch[i] = data[(int64_t)idx[i / n] * n + i % n];
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace unet
}  // namespace vfem

using namespace vfem::unet;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {

thread_local std::string u_err;

#define U_TRY(expr)                                                                                            \
    do {                                                                                                       \
        hipError_t _e = (expr);                                                                                \
        if (_e != hipSuccess) {                                                                                \
            u_err = std::string(#expr) + ": " + hipGetErrorString(_e);                                         \
            return VFEM_ERR_DEVICE;                                                                            \
        }                                                                                                      \
    } while (0)

struct Dims {
    int d, h, w;
    int64_t v() const { return (int64_t)d * h * w; }
};

// activations (and their gradients) for one batch shape
struct Acts {
    int B = 0;
    Dims dim[4] = {};                 // node grids of the four resolutions
    float *a[kNumLayers] = {};        // layer outputs
    float *up[3] = {};                // up(d3) at dim 2, up(u2) at dim 1, up(u1) at dim 0
    float *x = nullptr;               // input [B][4][dim 0]
    float *ga[kNumLayers] = {}, *gup[3] = {}, *gx = nullptr;  // gradients (training only)
    bool grads = false;
};

int layer_res(int l)
{
    static const int r[kNumLayers] = {0, 0, 1, 1, 2, 2, 3, 2, 1, 0, 0};
    return r[l];
}

}  // namespace

struct vfem_unet {
    float *w[kNumLayers] = {}, *g[kNumLayers] = {}, *m[kNumLayers] = {}, *v[kNumLayers] = {};
    __bf16 *wf[kNumLayers] = {}, *wb[kNumLayers] = {};
    int64_t n[kNumLayers] = {};
    int step = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs, act_allocs[3];
    Acts acts[3];  // 0: inference, 1: training, 2: test hooks
    float *loss = nullptr;
    int device = 0;

    int alloc(void **p, size_t bytes, int set = -1)
    {
        U_TRY(hipMalloc(p, std::max<size_t>(bytes, 16)));
        (set < 0 ? allocs : act_allocs[set]).push_back(*p);
        return 0;
    }
    ~vfem_unet()
    {
        for (void *p : allocs) (void)hipFree(p);
        for (auto &l : act_allocs)
            for (void *p : l) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int prepare(vfem_unet *N, int set, int B, int d, int h, int w, bool grads)
{
    Acts &A = N->acts[set];
    if (A.B == B && A.dim[0].d == d && A.dim[0].h == h && A.dim[0].w == w && (A.grads || !grads)) return 0;
    U_TRY(hipStreamSynchronize(N->stream));
    for (void *p : N->act_allocs[set]) (void)hipFree(p);
    N->act_allocs[set].clear();
    A = Acts();
    A.B = B;
    A.grads = grads;
    A.dim[0] = {d, h, w};
    for (int r = 1; r < 4; r++) A.dim[r] = {down(A.dim[r - 1].d), down(A.dim[r - 1].h), down(A.dim[r - 1].w)};
    auto buf = [&](float **p, int64_t count) { return N->alloc((void **)p, count * sizeof(float), set); };
    if (int rc = buf(&A.x, (int64_t)B * 4 * A.dim[0].v())) return rc;
    for (int l = 0; l < kNumLayers; l++)
        if (int rc = buf(&A.a[l], (int64_t)B * kLayers[l].cout * A.dim[layer_res(l)].v())) return rc;
    const int upc[3] = {64, 32, 24}, upr[3] = {2, 1, 0};
    for (int u = 0; u < 3; u++)
        if (int rc = buf(&A.up[u], (int64_t)B * upc[u] * A.dim[upr[u]].v())) return rc;
    if (grads) {
        if (int rc = buf(&A.gx, (int64_t)B * 4 * A.dim[0].v())) return rc;
        for (int l = 0; l < kNumLayers; l++)
            if (int rc = buf(&A.ga[l], (int64_t)B * kLayers[l].cout * A.dim[layer_res(l)].v())) return rc;
        for (int u = 0; u < 3; u++)
            if (int rc = buf(&A.gup[u], (int64_t)B * upc[u] * A.dim[upr[u]].v())) return rc;
    }
    return 0;
}

void conv(vfem_unet *N, Acts &A, int l, const float *in0, const float *in1)
{
    const LayerDef &d = kLayers[l];
    const Dims &di = A.dim[l == D1A ? 0 : l == D2A ? 1 : l == D3 ? 2 : layer_res(l)];
    const Dims &dout = A.dim[layer_res(l)];
    const int ncb = (d.cout + kCB - 1) / kCB;
    hipLaunchKernelGGL(k_conv_fwd, dim3(nblocks((int64_t)A.B * ncb * dout.v())), dim3(kT), 0, N->stream, in0, d.cin0,
                       in1, d.cin1, A.B, di.d, di.h, di.w, N->wf[l], d.cout, d.k, d.s, dout.d, dout.h, dout.w,
                       d.relu, A.a[l]);
}

void upsample(vfem_unet *N, Acts &A, const float *in, int C, int rin, float *out, int rout)
{
    const Dims &a = A.dim[rin], &b = A.dim[rout];
    hipLaunchKernelGGL(k_upsample, dim3(nblocks((int64_t)A.B * C * b.v())), dim3(kT), 0, N->stream, in, A.B * C, a.d,
                       a.h, a.w, out, b.d, b.h, b.w);
}

void forward(vfem_unet *N, Acts &A)
{
    conv(N, A, E0A, A.x, nullptr);
    conv(N, A, E0B, A.a[E0A], nullptr);
    conv(N, A, D1A, A.a[E0B], nullptr);
    conv(N, A, D1B, A.a[D1A], nullptr);
    conv(N, A, D2A, A.a[D1B], nullptr);
    conv(N, A, D2B, A.a[D2A], nullptr);
    conv(N, A, D3, A.a[D2B], nullptr);
    upsample(N, A, A.a[D3], 64, 3, A.up[0], 2);
    conv(N, A, U2, A.up[0], A.a[D2B]);
    upsample(N, A, A.a[U2], 32, 2, A.up[1], 1);
    conv(N, A, U1, A.up[1], A.a[D1B]);
    upsample(N, A, A.a[U1], 24, 1, A.up[2], 0);
    conv(N, A, U0, A.up[2], A.a[E0B]);
    conv(N, A, OUT, A.a[U0], nullptr);
}

void fill(vfem_unet *N, float *p, int64_t n)
{
    hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kT), 0, N->stream, n, p, 0.0f);
}

// layer l backward: ga[l] holds d loss / d output (post-activation); writes the
// weight gradient and accumulates the input gradients into gin0 / gin1
void conv_bwd(vfem_unet *N, Acts &A, int l, const float *in0, const float *in1, float *gin0, float *gin1)
{
    const LayerDef &d = kLayers[l];
    const Dims &di = A.dim[l == D1A ? 0 : l == D2A ? 1 : l == D3 ? 2 : layer_res(l)];
    const Dims &dout = A.dim[layer_res(l)];
    const int64_t nout = (int64_t)A.B * d.cout * dout.v();
    if (d.relu) hipLaunchKernelGGL(k_relu_bwd, dim3(nblocks(nout)), dim3(kT), 0, N->stream, nout, A.a[l], A.ga[l]);
    hipLaunchKernelGGL(k_conv_bwd_weight, dim3((d.cin0 + d.cin1) * d.cout), dim3(kT), 0, N->stream, A.ga[l], d.cout,
                       A.B, dout.d, dout.h, dout.w, in0, d.cin0, in1, d.cin1, di.d, di.h, di.w, d.k, d.s, N->g[l]);
    if (gin0 || gin1) {
        const int ncb = (d.cin0 + d.cin1 + kCB - 1) / kCB;
        hipLaunchKernelGGL(k_conv_bwd_data, dim3(nblocks((int64_t)A.B * ncb * di.v())), dim3(kT), 0, N->stream,
                           A.ga[l], d.cout, A.B, dout.d, dout.h, dout.w, N->wb[l], d.cin0, d.cin1, gin0, gin1, di.d,
                           di.h, di.w, d.k, d.s);
    }
}

void upsample_bwd(vfem_unet *N, Acts &A, const float *gout, int C, int rout, float *gin, int rin)
{
    const Dims &a = A.dim[rin], &b = A.dim[rout];
    hipLaunchKernelGGL(k_upsample_bwd, dim3(nblocks((int64_t)A.B * C * a.v())), dim3(kT), 0, N->stream, gout,
                       A.B * C, b.d, b.h, b.w, gin, a.d, a.h, a.w);
}

// ga[OUT] must hold d loss / d out; every other gradient buffer is rebuilt here
void backward(vfem_unet *N, Acts &A, bool input_grad)
{
    for (int l = 0; l < OUT; l++) fill(N, A.ga[l], (int64_t)A.B * kLayers[l].cout * A.dim[layer_res(l)].v());
    const int upc[3] = {64, 32, 24}, upr[3] = {2, 1, 0};
    for (int u = 0; u < 3; u++) fill(N, A.gup[u], (int64_t)A.B * upc[u] * A.dim[upr[u]].v());
    if (input_grad) fill(N, A.gx, (int64_t)A.B * 4 * A.dim[0].v());
    conv_bwd(N, A, OUT, A.a[U0], nullptr, A.ga[U0], nullptr);
    conv_bwd(N, A, U0, A.up[2], A.a[E0B], A.gup[2], A.ga[E0B]);
    upsample_bwd(N, A, A.gup[2], 24, 0, A.ga[U1], 1);
    conv_bwd(N, A, U1, A.up[1], A.a[D1B], A.gup[1], A.ga[D1B]);
    upsample_bwd(N, A, A.gup[1], 32, 1, A.ga[U2], 2);
    conv_bwd(N, A, U2, A.up[0], A.a[D2B], A.gup[0], A.ga[D2B]);
    upsample_bwd(N, A, A.gup[0], 64, 2, A.ga[D3], 3);
    conv_bwd(N, A, D3, A.a[D2B], nullptr, A.ga[D2B], nullptr);
    conv_bwd(N, A, D2B, A.a[D2A], nullptr, A.ga[D2A], nullptr);
    conv_bwd(N, A, D2A, A.a[D1B], nullptr, A.ga[D1B], nullptr);
    conv_bwd(N, A, D1B, A.a[D1A], nullptr, A.ga[D1A], nullptr);
    conv_bwd(N, A, D1A, A.a[E0B], nullptr, A.ga[E0B], nullptr);
    conv_bwd(N, A, E0B, A.a[E0A], nullptr, A.ga[E0A], nullptr);
    conv_bwd(N, A, E0A, A.x, nullptr, input_grad ? A.gx : nullptr, nullptr);
}

void pack_all(vfem_unet *N)
{
    for (int l = 0; l < kNumLayers; l++) {
        const LayerDef &d = kLayers[l];
        hipLaunchKernelGGL(k_pack, dim3(nblocks(N->n[l])), dim3(kT), 0, N->stream, N->n[l], N->w[l], d.cout,
                           d.cin0 + d.cin1, d.k * d.k * d.k, N->wf[l], N->wb[l]);
    }
}

}  // namespace

namespace vfem {

int unet_prepare(vfem_unet *net, int d, int h, int w)
{
    U_TRY(hipSetDevice(net->device));
    return prepare(net, 0, 1, d, h, w, false);
}

float *unet_input(vfem_unet *net) { return net->acts[0].x; }
const float *unet_output(vfem_unet *net) { return net->acts[0].a[OUT]; }

void unet_forward_enqueue(vfem_unet *net, hipStream_t stream)
{
    hipStream_t keep = net->stream;
    net->stream = stream;
    forward(net, net->acts[0]);
    net->stream = keep;
}

const char *unet_last_error() { return u_err.c_str(); }

}  // namespace vfem

// ===========================================================================
// C-ABI
// ===========================================================================

extern "C" {

int64_t vfem_unet_num_params(void)
{
    int64_t s = 0;
    for (int l = 0; l < kNumLayers; l++) s += layer_params(l);
    return s;
}

// host-to-device copy ordered on the network's stream and waited for: that
// stream is non-blocking, so a copy on the null stream would not be ordered
// before its kernels
static hipError_t to_device(hipStream_t s, void *dst, const void *src, size_t bytes)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

int vfem_unet_create(uint64_t seed, vfem_unet **out)
{
    if (!out) return VFEM_ERR_ARGUMENT;
    *out = nullptr;
    vfem_unet *N = new vfem_unet();
    auto fail = [&](int rc) {
        delete N;
        return rc;
    };
    if (hipGetDevice(&N->device) != hipSuccess) return fail(VFEM_ERR_DEVICE);
    if (hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking) != hipSuccess) return fail(VFEM_ERR_DEVICE);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kNumLayers; l++) {
        const LayerDef &d = kLayers[l];
        N->n[l] = layer_params(l);
        const size_t b4 = N->n[l] * sizeof(float), b2 = N->n[l] * sizeof(__bf16);
        if (N->alloc((void **)&N->w[l], b4) || N->alloc((void **)&N->g[l], b4) || N->alloc((void **)&N->m[l], b4) ||
            N->alloc((void **)&N->v[l], b4) || N->alloc((void **)&N->wf[l], b2) || N->alloc((void **)&N->wb[l], b2))
            return fail(VFEM_ERR_DEVICE);
        // He-normal initialisation
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / ((d.cin0 + d.cin1) * d.k * d.k * d.k)));
        std::vector<float> h(N->n[l]);
        for (auto &x : h) x = nd(rng);
        if (to_device(N->stream, N->w[l], h.data(), b4) != hipSuccess) return fail(VFEM_ERR_DEVICE);
        // on the network's own (non-blocking) stream, which a null-stream fill
        // would not order before the training kernels
        if (hipMemsetAsync(N->m[l], 0, b4, N->stream) != hipSuccess ||
            hipMemsetAsync(N->v[l], 0, b4, N->stream) != hipSuccess)
            return fail(VFEM_ERR_DEVICE);
    }
    if (N->alloc((void **)&N->loss, 4096 * sizeof(float))) return fail(VFEM_ERR_DEVICE);
    pack_all(N);
    if (hipStreamSynchronize(N->stream) != hipSuccess) return fail(VFEM_ERR_DEVICE);
    *out = N;
    return VFEM_OK;
}

void vfem_unet_destroy(vfem_unet *net) { delete net; }

int vfem_unet_get_weights(vfem_unet *net, float *w)
{
    if (!net || !w) return VFEM_ERR_ARGUMENT;
    U_TRY(hipStreamSynchronize(net->stream));  // the copies below run on the null stream
    for (int l = 0; l < kNumLayers; l++) {
        U_TRY(hipMemcpy(w, net->w[l], net->n[l] * sizeof(float), hipMemcpyDeviceToHost));
        w += net->n[l];
    }
    return VFEM_OK;
}

int vfem_unet_set_weights(vfem_unet *net, const float *w)
{
    if (!net || !w) return VFEM_ERR_ARGUMENT;
    for (int l = 0; l < kNumLayers; l++) {
        U_TRY(to_device(net->stream, net->w[l], w, net->n[l] * sizeof(float)));
        w += net->n[l];
    }
    pack_all(net);
    U_TRY(hipStreamSynchronize(net->stream));
    return VFEM_OK;
}

int vfem_unet_forward(vfem_unet *net, const float *x, int32_t batch, int32_t d, int32_t h, int32_t w, float *out)
{
    if (!net || !x || !out || batch < 1 || d < 2 || h < 2 || w < 2) return VFEM_ERR_ARGUMENT;
    if (int rc = prepare(net, 2, batch, d, h, w, true)) return rc;
    Acts &A = net->acts[2];
    const int64_t nin = (int64_t)batch * 4 * A.dim[0].v(), nout = (int64_t)batch * 3 * A.dim[0].v();
    U_TRY(to_device(net->stream, A.x, x, nin * sizeof(float)));
    forward(net, A);
    U_TRY(hipStreamSynchronize(net->stream));
    U_TRY(hipMemcpy(out, A.a[OUT], nout * sizeof(float), hipMemcpyDeviceToHost));
    return VFEM_OK;
}

int vfem_unet_backward(vfem_unet *net, const float *gout, float *gw, float *gx)
{
    if (!net || !gout || !gw) return VFEM_ERR_ARGUMENT;
    Acts &A = net->acts[2];
    if (!A.grads) return VFEM_ERR_ARGUMENT;
    const int64_t nout = (int64_t)A.B * 3 * A.dim[0].v(), nin = (int64_t)A.B * 4 * A.dim[0].v();
    U_TRY(to_device(net->stream, A.ga[OUT], gout, nout * sizeof(float)));
    backward(net, A, gx != nullptr);
    U_TRY(hipStreamSynchronize(net->stream));
    for (int l = 0; l < kNumLayers; l++) {
        U_TRY(hipMemcpy(gw, net->g[l], net->n[l] * sizeof(float), hipMemcpyDeviceToHost));
        gw += net->n[l];
    }
    if (gx) U_TRY(hipMemcpy(gx, A.gx, nin * sizeof(float), hipMemcpyDeviceToHost));
    return VFEM_OK;
}

int vfem_unet_train(vfem_unet *net, const float *x, const float *y, const float *mask, int64_t n_samples, int32_t d,
                    int32_t h, int32_t w, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss)
{
    if (!net || !x || !y || !mask || n_samples < batch || batch < 1 || batch > 4096 || epochs < 1) {
        u_err = "invalid argument";
        return VFEM_ERR_ARGUMENT;
    }
    if (int rc = prepare(net, 1, batch, d, h, w, true)) return rc;
    Acts &A = net->acts[1];
    const int64_t V = A.dim[0].v();
    float *dx = nullptr, *dy = nullptr, *dm = nullptr, *by = nullptr, *bm = nullptr;
    int32_t *didx = nullptr;
    std::vector<void *> tmp;
    auto talloc = [&](void **p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        tmp.push_back(*p);
        return true;
    };
    auto cleanup = [&]() {
        for (void *p : tmp) (void)hipFree(p);
    };
    if (!talloc((void **)&dx, n_samples * 4 * V * sizeof(float)) || !talloc((void **)&dy, n_samples * 3 * V * sizeof(float)) ||
        !talloc((void **)&dm, n_samples * V * sizeof(float)) || !talloc((void **)&by, batch * 3 * V * sizeof(float)) ||
        !talloc((void **)&bm, batch * V * sizeof(float)) || !talloc((void **)&didx, n_samples * sizeof(int32_t))) {
        cleanup();
        u_err = "out of device memory";
        return VFEM_ERR_DEVICE;
    }
    if (to_device(net->stream, dx, x, n_samples * 4 * V * sizeof(float)) != hipSuccess ||
        to_device(net->stream, dy, y, n_samples * 3 * V * sizeof(float)) != hipSuccess ||
        to_device(net->stream, dm, mask, n_samples * V * sizeof(float)) != hipSuccess) {
        cleanup();
        return VFEM_ERR_DEVICE;
    }
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(n_samples);
    const int64_t steps = n_samples / batch;
    std::vector<float> lossh(batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
        for (int64_t i = 0; i < n_samples; i++) perm[i] = (int32_t)i;
        std::shuffle(perm.begin(), perm.end(), rng);
        if (to_device(net->stream, didx, perm.data(), n_samples * sizeof(int32_t)) != hipSuccess) {
            cleanup();
            return VFEM_ERR_DEVICE;
        }
        double esum = 0;
        for (int64_t s = 0; s < steps; s++) {
            const int32_t *bi = didx + s * batch;
            hipLaunchKernelGGL(k_gather, dim3(nblocks(4 * V * batch)), dim3(kT), 0, net->stream, dx, bi, 4 * V, batch,
                               A.x);
            hipLaunchKernelGGL(k_gather, dim3(nblocks(3 * V * batch)), dim3(kT), 0, net->stream, dy, bi, 3 * V, batch,
                               by);
            hipLaunchKernelGGL(k_gather, dim3(nblocks(V * batch)), dim3(kT), 0, net->stream, dm, bi, V, batch, bm);
            forward(net, A);
            hipLaunchKernelGGL(k_mse, dim3(batch), dim3(kT), 0, net->stream, A.a[OUT], by, bm, V, batch, A.ga[OUT],
                               net->loss);
            backward(net, A, false);
            net->step++;
            // cosine decay of the learning rate over the run
            const double prog = (double)(e * steps + s) / (double)(epochs * steps);
            const float lr_t = (float)(lr * 0.5 * (1.0 + std::cos(3.141592653589793 * prog)));
            const float c1 = 1.0f / (1.0f - std::pow(b1, (float)net->step));
            const float c2 = 1.0f / (1.0f - std::pow(b2, (float)net->step));
            for (int l = 0; l < kNumLayers; l++) {
                const LayerDef &d = kLayers[l];
                hipLaunchKernelGGL(k_adam, dim3(nblocks(net->n[l])), dim3(kT), 0, net->stream, net->n[l], net->w[l],
                                   net->g[l], net->m[l], net->v[l], lr_t, b1, b2, eps, c1, c2, d.cout,
                                   d.cin0 + d.cin1, d.k * d.k * d.k, net->wf[l], net->wb[l]);
            }
            if (hipMemcpyAsync(lossh.data(), net->loss, batch * sizeof(float), hipMemcpyDeviceToHost, net->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(net->stream) != hipSuccess) {
                cleanup();
                u_err = "training step failed";
                return VFEM_ERR_DEVICE;
            }
            for (float q : lossh) esum += q;
        }
        if (epoch_loss) epoch_loss[e] = (float)(esum / (double)(steps * batch));
    }
    cleanup();
    if (hipGetLastError() != hipSuccess) return VFEM_ERR_DEVICE;
    return VFEM_OK;
}

}  // extern "C"
