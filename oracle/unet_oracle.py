"""CPU oracle of the learned coarse-grid correction (test infrastructure only:
imported by tests/ to check the HIP kernels, never by the product).

Float64 numpy restatement of the 3-D conv U-Net the library trains and runs
(vulkanfem_amd/csrc/vfem_unet.hip).  It has no counterpart in the reference,
which only has the multigrid coarse correction (include/MultigridPreconditioner.hpp:158-193);
the network stands in for that correction at one level of the damped cycle.

Tensors are [B][C][D][H][W] (W = x fastest).  Layers:

    name  inputs            Cin -> Cout  kernel stride  ReLU
    e0a   x                   4 -> 16      3      1      yes
    e0b   e0a                16 -> 16      3      1      yes     (skip 0)
    d1a   e0b                16 -> 24      3      2      yes
    d1b   d1a                24 -> 24      3      1      yes     (skip 1)
    d2a   d1b                24 -> 32      3      2      yes
    d2b   d2a                32 -> 32      3      1      yes     (skip 2)
    d3    d2b                32 -> 64      3      2      yes     (bottleneck)
    u2    up(d3) | d2b    64+32 -> 32      3      1      yes
    u1    up(u2) | d1b    32+24 -> 24      3      1      yes
    u0    up(u1) | e0b    24+16 -> 16      3      1      yes
    out   u0                 16 -> 3       1      1      no

3x3x3 convolutions use zero padding 1; a stride-2 convolution maps n nodes to
(n - 1) // 2 + 1.  up() is node-aligned trilinear interpolation to the skip
tensor's size: fine node j takes coarse node j/2 when j is even and the mean of
coarse nodes (j-1)/2 and min((j+1)/2, m-1) when j is odd, per axis.  No biases:
the network is positively homogeneous, N(a r) = a N(r) for a > 0.
"""
import numpy as np

LAYERS = [  # name, inputs, cin, cout, k, stride, relu
    ("e0a", ("x",), 4, 16, 3, 1, True),
    ("e0b", ("e0a",), 16, 16, 3, 1, True),
    ("d1a", ("e0b",), 16, 24, 3, 2, True),
    ("d1b", ("d1a",), 24, 24, 3, 1, True),
    ("d2a", ("d1b",), 24, 32, 3, 2, True),
    ("d2b", ("d2a",), 32, 32, 3, 1, True),
    ("d3", ("d2b",), 32, 64, 3, 2, True),
    ("u2", ("up:d3", "d2b"), 96, 32, 3, 1, True),
    ("u1", ("up:u2", "d1b"), 56, 24, 3, 1, True),
    ("u0", ("up:u1", "e0b"), 40, 16, 3, 1, True),
    ("out", ("u0",), 16, 3, 1, 1, False),
]


def num_params():
    return sum(ci * co * k ** 3 for _, _, ci, co, k, _, _ in LAYERS)


def init_weights(seed=0):
    """He-normal weights [cout][cin][k][k][k] (the library draws its own; this is
    for the oracle's self-tests)."""
    rng = np.random.default_rng(seed)
    return {n: rng.standard_normal((co, ci, k, k, k)) * np.sqrt(2.0 / (ci * k ** 3))
            for n, _, ci, co, k, _, _ in LAYERS}


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv_out_size(n, k, s):
    p = k // 2
    return (n + 2 * p - k) // s + 1


def conv3d(x, w, stride):
    """y[b, o, z, y, x] = sum_{c, t} w[o, c, t] * x[b, c, s*z + tz - p, ...] (zero padding)."""
    B, C, D, H, W = x.shape
    O, _, k, _, _ = w.shape
    p = k // 2
    Do, Ho, Wo = (conv_out_size(n, k, stride) for n in (D, H, W))
    xp = np.zeros((B, C, D + 2 * p + stride, H + 2 * p + stride, W + 2 * p + stride))
    xp[:, :, p:p + D, p:p + H, p:p + W] = x
    y = np.zeros((B, O, Do, Ho, Wo))
    for tz in range(k):
        for ty in range(k):
            for tx in range(k):
                patch = xp[:, :, tz:tz + stride * Do:stride, ty:ty + stride * Ho:stride, tx:tx + stride * Wo:stride]
                y += np.einsum("bczyx,oc->bozyx", patch, w[:, :, tz, ty, tx])
    return y


def conv3d_backward(x, w, stride, gy):
    """Gradients of conv3d w.r.t. x and w for upstream gradient gy."""
    B, C, D, H, W = x.shape
    O, _, k, _, _ = w.shape
    p = k // 2
    Do, Ho, Wo = gy.shape[2:]
    xp = np.zeros((B, C, D + 2 * p + stride, H + 2 * p + stride, W + 2 * p + stride))
    xp[:, :, p:p + D, p:p + H, p:p + W] = x
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    for tz in range(k):
        for ty in range(k):
            for tx in range(k):
                sl = (slice(None), slice(None), slice(tz, tz + stride * Do, stride), slice(ty, ty + stride * Ho, stride),
                      slice(tx, tx + stride * Wo, stride))
                gw[:, :, tz, ty, tx] = np.einsum("bozyx,bczyx->oc", gy, xp[sl])
                gxp[sl] += np.einsum("bozyx,oc->bczyx", gy, w[:, :, tz, ty, tx])
    return gxp[:, :, p:p + D, p:p + H, p:p + W], gw


def _up_matrix(m, n):
    """n x m node-aligned linear interpolation matrix."""
    U = np.zeros((n, m))
    for j in range(n):
        i0 = j // 2
        if j % 2 == 0:
            U[j, min(i0, m - 1)] += 1.0
        else:
            U[j, min(i0, m - 1)] += 0.5
            U[j, min(i0 + 1, m - 1)] += 0.5
    return U


def upsample(x, size):
    Uz, Uy, Ux = (_up_matrix(m, n) for m, n in zip(x.shape[2:], size))
    return np.einsum("Zz,Yy,Xx,bczyx->bcZYX", Uz, Uy, Ux, x)


def upsample_backward(gy, in_size):
    Uz, Uy, Ux = (_up_matrix(m, n) for m, n in zip(in_size, gy.shape[2:]))
    return np.einsum("Zz,Yy,Xx,bcZYX->bczyx", Uz, Uy, Ux, gy)


def forward(x, weights, keep=False):
    """Network output (and the activations when keep)."""
    act = {"x": x}
    for name, ins, ci, co, k, s, relu in LAYERS:
        parts = []
        for src in ins:
            if src.startswith("up:"):
                parts.append(upsample(act[src[3:]], act[ins[1]].shape[2:]))
            else:
                parts.append(act[src])
        inp = np.concatenate(parts, axis=1) if len(parts) > 1 else parts[0]
        act[name + ":in"] = inp
        y = conv3d(inp, weights[name], s)
        act[name] = np.maximum(y, 0.0) if relu else y
    return (act["out"], act) if keep else act["out"]


def backward(act, weights, gout):
    """Weight gradients and the input gradient for d loss / d out = gout."""
    grads = {}
    g = {"out": gout}
    for name, ins, ci, co, k, s, relu in reversed(LAYERS):
        gy = g.pop(name)
        if relu:
            gy = gy * (act[name] > 0)
        gin, grads[name] = conv3d_backward(act[name + ":in"], weights[name], s, gy)
        c0 = 0
        for src in ins:
            if src.startswith("up:"):
                base = src[3:]
                m = act[base].shape[1]
                gpart = upsample_backward(gin[:, c0:c0 + m], act[base].shape[2:])
            else:
                base = src
                m = act[base].shape[1]
                gpart = gin[:, c0:c0 + m]
            g[base] = g.get(base, 0.0) + gpart
            c0 += m
    return grads, g["x"]


def masked_mse(out, target, mask):
    """Loss and its gradient: sum over samples of |mask (out - t)|^2 / |mask t|^2, / B."""
    B = out.shape[0]
    m = mask[:, None]
    d = (out - target) * m
    den = np.maximum((target * m).reshape(B, -1).__pow__(2).sum(axis=1), 1e-30)
    loss = float(((d.reshape(B, -1) ** 2).sum(axis=1) / den).sum() / B)
    grad = 2.0 * d / den[:, None, None, None, None] / B
    return loss, grad
