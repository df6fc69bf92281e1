"""CPU checks of the U-Net oracle (oracle/unet_oracle.py): its analytic
backward pass against central finite differences, and the adjoint identities
of its convolution and upsampling (so a transposed index in either fails)."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_param_count():
    assert U.num_params() == 274800


def test_conv_backward_is_the_adjoint():
    rng = np.random.default_rng(0)
    for stride, k in ((1, 3), (2, 3), (1, 1)):
        x = rng.standard_normal((2, 3, 7, 5, 9))
        w = rng.standard_normal((4, 3, k, k, k))
        y = U.conv3d(x, w, stride)
        gy = rng.standard_normal(y.shape)
        gx, gw = U.conv3d_backward(x, w, stride, gy)
        # <conv(x), gy> is linear in x and in w
        assert np.isclose((y * gy).sum(), (gx * x).sum(), rtol=1e-12)
        assert np.isclose((y * gy).sum(), (gw * w).sum(), rtol=1e-12)
    # a stride-2 3^3 convolution maps n nodes to (n - 1) // 2 + 1
    assert [U.conv_out_size(n, 3, 2) for n in (9, 5, 3, 2, 33, 10)] == [5, 3, 2, 1, 17, 5]


def test_upsample_is_node_aligned_interpolation():
    rng = np.random.default_rng(1)
    c = rng.standard_normal((1, 2, 3, 5, 2))
    f = U.upsample(c, (5, 9, 3))
    assert np.allclose(f[:, :, ::2, ::2, ::2][:, :, :3, :5, :2], c)       # coarse nodes reproduced
    assert np.allclose(f[:, :, 1, 0, 0], 0.5 * (c[:, :, 0, 0, 0] + c[:, :, 1, 0, 0]))
    lin = np.fromfunction(lambda b, ch, z, y, x: 1 + 2 * z - y + 0.5 * x, (1, 1, 3, 5, 2))
    assert np.allclose(U.upsample(lin, (5, 9, 3))[0, 0, :, :, :3],
                       np.fromfunction(lambda z, y, x: 1 + z - 0.5 * y + 0.25 * x, (5, 9, 3)))
    g = rng.standard_normal(f.shape)
    assert np.isclose((f * g).sum(), (c * U.upsample_backward(g, c.shape[2:])).sum(), rtol=1e-12)


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    w = U.init_weights(1)
    x = rng.standard_normal((2, 4, 9, 9, 9))
    t = rng.standard_normal((2, 3, 9, 9, 9))
    mask = (rng.random((2, 9, 9, 9)) > 0.3).astype(float)
    out, act = U.forward(x, w, keep=True)
    _, g = U.masked_mse(out, t, mask)
    grads, gx = U.backward(act, w, g)

    def loss(w2, x2):
        return U.masked_mse(U.forward(x2, w2), t, mask)[0]

    eps = 1e-6
    for name in ("e0a", "d1b", "d3", "u2", "u0", "out"):
        idx = tuple(int(rng.integers(0, s)) for s in w[name].shape)
        wp = {k: v.copy() for k, v in w.items()}
        wp[name][idx] += eps
        wm = {k: v.copy() for k, v in w.items()}
        wm[name][idx] -= eps
        fd = (loss(wp, x) - loss(wm, x)) / (2 * eps)
        assert fd == pytest.approx(grads[name][idx], rel=1e-5, abs=1e-9), name
    idx = (1, 2, 4, 4, 5)
    xp, xm = x.copy(), x.copy()
    xp[idx] += eps
    xm[idx] -= eps
    assert (loss(w, xp) - loss(w, xm)) / (2 * eps) == pytest.approx(gx[idx], rel=1e-5, abs=1e-9)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 1e-30])
    r = U.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7      # ties to even
    assert abs(r[3] + 3.140625) < 1e-12


def test_conv_and_upsample_match_torch():
    """The oracle's convolution against torch.nn.functional.conv3d (zero
    padding 1, stride 1 and 2; 1x1x1 head), and its upsampling against
    trilinear interpolation with aligned corners where the fine extent is
    2m - 1 (node-aligned by construction)."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(4)
    for stride, k, shape in ((1, 3, (2, 5, 9, 7, 11)), (2, 3, (2, 5, 9, 7, 10)), (2, 3, (1, 3, 5, 4, 3)),
                             (1, 1, (2, 16, 6, 5, 4))):
        x = rng.standard_normal(shape)
        w = rng.standard_normal((4, shape[1], k, k, k))
        ref = F.conv3d(torch.from_numpy(x), torch.from_numpy(w), stride=stride, padding=k // 2).numpy()
        y = U.conv3d(x, w, stride)
        assert y.shape == ref.shape
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    c = rng.standard_normal((2, 3, 4, 5, 3))
    size = (7, 9, 5)
    ref = F.interpolate(torch.from_numpy(c), size=size, mode="trilinear", align_corners=True).numpy()
    assert np.abs(U.upsample(c, size) - ref).max() <= 1e-13
