"""CPU tests pinning the oracle's conv net (the checker of the device's BF16
MFMA kernels): the periodic 3-D convolution against numpy shifts, the
gradient of every parameter block against central finite differences, and
the bfloat16 rounding against its definition."""
import numpy as np
import pytest

import oracle_lib as orc


def numpy_conv(x, w, b):
    out = np.zeros(x.shape[:3] + (w.shape[2],)) + b
    for t in range(27):
        d = (t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1)
        # value at (p + d): roll by -d
        xs = np.roll(x, shift=(-d[0], -d[1], -d[2]), axis=(0, 1, 2))
        out += np.einsum("xyzi,io->xyzo", xs, w[t])
    return out


def random_params(rng, scale=0.2):
    return rng.normal(0, scale, orc.NET_NPARAM)


def test_conv3d_matches_numpy():
    rng = np.random.default_rng(1)
    for n, cin, cout in ((5, 8, 16), (4, 16, 3), (6, 16, 16)):
        x = rng.standard_normal((n, n, n, cin))
        w = rng.standard_normal((27, cin, cout))
        b = rng.standard_normal(cout)
        np.testing.assert_allclose(orc.conv3d(x, w, b), numpy_conv(x, w, b), rtol=1e-12, atol=1e-11)


def test_forward_matches_numpy_composition():
    rng = np.random.default_rng(2)
    n = 5
    p = random_params(rng)
    x = rng.standard_normal((n, n, n, 8))
    off = 0
    h = x
    for layer, (ci, co) in enumerate(zip(orc.NET_CIN, orc.NET_COUT)):
        w = p[off: off + 27 * ci * co].reshape(27, ci, co)
        off += 27 * ci * co
        b = p[off: off + co]
        off += co
        h = numpy_conv(h, w, b)
        if layer < 3:
            h = np.maximum(h, 0)
    y, _ = orc.net_forward(p, x)
    np.testing.assert_allclose(y, h, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n = 4
    p = random_params(rng, 0.3)
    x = rng.standard_normal((n, n, n, 8))
    target = rng.standard_normal((n, n, n, 3))
    loss, g = orc.net_gradient(p, x, target)
    assert loss == pytest.approx(orc.net_forward(p, x, target)[1], rel=1e-14)
    # a few parameters from every block (weights and biases of all 4 layers)
    idx = []
    off = 0
    for ci, co in zip(orc.NET_CIN, orc.NET_COUT):
        nw = 27 * ci * co
        idx += list(rng.choice(np.arange(off, off + nw), 6, replace=False)) + [off + nw, off + nw + co - 1]
        off += nw + co
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (orc.net_forward(pp, x, target)[1] - orc.net_forward(pm, x, target)[1]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-10), i


def test_bf16_rounding():
    for v, r in ((1.0, 1.0), (1.0 + 2 ** -8, 1.0), (1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7), (-2.5, -2.5),
                 (1.0 + 2 ** -7, 1.0 + 2 ** -7), (3.14159, 3.140625)):
        assert orc.lib().orc_bf16_round(v) == r, v
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1 (even)
    assert orc.lib().orc_bf16_round(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
