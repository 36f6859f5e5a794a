"""CPU checks pinning the U-Net oracle (oracle/unet_ref.py) before it judges the GPU
kernels: BF16 rounding against torch's bfloat16 cast, the stride-1 convolution against
scipy.ndimage.correlate, the stride-2 and transposed links against their index
relations by brute force, and the backward pass against central differences."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from oracle import unet_ref as U


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, np.inf, -np.inf], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(U.bf16(x), ref)


def test_stride1_conv_matches_scipy_correlate():
    rng = np.random.default_rng(2)
    g = (5, 4, 6)
    x = rng.normal(size=g)
    w = rng.normal(size=(3, 3, 3))
    lk = U.links(g, g, 1, 0)
    y = np.zeros(int(np.prod(g)))
    for t, (ov, iv) in enumerate(lk):
        y[ov] += x.ravel()[iv] * w.ravel()[t]
    ref = ndimage.correlate(x, w, mode="constant", cval=0.0)
    np.testing.assert_allclose(y.reshape(g), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("transposed", [0, 1])
def test_stride2_links_brute_force(transposed):
    gi, go = ((9, 5, 5), (5, 3, 3)) if not transposed else ((5, 3, 3), (9, 5, 5))
    lk = U.links(gi, go, 2, transposed)
    for t, (ov, iv) in enumerate(lk):
        d = np.array([t // 9, (t // 3) % 3, t % 3])
        want = set()
        for o in np.ndindex(*go):
            o = np.array(o)
            for i in np.ndindex(*gi):
                i = np.array(i)
                rel = (i == 2 * o + d - 1) if not transposed else (o == 2 * i + d - 1)
                if rel.all():
                    want.add((np.ravel_multi_index(o, go), np.ravel_multi_index(i, gi)))
        assert set(zip(ov.tolist(), iv.tolist())) == want


def test_backward_matches_central_differences():
    r = U.UNetRef((9, 5, 5), 8, exact=True)
    n = U.num_params(8)
    rng = np.random.default_rng(3)
    p = np.zeros(n)
    for (W, b), l in zip(r.split(np.arange(n)), r.L):  # He-normal weights, small biases
        p[W.ravel()] = rng.normal(size=W.size) * np.sqrt(2.0 / (27 * l[2]))
        p[b] = rng.normal(size=b.size) * 0.05
    x = rng.normal(size=(2, 9 * 5 * 5, U.IN_CH))
    y = rng.normal(size=(2, 9 * 5 * 5, U.OUT_CH))
    loss, g, _ = r.loss_grad(x, y, p)
    offs = np.cumsum([0] + [27 * l[2] * l[5] + l[5] for l in r.L])
    checked = 0
    for k in range(len(r.L)):  # a few parameters of every layer
        for i in rng.choice(np.arange(offs[k], offs[k + 1]), 3, replace=False):
            h = 1e-5
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (r.loss_grad(x, y, pp)[0] - r.loss_grad(x, y, pm)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max() + 1e-5 * abs(fd), (k, i, fd, g[i])
            checked += 1
    assert checked == 3 * len(r.L)


def test_forward_flops_counts_every_oracle_link_slot():
    """UNet.forward_flops (the FLOP count bench.py reports) equals 2 x 27 x cin x cout x
    output voxels summed over the oracle's layer table and grid levels."""
    from platoanalyze_amd.unet import UNet, conv_layers

    grid, C, batch = (65, 33, 33), 16, 4
    r = U.UNetRef(grid, C)
    want = 0
    for (bi, _, cin, bo, _, cout, *_rest) in r.L:
        go = r.g[U.BUF_LEVEL[bo]] if bo >= 0 else r.g[0]
        want += 2 * batch * int(np.prod(go)) * cout * 27 * cin
    assert [(l[2], l[5], U.BUF_LEVEL[l[0]], U.BUF_LEVEL[l[3]] if l[3] >= 0 else 0) for l in r.L] == conv_layers(C)
    net = UNet.__new__(UNet)  # no device handle: only the geometry is used
    net.grid, net.base = grid, C
    assert net.forward_flops(batch) == want
