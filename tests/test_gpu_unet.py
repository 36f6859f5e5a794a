"""GPU checks of the U-Net kernels (BF16 MFMA forward, data and weight gradients) against
the NumPy oracle with the same rounding points, and of training."""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu

GRID = (9, 5, 5)
BASE = 8


def _setup(seed=0, batch=2):
    from platoanalyze_amd.unet import UNet

    net = UNet(GRID, BASE, seed=seed)
    rng = np.random.default_rng(seed + 10)
    V = int(np.prod(GRID))
    x = rng.normal(size=(batch, V, 8)).astype(np.float32)
    y = rng.normal(size=(batch, V, 8)).astype(np.float32) * 0.1
    return net, x, y


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gpu_unet_forward_matches_oracle():
    print("case: unet forward 9x5x5 base 8 batch 2", flush=True)
    net, x, y = _setup()
    p = net.params().cpu().numpy().astype(np.float64)
    ref = U.UNetRef(GRID, BASE)
    out_ref, _ = ref.forward(x.astype(np.float64), p)
    xt = torch.from_numpy(x).reshape(2, *GRID, 8).cuda()
    out = net.forward(xt).reshape(2, -1, 8).cpu().numpy()
    err = _rel(out, out_ref)
    print(f"forward rel L2 error {err:.2e}", flush=True)
    assert err < 1e-2


def test_gpu_unet_gradient_matches_oracle():
    print("case: unet gradient 9x5x5 base 8 batch 2", flush=True)
    net, x, y = _setup(seed=1)
    p = net.params().cpu().numpy().astype(np.float64)
    ref = U.UNetRef(GRID, BASE)
    loss_ref, g_ref, _ = ref.loss_grad(x.astype(np.float64), y.astype(np.float64), p)
    xt = torch.from_numpy(x).reshape(2, *GRID, 8).cuda()
    yt = torch.from_numpy(y).reshape(2, *GRID, 8).cuda()
    loss, g = net.gradient(xt, yt)
    g = g.cpu().numpy()
    err = _rel(g, g_ref)
    print(f"loss {loss:.6e} ref {loss_ref:.6e}; gradient rel L2 error {err:.2e}", flush=True)
    assert abs(loss - loss_ref) <= 1e-2 * abs(loss_ref)
    assert err < 3e-2
    # per layer, so a wrong layer cannot hide behind the large ones
    offs = np.cumsum([0] + [27 * l[2] * l[5] + l[5] for l in ref.L])
    for k in range(len(ref.L)):
        a, b = g[offs[k]:offs[k + 1]], g_ref[offs[k]:offs[k + 1]]
        if np.linalg.norm(b) > 0:
            assert _rel(a, b) < 5e-2, (k, _rel(a, b))


def test_gpu_unet_training_reduces_loss():
    print("case: unet overfit one batch", flush=True)
    net, x, y = _setup(seed=2)
    xt = torch.from_numpy(x).reshape(2, *GRID, 8).cuda()
    yt = torch.from_numpy(y).reshape(2, *GRID, 8).cuda()
    l0 = net.train_step(xt, yt, 1e-3)
    for _ in range(200):
        l1 = net.train_step(xt, yt, 1e-3)
    print(f"loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < 0.25 * l0
