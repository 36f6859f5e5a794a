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
    x = rng.normal(size=(batch, V, U.IN_CH)).astype(np.float32)
    y = rng.normal(size=(batch, V, U.OUT_CH)).astype(np.float32) * 0.1
    return net, x, y


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gpu_unet_forward_matches_oracle():
    print("case: unet forward 9x5x5 base 8 batch 2", flush=True)
    net, x, y = _setup()
    p = net.params().cpu().numpy().astype(np.float64)
    ref = U.UNetRef(GRID, BASE)
    out_ref, _ = ref.forward(x.astype(np.float64), p)
    xt = torch.from_numpy(x).reshape(2, *GRID, U.IN_CH).cuda()
    out = net.forward(xt).reshape(2, -1, U.OUT_CH).cpu().numpy()
    err = _rel(out, out_ref)
    print(f"forward rel L2 error {err:.2e}", flush=True)
    # measured 1.2e-4 to 2.8e-4 (BF16 operands rounded at the same points as the oracle;
    # the FP32 accumulation order of the MFMA tiles and of the split reductions moves a
    # few activations across a BF16 rounding boundary, one BF16 ulp = 3.9e-3 relative)
    assert err < 2e-3


def test_gpu_unet_gradient_matches_oracle():
    print("case: unet gradient 9x5x5 base 8 batch 2", flush=True)
    net, x, y = _setup(seed=1)
    p = net.params().cpu().numpy().astype(np.float64)
    ref = U.UNetRef(GRID, BASE)
    loss_ref, g_ref, _ = ref.loss_grad(x.astype(np.float64), y.astype(np.float64), p)
    xt = torch.from_numpy(x).reshape(2, *GRID, U.IN_CH).cuda()
    yt = torch.from_numpy(y).reshape(2, *GRID, U.OUT_CH).cuda()
    loss, g = net.gradient(xt, yt)
    g = g.cpu().numpy()
    err = _rel(g, g_ref)
    print(f"loss {loss:.6e} ref {loss_ref:.6e}; gradient rel L2 error {err:.2e}", flush=True)
    # measured: loss to 6e-7, gradient 6.1e-8 to 2.4e-4 (the same BF16 rounding flips of
    # activations as in the forward check)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    assert err < 2e-3
    # per layer, so a wrong layer cannot hide behind the large ones
    offs = np.cumsum([0] + [27 * l[2] * l[5] + l[5] for l in ref.L])
    per_layer = [_rel(g[offs[k]:offs[k + 1]], g_ref[offs[k]:offs[k + 1]]) for k in range(len(ref.L))
                 if np.linalg.norm(g_ref[offs[k]:offs[k + 1]]) > 0]
    print(f"largest per-layer gradient rel L2 error {max(per_layer):.2e}", flush=True)
    assert max(per_layer) < 1e-2


def test_gpu_unet_training_reduces_loss():
    print("case: unet overfit one batch", flush=True)
    net, x, y = _setup(seed=2)
    xt = torch.from_numpy(x).reshape(2, *GRID, U.IN_CH).cuda()
    yt = torch.from_numpy(y).reshape(2, *GRID, U.OUT_CH).cuda()
    l0 = net.train_step(xt, yt, 1e-3)
    for _ in range(200):
        l1 = net.train_step(xt, yt, 1e-3)
    print(f"loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < 0.25 * l0


def test_gpu_warm_start_end_to_end():
    """Train a small U-Net on SIMP OC trajectories of a 16x8x8 box (library solves and
    kernels end to end), then solve held-out designs from zero and from the network's
    warm start.  Both must reach the same solution on the true FP64 residual; the
    iteration counts of both are printed: that difference is what the network gains.
    Deflating FP64 block-Jacobi PCG by the network's guess must not cost iterations
    (deterministic counts) and must reach the same solution."""
    from platoanalyze_amd import workloads as wl
    from platoanalyze_amd.elliptic import Problem
    from platoanalyze_amd.warmstart import WarmStart, train

    print("case: warm start 16x8x8, 4 trajectories x 6 steps", flush=True)
    base = wl.make("cantilever_64x32x32", tiny=(16, 8, 8))
    net, c0, hist, held = train(base, 4, 6, epochs=20, batch=4, lr=2e-3, base_channels=8, seed=3,
                                log=lambda s: None)
    assert hist[-1] < hist[0]  # the library's own kernels trained the network
    cold, warm, plain, defl = [], [], [], []
    for w, z, f in held[::2]:
        m = wl.model(w)
        pb = Problem(m, tolerance=1e-9, check_every=8, precision="mixed")
        u_c = pb.solution(z).clone()
        cold.append(pb.stats.iterations)
        ws = WarmStart(net, w, m, c0)
        u0 = ws.initial_guess(z, f)
        u_w = pb.solution(z, u0=u0)
        warm.append(pb.stats.iterations)
        assert pb.stats.converged and pb.stats.rel_residual <= 1e-9
        assert float((u_w - u_c).norm() / u_c.norm()) < 1e-6
        pj = Problem(m, tolerance=1e-9, check_every=1)
        pj.solution(z)
        plain.append(pj.stats.iterations)
        u_d = pj.solution_deflated(z, ws.deflation_space(z, f))
        defl.append(pj.stats.iterations)
        assert pj.stats.converged and float((u_d - u_c).norm() / u_c.norm()) < 1e-6
    print(f"loss {hist[0]:.3e} -> {hist[-1]:.3e}; iterations cold {cold} warm {warm}; "
          f"FP64 Jacobi plain {plain} deflated by the guess {defl}", flush=True)
    assert all(d <= p for d, p in zip(defl, plain))


def test_gpu_warm_start_update_mode():
    """Update mode on SIMP OC trajectories of a 16x8x8 box: from the previous design and
    its solution the network predicts the displacement change and the guess takes the
    energy-optimal step along it.  Checked: training lowers the loss; every guess has
    at most the previous solution's energy (exact: the step size minimises the energy
    along the prediction, zero allowed); the warm-started solves reach the cold solution
    in fewer iterations in total.  Iterations from zero, from the previous solution and
    from the network's guess are printed."""
    from platoanalyze_amd import workloads as wl
    from platoanalyze_amd.elliptic import Problem
    from platoanalyze_amd.warmstart import WarmStart, energy, train

    print("case: update-mode warm start 16x8x8, 4 trajectories x 8 steps", flush=True)
    base = wl.make("cantilever_64x32x32", tiny=(16, 8, 8))
    net, c0, hist, held = train(base, 4, 8, epochs=20, batch=4, lr=2e-3, base_channels=8, seed=3, mode="update",
                                log=lambda s: None)
    assert hist[-1] < hist[0]
    cold, prev, warm = [], [], []
    for w, zp, up, z, f in held[::2]:
        m = wl.model(w)
        pb = Problem(m, tolerance=1e-9, check_every=8, precision="mixed")
        u_c = pb.solution(z).clone()
        cold.append(pb.stats.iterations)
        pb.solution(z, u0=up)
        prev.append(pb.stats.iterations)
        u0 = WarmStart(net, w, m, c0).initial_guess(z, f, zp, up)
        e_star, e_prev, e_warm = energy(m, z, u_c, f), energy(m, z, up, f), energy(m, z, u0, f)
        assert e_warm - e_star <= (e_prev - e_star) * (1 + 1e-9) + 1e-14 * abs(e_star)
        u_w = pb.solution(z, u0=u0)
        warm.append(pb.stats.iterations)
        assert pb.stats.converged and pb.stats.rel_residual <= 1e-9
        assert float((u_w - u_c).norm() / u_c.norm()) < 1e-6
    print(f"loss {hist[0]:.3e} -> {hist[-1]:.3e}; iterations cold {cold} previous {prev} network {warm}", flush=True)
    assert sum(warm) < sum(cold)
