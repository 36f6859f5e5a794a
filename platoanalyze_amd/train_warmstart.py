"""Train the warm-start U-Net on SIMP optimality-criteria trajectories and evaluate it.

    python -m platoanalyze_amd.train_warmstart --config cantilever_64x32x32 \
        --traj 20 --iters 100 --epochs 10 --out weights.npz

Data: one multigrid solve per OC iteration (random volume fraction and traction
direction per trajectory); the last trajectory is held out.  Evaluation: on the held-out
designs, the mixed-precision (or, with --eval-precision fp64, the FP64) block-Jacobi PCG
solve to 1e-8 from zero against the same solve from the network's warm start (iterations
and seconds), both judged by the true FP64 residual;
in update mode also from the previous design's solution, with Jacobi and multigrid
preconditioning.  `--apply WORKLOAD` re-instantiates the trained network on that
workload's node grid (the network is fully convolutional) and compares mixed multigrid
solves of its design from zero and from the network's guess.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch


def evaluate(net, c0, held, tol=1e-8, log=print, mode=None, precision="mixed"):
    from . import workloads as wl
    from .elliptic import Problem
    from .warmstart import WarmStart

    rows = []
    for w, z, f in held:
        m = wl.model(w)
        ws = WarmStart(net, w, m, c0, mode=mode)
        pb = Problem(m, tolerance=tol, check_every=16, precision=precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u_cold = pb.solution(z)
        torch.cuda.synchronize()
        t_cold, it_cold = time.perf_counter() - t0, pb.stats.iterations
        t0 = time.perf_counter()
        u0 = ws.initial_guess(z, f)
        u_warm = pb.solution(z, u0=u0)
        torch.cuda.synchronize()
        t_warm, it_warm = time.perf_counter() - t0, pb.stats.iterations
        r0 = float((u0 - u_cold).norm() / u_cold.norm())
        rows.append({"iters_cold": it_cold, "iters_warm": it_warm, "s_cold": t_cold, "s_warm": t_warm,
                     "u0_rel_err": r0, "u_rel_diff": float((u_warm - u_cold).norm() / u_cold.norm())})
        log(json.dumps(rows[-1]))
    return rows


def evaluate_update(net, c0, held, tol=1e-8, log=print, preconditioner="jacobi"):
    """Update mode: each held-out design solved from zero, from the previous design's
    solution and from the network's update of it (iterations, energy excess)."""
    from . import workloads as wl
    from .elliptic import Problem
    from .warmstart import WarmStart, energy

    rows = []
    for w, zp, up, z, f in held:
        m = wl.model(w)
        ws = WarmStart(net, w, m, c0, mode="update")
        pb = Problem(m, tolerance=tol, check_every=4 if preconditioner == "multigrid" else 16, precision="mixed",
                     preconditioner=preconditioner)
        u = pb.solution(z).clone()
        it_cold = pb.stats.iterations
        pb.solution(z, u0=up)
        it_prev = pb.stats.iterations
        u0 = ws.initial_guess(z, f, zp, up)
        u_w = pb.solution(z, u0=u0)
        it_warm = pb.stats.iterations
        e_star = energy(m, z, u, f)
        rows.append({"iters_cold": it_cold, "iters_prev": it_prev, "iters_warm": it_warm,
                     "energy_excess_prev": energy(m, z, up, f) - e_star, "energy_excess_warm": energy(m, z, u0, f) - e_star,
                     "u_rel_diff": float((u_w - u).norm() / u.norm())})
        log(json.dumps(rows[-1]))
    return rows


def evaluate_transfer(net, c0, mode, config, tol=1e-8, log=print):
    """The trained network (fully convolutional) re-instantiated with the same parameters
    on another box's node grid and applied to that workload's design: mixed multigrid
    solves from zero and from the network's guess (iterations, seconds including the
    network's forward pass, the guess's error against the solution)."""
    from . import workloads as wl
    from .elliptic import Problem
    from .unet import UNet
    from .warmstart import Features, WarmStart

    w = wl.make(config)
    m = wl.model(w)
    big = UNet(tuple(Features(w).dims), net.base)
    big.set_params(net.params())
    ws = WarmStart(big, w, m, c0, mode=mode)
    z = torch.as_tensor(w.z, dtype=torch.float64, device="cuda")
    pb = Problem(m, tolerance=tol, check_every=4, precision="mixed", preconditioner="multigrid")
    pb.solution(z)  # builds the multigrid hierarchy (once per mesh)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = pb.solution(z).clone()
    torch.cuda.synchronize()
    t_cold, it_cold = time.perf_counter() - t0, pb.stats.iterations
    t0 = time.perf_counter()
    u0 = ws.initial_guess(z)
    u_w = pb.solution(z, u0=u0)
    torch.cuda.synchronize()
    t_warm, it_warm = time.perf_counter() - t0, pb.stats.iterations
    row = {"config": w.cfg.name, "grid": list(big.grid), "mode": ws.mode, "iters_cold": it_cold,
           "iters_warm": it_warm, "s_cold": t_cold, "s_warm": t_warm,
           "u0_rel_err": float((u0 - u).norm() / u.norm()), "u_rel_diff": float((u_w - u).norm() / u.norm())}
    log(json.dumps(row))
    return row


def study_config3(base, n_traj: int = 20, n_iter: int = 100, epochs: int = 30, channels: int = 16,
                  n_eval: int = 10, seed: int = 0, mode: str = "load", lr: float = 3e-4, batch: int = 4,
                  log=print) -> dict:
    """The learned warm start on the mixed-precision / warm-start workload: n_traj seeded
    SIMP OC trajectories of n_iter designs each, one FP64-accurate solve per design
    (n_traj * n_iter solves), the last trajectory held out; the U-Net (4 levels, channels
    .. 8 * channels) trained on the others by the library's BF16-MFMA kernels.  On n_eval
    held-out designs, deterministic FP64 block-Jacobi PCG iteration counts to 1e-8: from
    zero, from the network's guess, deflated by the guess (pa_solution_deflated), deflated
    by the previous design's solution, and by both; FP64 multigrid from zero and from the
    guess; and the mixed multigrid solve (the fast path) from zero and from the guess."""
    from . import workloads as wl
    from .elliptic import Problem
    from .warmstart import WarmStart, train

    t0 = time.perf_counter()
    # n_traj - 1 training trajectories and one held out: n_traj * n_iter solves
    net, c0, hist, held = train(base, n_traj - 1, n_iter, epochs, batch=batch, lr=lr, base_channels=channels,
                                seed=seed, mode=mode, holdout=1, log=log)
    t_train = time.perf_counter() - t0
    idx = np.linspace(1, len(held) - 1, min(n_eval, len(held) - 1)).astype(int)
    rows = []
    for i in idx:
        w, z, f = held[i]
        m = wl.model(w)
        ref = Problem(m, tolerance=1e-10, check_every=2, preconditioner="multigrid")
        u_prev = ref.solution(held[i - 1][1]).clone()
        u_ref = ref.solution(z).clone()
        ws = WarmStart(net, w, m, c0, mode=mode)
        g = ws.initial_guess(z, f)
        row = {"design": int(i)}
        pj = Problem(m, tolerance=1e-8, check_every=1)
        pj.solution(z)
        row["fp64_jacobi_zero"] = pj.stats.iterations
        pj.solution(z, u0=g)
        row["fp64_jacobi_net_start"] = pj.stats.iterations
        for name, W in (("net", ws.deflation_space(z, f)), ("prev", u_prev[None, :]),
                        ("net_prev", ws.deflation_space(z, f, u_prev))):
            pj.solution_deflated(z, W)
            row[f"fp64_jacobi_deflated_{name}"] = pj.stats.iterations
        pm = Problem(m, tolerance=1e-8, check_every=1, preconditioner="multigrid")
        pm.solution(z)
        row["fp64_multigrid_zero"] = pm.stats.iterations
        pm.solution(z, u0=g)
        row["fp64_multigrid_net_start"] = pm.stats.iterations
        px = Problem(m, tolerance=1e-8, check_every=2, precision="mixed", preconditioner="multigrid")
        px.solution(z)
        row["mixed_multigrid_zero"] = px.stats.iterations
        px.solution(z, u0=g)
        row["mixed_multigrid_net_start"] = px.stats.iterations
        row["guess_rel_err"] = float((g - u_ref).norm() / u_ref.norm())
        rows.append(row)
        log(json.dumps(row))
    keys = [k for k in rows[0] if k not in ("design",)]
    summary = {k: float(np.mean([r[k] for r in rows])) for k in keys}
    out = {"config": base.cfg.name, "trajectories": n_traj, "designs_per_trajectory": n_iter,
           "solves": n_traj * n_iter, "channels": [channels * 2 ** i for i in range(4)], "epochs": epochs,
           "mode": mode, "lr": lr, "batch": batch, "train_seconds": t_train,
           "loss_history": [float(v) for v in hist], "final_loss": hist[-1] if hist else None,
           "mean": summary, "checks": study_checks(rows), "rows": rows}
    return out


def study_checks(rows: list[dict]) -> dict:
    """Per-design comparisons of the deterministic FP64 block-Jacobi counts: does each
    use of the network cost no iterations, on every held-out design?"""
    def every(a, b):
        return all(r[a] <= r[b] for r in rows)

    return {"deflated_net_le_zero": every("fp64_jacobi_deflated_net", "fp64_jacobi_zero"),
            "deflated_net_prev_le_deflated_prev": every("fp64_jacobi_deflated_net_prev", "fp64_jacobi_deflated_prev"),
            "net_start_le_zero": every("fp64_jacobi_net_start", "fp64_jacobi_zero"),
            "multigrid_net_start_le_zero": every("fp64_multigrid_net_start", "fp64_multigrid_zero")}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cantilever_64x32x32")
    ap.add_argument("--tiny", default=None, help="nx,ny,nz: same physics on a smaller grid")
    ap.add_argument("--traj", type=int, default=20)
    ap.add_argument("--iters", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=None, help="Adam learning rate (default 1e-3; 3e-4 for --study)")
    ap.add_argument("--channels", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", default="load", choices=["load", "shape", "basis", "update"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--eval-precision", default="mixed", choices=["mixed", "fp64"],
                    help="solver of the held-out evaluation (fp64: deterministic iteration counts)")
    ap.add_argument("--apply", default=None,
                    help="also apply the trained network, fully convolutionally, to this workload")
    ap.add_argument("--study", action="store_true",
                    help="the warm-start study (study_config3): deterministic FP64 iteration counts with and "
                         "without the network, written as JSON to --out")
    a = ap.parse_args(argv)
    if a.study:
        from . import workloads as wl
        tiny = tuple(int(v) for v in a.tiny.split(",")) if a.tiny else None
        res = study_config3(wl.make(a.config, tiny=tiny), a.traj, a.iters, a.epochs, a.channels, seed=a.seed,
                            mode=a.mode, lr=a.lr if a.lr is not None else 3e-4, batch=a.batch,
                            log=lambda s: print(s, flush=True))
        with open(a.out, "w") as fh:
            json.dump(res, fh, indent=1)
        print(json.dumps({"study": a.out, "mean": res["mean"]}), flush=True)
        return 0

    from . import workloads as wl
    from .warmstart import save, train

    log = lambda s: print(s, flush=True)  # noqa: E731
    tiny = tuple(int(v) for v in a.tiny.split(",")) if a.tiny else None
    base = wl.make(a.config, tiny=tiny)
    t0 = time.perf_counter()
    lr = a.lr if a.lr is not None else 1e-3
    net, c0, hist, held = train(base, a.traj, a.iters, a.epochs, batch=a.batch, lr=lr, base_channels=a.channels,
                            seed=a.seed, mode=a.mode, log=log)
    log(f"training done in {time.perf_counter() - t0:.1f} s")
    idx = np.linspace(0, len(held) - 1, min(8, len(held))).astype(int)
    if a.mode == "update":
        rows = []
        for pre in ("jacobi", "multigrid"):
            log(f"evaluation, {pre} preconditioner")
            rows += evaluate_update(net, c0, [held[i] for i in idx], log=log, preconditioner=pre)
    else:
        rows = evaluate(net, c0, [held[i] for i in idx], log=log, mode=a.mode, precision=a.eval_precision)
    meta = {"config": base.cfg.name, "mode": a.mode, "traj": a.traj, "iters": a.iters, "epochs": a.epochs,
            "loss_history": np.array(hist), "eval_iters_cold": np.array([r["iters_cold"] for r in rows]),
            "eval_iters_warm": np.array([r["iters_warm"] for r in rows])}
    if a.apply:
        if a.mode == "update":
            raise SystemExit("--apply needs a mode that predicts from the design (load, shape or basis)")
        log(f"applied to {a.apply}")
        evaluate_transfer(net, c0, a.mode, a.apply, log=log)
    save(net, c0, a.out, meta)
    log(json.dumps({"saved": a.out, "final_loss": hist[-1] if hist else None,
                    "mean_iters_cold": float(np.mean([r["iters_cold"] for r in rows])),
                    "mean_iters_warm": float(np.mean([r["iters_warm"] for r in rows])),
                    "mean_iters_prev": (float(np.mean([r["iters_prev"] for r in rows])) if a.mode == "update"
                                        else None)}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
