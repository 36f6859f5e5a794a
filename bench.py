"""Benchmark: Tet4 SIMP Jacobian + residual assembly throughput on one MI355X.

A step is one full assembly (VectorFunction::gradient_u with the fused residual
R = K(z)u - f) of the 256x128x128 x 6 Tet4 cantilever (25.2M tets, 4.28M nodes, 63.5M
graph blocks), inputs resident in HBM.  Prints ONE JSON line.  Also measured (reported beside `value`):
solve + sensitivity time (Problem::solution to 1e-8 relative true residual, then
dC/dz) for FP64 and mixed-precision PCG with block-Jacobi and multigrid
preconditioning, and the CPU baseline (the oracle restatement of the reference
algorithm on a bounded sample, all host threads and one thread).

    python bench.py [--steps K] [--warmup W] [--config NAME] [--no-cpu] [--no-solve] [--no-general]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Tet4 SIMP assembly Melem/s; solve+sensitivity time to 1e-8 rel residual"
HBM_PEAK_GBS = 8000.0


def asm_bytes(E: int, N: int, nnzb: int) -> int:
    """Compulsory HBM bytes of one Jacobian+residual assembly (int32 indices, FP64 values):
    conn 16E, coords/z/u 56N, rowmap 4(N+1), colidx 4nnz_b, values 72nnz_b, R 24N."""
    return 16 * E + 56 * N + 4 * (N + 1) + 4 * nnzb + 72 * nnzb + 24 * N


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def slab_sample(w, nxs: int):
    """The first `nxs` hex layers along x of a box workload: same lattice spacing, element
    shapes, material and densities (the workload's own density field restricted to the
    slab), as a bounded CPU sample whose per-element work is that of the full mesh."""
    from platoanalyze_amd.mesh import build_box, density_lattice, lattice_to_nodes

    c = w.cfg
    h = (c.lx / c.nx, c.ly / c.ny, c.lz / c.nz)
    mesh = build_box(nxs, c.ny, c.nz, nxs * h[0], c.ly, c.lz)
    lat = density_lattice(c.nx, c.ny, c.nz, c.seed, rho_min=c.rho_min)
    return mesh, lattice_to_nodes(lat, mesh.coords, *h)


def cpu_baseline(w, threads: int) -> dict:
    """The oracle (the reference algorithm restated: SFad worksets, atomic scatter after a
    linear row search, two Jacobian passes) timed on slabs of the benched workload:
    gradient_u + value on a 32-hex slab with `threads` OpenMP threads (median of 5 after
    one warm-up) and on a 4-hex slab with one thread (median of 3)."""
    import numpy as np

    from oracle import oracle as orc

    isa = orc.select_isa()

    def timed(nxs: int, nthreads: int, reps: int):
        mesh, z = slab_sample(w, nxs)
        o = orc.Model(mesh.coords, mesh.conn, w.cfg.E, w.cfg.nu, w.cfg.p, w.cfg.rho_min)
        rm, ci = orc.graph(mesh.conn, mesh.num_nodes)
        u = np.random.default_rng(1).normal(size=3 * mesh.num_nodes) * 1e-3
        orc.set_threads(nthreads)
        times = []
        for _ in range(reps + 1):  # 1 warm-up
            t = time.perf_counter()
            o.jacobian(u, z, rm, ci)
            o.residual(u, z)
            times.append(time.perf_counter() - t)
        return mesh.num_elems, sorted(times[1:])[reps // 2]

    ne, med = timed(min(32, w.cfg.nx), threads, 5)
    ne1, med1 = timed(min(4, w.cfg.nx), 1, 3)
    orc.set_threads(threads)
    return {"value": ne / med / 1e6, "unit": "Melem/s", "cores": threads, "kind": "port",
            "value_1_thread": ne1 / med1 / 1e6, "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "oracle_isa": isa,
            "sample": f"x-slabs of {w.cfg.name} (same spacing and densities): gradient_u+value, "
                      f"{min(32, w.cfg.nx)}-hex slab ({ne} tets) on {threads} threads, median of 5 after 1 "
                      f"warm-up, {med:.3f} s each; {min(4, w.cfg.nx)}-hex slab ({ne1} tets) on 1 thread, "
                      f"median of 3, {med1:.3f} s"}


def density_checksum(w) -> str:
    """First 16 hex digits of the SHA-256 of the workload's densities in node order (FP64)."""
    import hashlib

    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(w.z, np.float64).tobytes()).hexdigest()[:16]


def cpu_threads() -> int:
    """Threads of this process's CPU share: OMP_NUM_THREADS when set (a GPU box grants 16
    cores per GPU and sets it), else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def profile_unet(args) -> int:
    """The warm-start U-Net on the node grid of the workload (BF16 MFMA convolutions):
    HIP-event time of a forward pass and of a gradient (forward + backward), and the
    forward convolutions' dense FLOP rate against the BF16 MFMA peak."""
    import torch

    from platoanalyze_amd import workloads as wl
    from platoanalyze_amd.unet import IN_CH, OUT_CH, UNet

    c = wl.CONFIGS[args.config]
    grid, batch = (c.nx + 1, c.ny + 1, c.nz + 1), 4
    net = UNet(grid, args.unet_channels, seed=0)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((batch, *grid, IN_CH), generator=g, device="cuda")
    y = torch.randn((batch, *grid, OUT_CH), generator=g, device="cuda") * 0.1
    out = {}
    passes = [("forward", lambda: net.forward(x))]
    if not args.unet_forward_only:
        passes.append(("gradient", lambda: net.gradient(x, y)))
    for name, fn in passes:
        for _ in range(max(args.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.profile_unet):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = {"ms": e0.elapsed_time(e1) / args.profile_unet}
    fl = net.forward_flops(batch)
    out["forward"].update({"flops": fl, "TFLOPs": fl / (out["forward"]["ms"] * 1e-3) / 1e12})
    out["forward"]["frac_bf16_dense_peak"] = out["forward"]["TFLOPs"] / 2500.0
    print(json.dumps({"profile_unet": out, "grid": grid, "batch": batch, "channels": args.unet_channels,
                      "config": args.config}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cantilever_256x128x128")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-general", action="store_true", help="skip timing the general chunk kernel")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (for scaling HBM counters)")
    ap.add_argument("--profile-solve", type=int, default=0, metavar="K",
                    help="run only the solves (FP64/mixed x Jacobi/multigrid), each capped at K iterations "
                         "(for profilers)")
    ap.add_argument("--profile-unet", type=int, default=0, metavar="K",
                    help="run only the warm-start U-Net: K timed forward passes, then K gradients "
                         "(node grid of --config, --unet-channels, batch 4)")
    ap.add_argument("--unet-channels", type=int, default=16)
    ap.add_argument("--unet-forward-only", action="store_true")
    ap.add_argument("--warm-start-study", action="store_true",
                    help="also run the learned warm-start study at 64x32x32 (about 4 minutes) and report it "
                         "in place of the committed one")
    ap.add_argument("--solvers", default="fp64_jacobi,mixed_jacobi,fp64_multigrid,mixed_multigrid",
                    help="comma-separated precision_preconditioner variants for --profile-solve")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the assembly/solve path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built). Use --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from platoanalyze_amd import _capi
    from platoanalyze_amd import workloads as wl
    from platoanalyze_amd.elliptic import Problem

    dev = "cuda"
    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(args.warmup + args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": args.warmup + args.steps}))
        return 0
    if args.profile_unet:
        return profile_unet(args)
    if args.profile_solve:
        w = wl.make(args.config)
        m = wl.model(w, device=dev)
        z = torch.as_tensor(w.z, device=dev)
        out = {}
        for name in args.solvers.split(","):
            prec, pre = name.split("_")
            pb = Problem(m, tolerance=1e-8, check_every=32, precision=prec, preconditioner=pre,
                         max_iters=args.profile_solve)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            try:
                pb.solution(z)
            except _capi.PaError as e:  # the cap is expected to stop the solve
                if e.status != 5:
                    raise
            torch.cuda.synchronize()
            out[f"{prec}_{pre}"] = {"seconds": time.perf_counter() - t0, "iterations": pb.stats.iterations}
        print(json.dumps({"profile_solve": out, "config": args.config}))
        return 0
    t0 = time.perf_counter()
    w = wl.make(args.config)
    t_mesh = time.perf_counter() - t0
    t0 = time.perf_counter()
    m = wl.model(w, device=dev)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    E, N, nnzb = w.mesh.num_elems, w.mesh.num_nodes, m.nnz_blocks
    u = torch.as_tensor(np.random.default_rng(7).normal(size=3 * N) * 1e-3, device=dev)
    z = torch.as_tensor(w.z, device=dev)
    vals = torch.empty(9 * nnzb, dtype=torch.float64, device=dev)
    R = torch.empty(3 * N, dtype=torch.float64, device=dev)
    h = m.handle
    L = _capi.lib()
    pu, pz, pv, pr = (C.c_void_p(t.data_ptr()) for t in (u, z, vals, R))
    m._stream()

    def step():
        rc = L.pa_vf_gradient_u(h, pu, pz, pv, pr)
        if rc:
            _capi.check(rc, "pa_vf_gradient_u")

    box = C.c_int(0)
    _capi.call("pa_box_assembly", h, C.byref(box))
    # the library takes k_box_jac_ws_ahead when the lattice weights and u (72 B per padded
    # hex) exceed 256 MB, unless PA_BOX_AHEAD says otherwise
    c = w.cfg
    ahead = 72 * (c.nx + 2) * (c.ny + 2) * (c.nz + 2) > (256 << 20)
    if os.environ.get("PA_BOX_AHEAD"):
        ahead = os.environ["PA_BOX_AHEAD"] != "0"
    kernel_name = ("k_jac_chunk<4,true> (Jacobian+residual)" if not box.value else
                   "k_box_scatter + k_box_w + k_box_jac_ws_ahead<true> + k_add_load (lattice Jacobian+residual)"
                   if ahead else "k_box_scatter + k_box_w + k_box_jac_ws<true> (lattice Jacobian+residual)")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        t0 = time.perf_counter()
        for i in range(steps):
            starts[i].record()
            fn()
            ends[i].record()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / steps

    wall, kern_ms = timed(step, args.steps, args.warmup)
    ms_per_step = wall / args.steps * 1e3
    value = E / (wall / args.steps) / 1e6
    B = asm_bytes(E, N, nnzb)
    achieved = B / (kern_ms * 1e-3) / 1e9

    # measured ceilings on this box: a 1 GiB device copy (read + write bytes) and a 1 GiB
    # fill (write bytes), HIP events over 20 launches each
    def ceiling(fn, nbytes, reps=20):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9

    src_c = torch.ones(1 << 27, dtype=torch.float64, device=dev)
    dst_c = torch.empty_like(src_c)
    copy_gbs = ceiling(lambda: dst_c.copy_(src_c), 2 * src_c.numel() * 8)
    fill_gbs = ceiling(lambda: dst_c.fill_(1.0), dst_c.numel() * 8)
    del src_c, dst_c

    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.config}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")

    extra = {}
    if box.value and not args.no_general:
        # the general chunk kernel on the same problem (a second handle without the
        # lattice kernel), for comparison
        os.environ["PA_BOX_ASSEMBLY"] = "0"
        try:
            mg_ = wl.model(w, device=dev)
        finally:
            del os.environ["PA_BOX_ASSEMBLY"]
        mg_._stream()
        hg = mg_.handle

        def step_general():
            rc = L.pa_vf_gradient_u(hg, pu, pz, pv, pr)
            if rc:
                _capi.check(rc, "pa_vf_gradient_u")

        _, kg_ms = timed(step_general, args.steps, args.warmup)
        extra["general_kernel"] = {"kernel": "k_jac_chunk<4,true> (Jacobian+residual)", "kernel_ms": kg_ms,
                                   "achieved_GBs": B / (kg_ms * 1e-3) / 1e9,
                                   "frac": B / (kg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        del mg_
        m._stream()
    if not args.no_solve:
        # solve K(z)u = f to 1e-8 relative residual, then C and dC/dz, with each solver:
        # FP64 block-CRS PCG and FP64 PCG with the FP32 matrix-free operator and
        # true-residual replacement, each with 3x3 block-Jacobi or geometric multigrid
        # preconditioning.  Each variant runs twice: the first call builds the multigrid
        # hierarchy (once per mesh), the second is timed; errors are against FP64 Jacobi.
        methods = {
            ("fp64", "jacobi"): "FP64 block-CRS PCG, 3x3 block Jacobi",
            ("mixed", "jacobi"): "FP64 PCG, FP32 matrix-free operator, true-residual replacement, 3x3 block Jacobi",
            ("fp64", "multigrid"): "FP64 block-CRS PCG, geometric multigrid V-cycle",
            ("mixed", "multigrid"): "FP64 PCG, FP32 matrix-free operator, true-residual replacement, "
                                    "geometric multigrid V-cycle",
        }
        res = {}
        for (prec, pre) in methods:
            pb = Problem(m, tolerance=1e-8, check_every=32 if pre == "jacobi" else 4, precision=prec,
                         preconditioner=pre)
            secs = []
            for _ in range(2):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                u_s = pb.solution(z)
                g, Cval = pb.criterionGradient(z, u_s)
                torch.cuda.synchronize()
                secs.append(time.perf_counter() - t0)
            res[(prec, pre)] = {"seconds": secs[1], "first_call_seconds": secs[0],
                                "iterations": pb.stats.iterations, "outer_iterations": pb.stats.outer_iterations,
                                "true_rel_residual": pb.stats.rel_residual, "compliance": Cval, "u": u_s, "g": g}
        a = res[("fp64", "jacobi")]
        extra["solve_sensitivity_s"] = a["seconds"]  # the plain FP64 path (block-CRS PCG, block Jacobi)
        best = min(res, key=lambda k: res[k]["seconds"])
        extra["solve_sensitivity_best"] = {"variant": f"{best[0]}_{best[1]}", "seconds": res[best]["seconds"],
                                           "speedup_vs_fp64_jacobi": a["seconds"] / res[best]["seconds"]}
        solve_out = {}
        for key, r in res.items():
            d = {"method": methods[key], **{k: r[k] for k in ("seconds", "first_call_seconds", "iterations",
                                                              "outer_iterations", "true_rel_residual", "compliance")}}
            if key != ("fp64", "jacobi"):
                d.update({"rel_err_u": float((r["u"] - a["u"]).norm() / a["u"].norm()),
                          "rel_err_compliance": abs(r["compliance"] - a["compliance"]) / abs(a["compliance"]),
                          "rel_err_gradient": float((r["g"] - a["g"]).norm() / a["g"].norm()),
                          "speedup_vs_fp64_jacobi": a["seconds"] / r["seconds"]})
            solve_out[f"{key[0]}_{key[1]}"] = d
        extra["solve"] = solve_out

    # the learned warm start on the 64x32x32 cantilever: deterministic FP64 iteration
    # counts with and without the network (20 OC trajectories, 2000 solves, U-Net 16-128
    # channels trained by the library's kernels), live or from the committed study
    ws_file = ROOT / "profiles" / "warm_start_cantilever_64x32x32.json"
    if args.warm_start_study:
        from platoanalyze_amd.train_warmstart import study_config3

        st = study_config3(wl.make("cantilever_64x32x32"), 20, 100, 30, 16, log=lambda s_: None)
        st["source"] = "measured in this run"
    elif ws_file.exists():
        st = json.loads(ws_file.read_text())
        st["source"] = f"{ws_file.relative_to(ROOT)} (committed; bench.py --warm-start-study reruns it)"
    else:
        st = None
    if st is not None:
        extra["warm_start_study"] = {k: st.get(k) for k in ("config", "source", "solves", "channels", "epochs",
                                                            "train_seconds", "final_loss", "mean", "checks")}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_threads or cpu_threads())

    out = {
        "metric": METRIC, "value": value, "unit": "Melem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.config, "elements": E, "nodes": N, "dofs": 3 * N, "nnz_blocks": nnzb,
                   "step": "gradient_u + fused residual (R = K(z)u - f)", "parallelism": "single GPU",
                   "density_seed": w.cfg.seed, "density_sha256_16": density_checksum(w)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel_name, "kernel_ms": kern_ms,
                     "algorithmic_bytes": B,
                     "measured_copy_GBs": copy_gbs, "measured_fill_GBs": fill_gbs,
                     "frac_of_measured_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "setup_s": {"mesh": t_mesh, "graph_chunks_upload": t_setup},
    }
    out.update(extra)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
