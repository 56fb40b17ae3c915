#!/usr/bin/env python3
"""Headline benchmark of the LinearBasis interpolation engine on one MI355X.

Workload: single-state forward, adaptive sparse grid D=16 (seeded refinement
p=0.3 from level sum <= 2, max level 7, stopped at 60,018 nodes), dofs=32,
1,048,576 points x ~ U[0,1)^16, all resident in HBM before timing.

A step is one forward over the whole batch. The headline path is the
mixed-precision MFMA sweep (basis formed in FP64, contraction on
v_mfma_f32_32x32x2f32 with FP32 accumulation folded into FP64 after every
32-node chunk); the plain FP64 GPU sweep (same walk, FP64 contraction) is
timed beside it. Prints one JSON line.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--precision f32]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hddm-solver-postprocessors_amd"))

METRIC = "point·dof evals/sec on 1 MI355X, mixed/MFMA vs FP64 plain at ≤1e-6 rel err"
PEAK_TFLOPS = {"f16x2": 2500.0, "bf16x3": 2500.0, "f32": 157.3, "fp64": 78.6}
# FLOP of one MFMA instruction of each path's contraction
MFMA_FLOPS = {"f32": 2 * 32 * 32 * 2, "f16x2": 2 * 32 * 32 * 16, "bf16x3": 2 * 32 * 32 * 16}
DTYPE = {"f16x2": "f16", "bf16x3": "bf16", "f32": "f32", "fp64": "f64"}


def build_workload(npoints):
    import hddm_lb
    ii, jj, lv, _ = hddm_lb.generate_grid(16, 2, 7, 0.3, 1, 60000)
    S = hddm_lb.synthetic_surplus(lv, 32, 2)
    X = hddm_lb.synthetic_points(npoints, 16, 3)
    return ii, jj, S, X


def time_path(ctx, x, y, prec, steps, warmup):
    """Wall time per step over `steps` forwards, and the sweep kernels' mean
    duration from hipEvents recorded on the same stream in the same steps."""
    import torch
    import hddm_lb
    for _ in range(warmup):
        ctx.forward(x, y, None, prec)
    torch.cuda.synchronize()
    ctx.set_option(hddm_lb.OPT_TIMING, 1)
    ctx.sweep_time_summary()  # drop the warmup events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.forward(x, y, None, prec)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n, tot, mn, mx = ctx.sweep_time_summary()
    ctx.set_option(hddm_lb.OPT_TIMING, 0)
    return (t1 - t0) * 1e3 / steps, tot / max(n, 1), mn, mx


def time_multistate(args, prec, which):
    """Multistate forwards in one launch, timed like the headline (events
    around whole calls, inputs resident):
    d20x12: 12 states, D=20, adaptive grids (max level 6) with node counts
            log-uniform in [2k, 200k], dofs per state from {16, 32, 64},
            262,144 points with seeded state ids;
    d59x4:  4 states, D=59, regular grids sum(l-1) <= 3 (281,077 nodes each),
            dofs=128, 2^20 points with seeded state ids.
    The unit counts every output element: sum over points of the dofs of the
    point's state."""
    import numpy as np
    import torch
    import hddm_lb
    ctx = hddm_lb.Context(0)
    rng = np.random.default_rng(2024)
    if which == "d20x12":
        nnos = np.exp(rng.uniform(np.log(2000), np.log(200000), size=12)).astype(int)
        dofs = rng.choice([16, 32, 64], size=12)
        specs = [(20, 1, 6, 0.3, 300 + s, int(nnos[s]), int(dofs[s])) for s in range(12)]
        n, dim, steps = 262144, 20, max(2, args.steps // 2)
    else:
        specs = [(59, 3, 4, 0.0, 500 + s, 0, 128) for s in range(4)]
        n, dim, steps = 1 << 20, 59, 2
    nno_tot, dofs_of = 0, []
    for s, (d, start, maxl, p, seed, target, nd) in enumerate(specs):
        ii, jj, lv, _ = hddm_lb.generate_grid(d, start, maxl, p, seed, target)
        S = hddm_lb.synthetic_surplus(lv, nd, seed + 100)
        ctx.upload_state(s, ii, jj, S)
        nno_tot += ii.shape[0]
        dofs_of.append(nd)
    X = hddm_lb.synthetic_points(n, dim, 77)
    w = rng.uniform(0.5, 1.5, size=len(specs))
    sid = rng.choice(len(specs), size=n, p=w / w.sum()).astype(np.int32)
    work = float(np.asarray(dofs_of)[sid].sum())
    x = torch.from_numpy(X).cuda()
    sd = torch.from_numpy(sid).cuda()
    y = torch.empty((n, max(dofs_of)), dtype=torch.float64, device="cuda")
    res = {"workload": "%d states, D=%d, %d nodes in all, %d points" % (len(specs), dim, nno_tot, n)}
    for name, pr in (("mixed", prec), ("fp64", hddm_lb.FP64)):
        ctx.forward(x, y, sd, pr)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.forward(x, y, sd, pr)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        res[name] = {"ms_per_step": ms, "value": work / (ms * 1e-3)}
    res["unit"] = "point·dof/s"
    ctx.close()
    return res


def time_adjoint(args, prec):
    """Forward + adjoint on the differentiated workload: D=20 adaptive grid of
    ~100k nodes (max level 6), dofs=32, 65,536 points in [0.01, 0.99]^20,
    upstream gradient N(0, 1). A step is one forward (mixed path) and one FP64
    adjoint (dx and dsurplus) over the batch."""
    import numpy as np
    import torch
    import hddm_lb
    ii, jj, lv, _ = hddm_lb.generate_grid(20, 1, 6, 0.3, 404, 100000)
    S = hddm_lb.synthetic_surplus(lv, 32, 405)
    n = 65536
    X = hddm_lb.synthetic_points(n, 20, 406, 0.01, 0.99, kink_level=6, kink_gap=1e-3)
    G = np.random.default_rng(407).standard_normal((n, 32))
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    x = torch.from_numpy(X).cuda()
    g = torch.from_numpy(G).cuda()
    y = torch.empty((n, 32), dtype=torch.float64, device="cuda")
    dx = torch.empty((n, 20), dtype=torch.float64, device="cuda")
    _, total = ctx.surplus_offset(-1)
    ds = torch.empty((total,), dtype=torch.float64, device="cuda")

    def step():
        ctx.forward(x, y, None, prec)
        ctx.adjoint(x, g, dx, ds, None, hddm_lb.FP64)

    for _ in range(max(1, args.warmup // 2)):
        step()
    torch.cuda.synchronize()
    steps = max(1, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.forward(x, y, None, prec)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.close()
    return {"workload": "forward + adjoint, adaptive grid D=20, %d nodes, dofs=32, %d points" % (ii.shape[0], n),
            "ms_per_step": ms, "forward_ms": fwd_ms, "adjoint_ms": ms - fwd_ms,
            "point_dof_per_s": n * 32 / (ms * 1e-3),
            "adjoint_precision": "fp64", "forward_precision": args.precision}


def normwise_err(y, yref):
    import numpy as np
    den = np.abs(yref).max()
    return float(np.abs(y - yref).max() / den) if den > 0 else float(np.abs(y - yref).max())


def cpu_baseline(ii, jj, S, X, ygpu, target_s):
    """The reference's x86 xps/chains algorithm (the oracle's port of it) on
    the host cores, on a bounded prefix of the same points; also the checker
    for the GPU results on that prefix (the port against the FP64 gold
    restatement on its first 256 points)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    st = oracle.X86State(ii, jj, S)
    probe = 256
    t0 = time.perf_counter()
    st.forward(X[:probe], threads)
    dt = time.perf_counter() - t0
    m = int(min(X.shape[0], max(probe, probe * target_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    yc = st.forward(X[:m], threads)
    dt = time.perf_counter() - t0
    gold = oracle.forward(ii, jj, S, X[:probe])
    errs = {k: normwise_err(v[:m], yc) for k, v in ygpu.items()}
    return {
        "value": m * S.shape[1] / dt,
        "unit": "point·dof/s",
        "cores": threads,
        "kind": "port",
        "sample": "first %d of the %d points, x86 xps/chains algorithm (FP64), OpenMP static over points, %.1f s"
                  % (m, X.shape[0], dt),
    }, errs, normwise_err(yc[:probe], gold)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="f32", choices=["f16x2", "bf16x3", "f32"])
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and error check")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint measurement on the differentiated workload")
    ap.add_argument("--no-multistate", action="store_true",
                    help="skip the two multistate measurements (D=20 x 12 states, D=59 x 4 states)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device 3 times (calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a forward shards by points; "
              "run one process per GPU on its own share). Use --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import hddm_lb

    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    precs = {"f16x2": hddm_lb.F16X2, "bf16x3": hddm_lb.BF16X3, "f32": hddm_lb.F32, "fp64": hddm_lb.FP64}
    ii, jj, S, X = build_workload(args.points)
    nno, dofs, N = ii.shape[0], S.shape[1], X.shape[0]
    factors = int(np.count_nonzero((ii != 0) | (jj != 0)))
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    x = torch.from_numpy(X).cuda()
    y = torch.zeros((N, dofs), dtype=torch.float64, device="cuda")

    res, ygpu = {}, {}
    for name in (args.precision, "fp64"):
        ms, sweep_ms, mn, mx = time_path(ctx, x, y, precs[name], args.steps, args.warmup)
        res[name] = (ms, sweep_ms, mn, mx)
        ygpu[name] = y.cpu().numpy()
    ms, sweep_ms, _, _ = res[args.precision]
    # executed work of one sweep (outside the timed steps): trie ops evaluated
    # per 64-point block, chunks contracted and MFMA instructions issued
    ctx.set_option(hddm_lb.OPT_STATS, 1)
    ctx.forward(x, y, None, precs[args.precision])
    nops_ex, nprune, nchunk, nitem, nmfma = ctx.last_stats(5)
    ctx.set_option(hddm_lb.OPT_STATS, 0)

    # algorithmic work of one launch: dense contraction 2*nno*dofs plus
    # 4 flops per active factor, per point
    flops = N * (2.0 * nno * dofs + 4.0 * factors)
    # contraction work executed: the MFMA instructions issued (empty K-steps
    # and all-zero 32-point halves are skipped); on the FP64 path (VALU) the
    # 8-node chunks counted whole
    if args.precision in MFMA_FLOPS:
        exec_flops = nmfma * float(MFMA_FLOPS[args.precision])
    else:
        exec_flops = nchunk * 2.0 * 8 * 64 * 32
    chunk_tflops = exec_flops / (sweep_ms * 1e-3) / 1e12
    achieved = flops / (sweep_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.precision]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("sweep_%s_d16_60k_1m" % args.precision)

    out = {
        "metric": METRIC,
        "value": N * dofs / (ms * 1e-3),
        "unit": "point·dof/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (seeded grid, surpluses U[-1,1]*2^-sum(l), points U[0,1)^16)",
        "config": {"workload": "single-state forward, adaptive sparse grid D=16, %d nodes, dofs=%d, %d points"
                               % (nno, dofs, N),
                   "dim": 16, "nno": nno, "dofs": dofs, "points": N, "precision": args.precision},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "kernel": "sweep_kernel (%s)" % args.precision, "kernel_ms": sweep_ms,
                     "flops_per_launch": flops,
                     "frac_executed": chunk_tflops / peak,
                     "note": "achieved counts the dense contraction over all nodes; the trie walk "
                             "skips nodes whose basis is zero on a whole 64-point block, so frac "
                             "can exceed 1. frac_executed: FLOP of the MFMA instructions the "
                             "contraction issued over the same peak"},
        "fp64_plain": {"value": N * dofs / (res["fp64"][0] * 1e-3), "ms_per_step": res["fp64"][0],
                       "kernel_ms": res["fp64"][1],
                       "roofline_frac": flops / (res["fp64"][1] * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"]},
        "mixed_over_fp64": res["fp64"][0] / ms,
        "executed": {"walk_ops_per_block": nops_ex / max(nitem, 1),
                     "pruned_runs_per_block": nprune / max(nitem, 1),
                     "chunks_per_block": nchunk / max(nitem, 1), "blocks": nitem,
                     "mfma_per_block": nmfma / max(nitem, 1), "mfma_tflops": chunk_tflops},
    }
    # the same forward through host buffers (what the plugin's interpolate()
    # uses): x in and y out over PCIe, synchronous
    yh = np.empty((N, dofs))
    ctx.forward_host(X, yh, None, precs[args.precision])
    t0 = time.perf_counter()
    ctx.forward_host(X, yh, None, precs[args.precision])
    host_ms = (time.perf_counter() - t0) * 1e3
    out["host_buffers"] = {"ms_per_step": host_ms, "value": N * dofs / (host_ms * 1e-3),
                           "note": "x and y in pageable host memory, copies included"}
    if not args.no_adjoint:
        out["forward_adjoint"] = time_adjoint(args, precs[args.precision])
    if not args.no_multistate:
        out["multistate"] = {k: time_multistate(args, precs[args.precision], k) for k in ("d20x12", "d59x4")}
    if not args.no_cpu:
        cb, errs, port_err = cpu_baseline(ii, jj, S, X, ygpu, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["max_rel_err"] = {k: v for k, v in errs.items()}
        out["max_rel_err"]["cpu_port_vs_gold"] = port_err
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
