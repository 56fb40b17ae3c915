#!/usr/bin/env python3
"""Headline benchmark of the LinearBasis interpolation engine on one MI355X.

Headline workload: high-dimensional multistate forward + adjoint.
  4 states, D=59, regular sparse grids sum(l-1) <= 3 (281,077 nodes each),
  dofs=128, 1,048,576 points x ~ U[0,1)^59 with seeded state ids, upstream
  gradient N(0, 1); everything resident in HBM before timing.
A step is one forward (y) and one adjoint (dx and dsurplus) over a batch of
262,144 of those points (all four states; the workload's 2^20 points are four
steps); one step over all 2^20 points is timed after the steps as a check
that the rate does not depend on the batch. The mixed step runs HDDM_LB_F32:
the forward on FP32 MFMA, the adjoint's products on the bf16 MFMA with every
F32 operand split into three bf16 parts (the split f16 forward, HDDM_LB_F16X2,
misses 1e-6 on this workload and is reported beside it, not in the value);
the plain FP64 step runs both in FP64 with the same work queue, trie walk and
zero-skip.
value = points x dofs / mixed step time (point.dof evaluations per second,
each with its gradient). The sweep formulation and launch shape of every
forward are the autotuner's choice, reported in config.

Secondary lines (shorter): single-state forward D=16 (60k nodes, 32 dofs,
2^20 points), 12-state multistate forward D=20, and forward + adjoint on the
differentiated workload D=20 (100k nodes, 32 dofs, 65,536 points), each
mixed against FP64 with its error against the oracle; the per-point legacy
API through the drop-in plugin; the reference's CPU algorithm on the host.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hddm-solver-postprocessors_amd"))

METRIC = "point·dof evals/sec on 1 MI355X, mixed/MFMA vs FP64 plain at ≤1e-6 rel err"
PEAK_TFLOPS = {"f16x2": 2500.0, "bf16x3": 2500.0, "f32": 157.3, "fp64": 78.6}
# published per-state time of the reference's CUDA backend on the per-point API
# (LinearBasis/tests/InterpolateArrayManyMultistate/medium/daint.log:152)
REF_CUDA_US_PER_STATE = 122.467


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def normwise_err(y, yref):
    """max over dof columns of max |y - yref| / max |yref| (columns of zeros: absolute)."""
    import numpy as np
    y = np.asarray(y, np.float64)
    yref = np.asarray(yref, np.float64)
    den = np.abs(yref).max(axis=0)
    num = np.abs(y - yref).max(axis=0)
    den = np.where(den > 0, den, 1.0)
    return float((num / den).max()) if y.size else 0.0


def oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


# ------------------------------------------------------------------ workloads

def d59_states(ctx):
    import hddm_lb
    states = []
    for s in range(4):
        ii, jj, lv, _ = hddm_lb.generate_grid(59, 3, 4, 0.0, 500 + s, 0)
        S = hddm_lb.synthetic_surplus(lv, 128, 600 + s)
        ctx.upload_state(s, ii, jj, S)
        states.append((ii, jj, S))
    return states


def timed_steps(step, steps, warmup, ctx=None, label=None):
    """Wall ms per step over `steps` calls after `warmup`; with ctx, the mean
    sweep (kind 0) and adjoint (kind 1) kernel ms per step from the library's
    HIP events on the calls' stream, recorded in the same steps. With a
    label, every step is synchronized and reported on stderr (long steps)."""
    import torch
    import hddm_lb
    for i in range(warmup):
        step()
        if label:
            torch.cuda.synchronize()
            say("bench: %s warm-up step %d/%d" % (label, i + 1, warmup))
    torch.cuda.synchronize()
    if ctx is not None:
        ctx.set_option(hddm_lb.OPT_TIMING, 1)
        ctx.sweep_time_summary(0)
        ctx.sweep_time_summary(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        if label:
            torch.cuda.synchronize()
            say("bench: %s step %d/%d, %.0f ms so far" % (label, i + 1, steps, (time.perf_counter() - t0) * 1e3))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    kern = None
    if ctx is not None:
        kern = []
        for kind in (0, 1):
            n, tot, mn, mx = ctx.sweep_time_summary(kind)
            kern.append({"launches": n, "ms_per_step": tot / steps, "ms_per_launch": tot / max(n, 1),
                         "min_ms": mn, "max_ms": mx})
        ctx.set_option(hddm_lb.OPT_TIMING, 0)
    return ms, kern


def headline(args):
    """4 states x D=59 forward + adjoint: mixed vs FP64."""
    import numpy as np
    import torch
    import hddm_lb
    ctx = hddm_lb.Context(0)
    ctx.set_option(hddm_lb.OPT_AUTOTUNE, 1)
    t0 = time.perf_counter()
    states = d59_states(ctx)
    say("bench: D=59 states uploaded in %.1f s" % (time.perf_counter() - t0))
    n = args.points
    X = hddm_lb.synthetic_points(n, 59, 77)
    sid = np.random.default_rng(2024).integers(0, 4, size=n).astype(np.int32)
    G = np.random.default_rng(94).standard_normal((n, 128))
    x = torch.from_numpy(X).cuda()
    sd = torch.from_numpy(sid).cuda()
    g = torch.from_numpy(G).cuda()
    y = torch.empty((n, 128), dtype=torch.float64, device="cuda")
    dx = torch.empty((n, 59), dtype=torch.float64, device="cuda")
    _, total = ctx.surplus_offset(-1)
    ds = torch.empty((total,), dtype=torch.float64, device="cuda")
    legs = {"mixed": (hddm_lb.F32, hddm_lb.F32), "fp64": (hddm_lb.FP64, hddm_lb.FP64)}
    res = {}
    # oracle references on a seeded subset: y and dx at 16 points per state
    refs = []
    if not args.no_errors:
        oracle = oracle_mod()
        rng = np.random.default_rng(95)
        t0 = time.perf_counter()
        for s, (ii, jj, S) in enumerate(states):
            pick = rng.choice(np.flatnonzero(sid == s), size=16, replace=False)
            dxo, _ = oracle.adjoint(ii, jj, S, X[pick], G[pick])
            refs.append((pick, oracle.forward(ii, jj, S, X[pick]), dxo))
        say("bench: oracle references in %.1f s" % (time.perf_counter() - t0))

    def errors():
        Y = y.cpu().numpy()
        DX = dx.cpu().numpy()
        return {"y": max(normwise_err(Y[p], yo) for p, yo, _ in refs),
                "dx": max(normwise_err(DX[p], dxo) for p, _, dxo in refs)}

    for name, (pf, pa) in legs.items():
        def step(pf=pf, pa=pa):
            ctx.forward(x, y, sd, pf)
            ctx.adjoint(x, g, dx, ds, sd, pa)
        # the plain FP64 leg is the slower one: it runs a fifth of the steps
        # (at least 2) after one warm-up step, and says so in its line
        steps = args.steps if name == "mixed" else max(2, args.steps // 10)
        warm = args.warmup if name == "mixed" else 1
        ms, kern = timed_steps(step, steps, warm, ctx, label="D=59 " + name)
        cfg = ctx.last_config()
        status = ctx.device_status()
        res[name] = {"ms_per_step": ms, "steps": steps, "warmup": warm, "kernels": kern, "forward_config": cfg,
                     "status": status}
        say("bench: D=59 %s step %.1f ms (forward sweep %.1f, adjoint %.1f), config %s"
            % (name, ms, kern[0]["ms_per_step"], kern[1]["ms_per_step"], cfg))
        assert status == 0, "device status %d after the %s leg" % (status, name)
        res[name]["max_rel_err"] = None if args.no_errors else errors()
    # the split f16 forward alone (not in the value: it misses 1e-6 here)
    ms16, k16 = timed_steps(lambda: ctx.forward(x, y, sd, hddm_lb.F16X2), max(2, args.steps // 4), 1, ctx)
    res["forward_f16x2"] = {"ms_per_forward": ms16, "sweep_ms": k16[0]["ms_per_step"], "config": ctx.last_config(),
                            "max_rel_err_y": None if args.no_errors else errors()["y"]}
    say("bench: D=59 f16x2 forward %.1f ms" % ms16)
    # dS error on a 128-point slice (the oracle's dS cost grows with the points)
    m = 128
    for name, (pf, pa) in (() if args.no_errors else legs.items()):
        ds2 = torch.empty_like(ds)
        ctx.adjoint(x[:m], g[:m], None, ds2, sd[:m], pa)
        torch.cuda.synchronize()
        D = ds2.cpu().numpy()
        oracle = oracle_mod()
        e = 0.0
        for s, (ii, jj, S) in enumerate(states):
            sel = sid[:m] == s
            _, dso = oracle.adjoint(ii, jj, S, X[:m][sel], G[:m][sel], nthreads=min(16, os.cpu_count() or 1))
            off, _ = ctx.surplus_offset(s)
            e = max(e, normwise_err(D[off:off + ii.shape[0] * 128].reshape(-1, 128), dso))
        res[name]["max_rel_err"]["dsurplus"] = e
    # one mixed step over all 2^20 points of the workload (rate check)
    full = None
    if args.full_check and n < (1 << 20):
        nf = 1 << 20
        Xf = hddm_lb.synthetic_points(nf, 59, 78)
        sf = np.random.default_rng(2025).integers(0, 4, size=nf).astype(np.int32)
        xf = torch.from_numpy(Xf).cuda()
        sdf = torch.from_numpy(sf).cuda()
        gf = torch.from_numpy(np.random.default_rng(96).standard_normal((nf, 128))).cuda()
        yf = torch.empty((nf, 128), dtype=torch.float64, device="cuda")
        dxf = torch.empty((nf, 59), dtype=torch.float64, device="cuda")
        pf, pa = legs["mixed"]

        def fstep():
            ctx.forward(xf, yf, sdf, pf)
            ctx.adjoint(xf, gf, dxf, ds, sdf, pa)
        # one warm-up step: the first forward of a new size class is tuned
        fms, _ = timed_steps(fstep, 1, 1, None, label="D=59 mixed, all 2^20 points")
        full = {"points": nf, "ms_per_step": fms, "value": nf * 128 / (fms * 1e-3)}
        del xf, sdf, gf, yf, dxf
    res["full_batch_check"] = full
    # executed work, on statistics passes outside the timed steps: non-zero
    # (point, node) pairs and walk ops of the wave-specialized F32 forward,
    # 32-node chunks of the mixed adjoint
    ctx.set_option(hddm_lb.OPT_STATS, 1)
    ctx.set_option(hddm_lb.OPT_FORMULATION, 2)
    ctx.forward(x, y, sd, hddm_lb.F32)
    st = ctx.last_stats(6)
    ctx.set_option(hddm_lb.OPT_FORMULATION, 0)
    ctx.adjoint(x, g, dx, ds, sd, hddm_lb.F32)
    ast = ctx.last_adjoint_stats(9)
    ctx.set_option(hddm_lb.OPT_STATS, 0)
    ctx.close()
    return res, {"nnz": st[5], "walk_ops": st[0], "adjoint_chunks32": ast[8], "adjoint_items": ast[7],
                 "nno": states[0][0].shape[0]}


def single_state_d16(args):
    """Single-state forward: D=16, ~60k nodes (seeded refinement p=0.3 from
    level sum <= 2, max level 7), dofs=32, 2^20 points."""
    import numpy as np
    import torch
    import hddm_lb
    ii, jj, lv, _ = hddm_lb.generate_grid(16, 2, 7, 0.3, 1, 60000)
    S = hddm_lb.synthetic_surplus(lv, 32, 2)
    n = 1 << 20
    X = hddm_lb.synthetic_points(n, 16, 3)
    ctx = hddm_lb.Context(0)
    ctx.set_option(hddm_lb.OPT_AUTOTUNE, 1)
    ctx.upload_state(0, ii, jj, S)
    x = torch.from_numpy(X).cuda()
    y = torch.empty((n, 32), dtype=torch.float64, device="cuda")
    out = {"workload": "single-state forward, D=16, %d nodes, dofs=32, %d points" % (ii.shape[0], n)}
    oracle = oracle_mod()
    pick = np.random.default_rng(5).choice(n, size=256, replace=False)
    yo = oracle.forward(ii, jj, S, X[pick])
    for name, pr in (("mixed", hddm_lb.F32), ("fp64", hddm_lb.FP64)):
        ms, kern = timed_steps(lambda: ctx.forward(x, y, None, pr), max(3, args.steps // 2), 2, ctx)
        out[name] = {"ms_per_step": ms, "sweep_ms": kern[0]["ms_per_launch"], "value": n * 32 / (ms * 1e-3),
                     "precision": "f32" if name == "mixed" else "fp64", "config": ctx.last_config(),
                     "max_rel_err": normwise_err(y.cpu().numpy()[pick], yo)}
    out["mixed_over_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    # the same forward through host buffers: x in and y out over PCIe
    yh = np.empty((n, 32))
    ctx.forward_host(X, yh, None, hddm_lb.F32)
    t0 = time.perf_counter()
    ctx.forward_host(X, yh, None, hddm_lb.F32)
    host_ms = (time.perf_counter() - t0) * 1e3
    out["host_buffers_f32"] = {"ms_per_call": host_ms, "value": n * 32 / (host_ms * 1e-3),
                               "note": "x and y in pageable host memory, copies included"}
    ctx.close()
    return out, (ii, jj, S, X)


def multistate_d20(args):
    """12 states, D=20, adaptive grids (max level 6) with node counts
    log-uniform in [2k, 200k], dofs per state from {16, 32, 64}, 262,144
    points with seeded state ids, one launch."""
    import numpy as np
    import torch
    import hddm_lb
    rng = np.random.default_rng(2024)
    nnos = np.exp(rng.uniform(np.log(2000), np.log(200000), size=12)).astype(int)
    dofs = rng.choice([16, 32, 64], size=12)
    ctx = hddm_lb.Context(0)
    ctx.set_option(hddm_lb.OPT_AUTOTUNE, 1)
    states = []
    for s in range(12):
        ii, jj, lv, _ = hddm_lb.generate_grid(20, 1, 6, 0.3, 300 + s, int(nnos[s]))
        S = hddm_lb.synthetic_surplus(lv, int(dofs[s]), 400 + s)
        ctx.upload_state(s, ii, jj, S)
        states.append((ii, jj, S))
    n = 262144
    X = hddm_lb.synthetic_points(n, 20, 77)
    w = rng.uniform(0.5, 1.5, size=12)
    sid = rng.choice(12, size=n, p=w / w.sum()).astype(np.int32)
    work = float(dofs[sid].sum())
    x = torch.from_numpy(X).cuda()
    sd = torch.from_numpy(sid).cuda()
    y = torch.empty((n, 64), dtype=torch.float64, device="cuda")
    out = {"workload": "12 states, D=20, %d nodes in all, dofs 16/32/64, %d points"
                       % (sum(s[0].shape[0] for s in states), n)}
    oracle = oracle_mod()
    for name, pr in (("mixed", hddm_lb.F32), ("fp64", hddm_lb.FP64)):
        ms, _ = timed_steps(lambda: ctx.forward(x, y, sd, pr), max(3, args.steps // 2), 2)
        Y = y.cpu().numpy()
        e = 0.0
        r2 = np.random.default_rng(6)
        for s, (ii, jj, S) in enumerate(states):
            pick = r2.choice(np.flatnonzero(sid == s), size=8, replace=False)
            e = max(e, normwise_err(Y[pick, :S.shape[1]], oracle.forward(ii, jj, S, X[pick])))
        out[name] = {"ms_per_step": ms, "value": work / (ms * 1e-3), "config": ctx.last_config(), "max_rel_err": e}
    out["mixed_over_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    ctx.close()
    return out


def forward_adjoint_d20(args):
    """Forward + adjoint on the differentiated workload: D=20 adaptive grid of
    ~100k nodes (max level 6), dofs=32, 65,536 points in [0.01, 0.99]^20 kept
    1e-3 off every kink, upstream gradient N(0, 1)."""
    import numpy as np
    import torch
    import hddm_lb
    ii, jj, lv, _ = hddm_lb.generate_grid(20, 1, 6, 0.3, 404, 100000)
    S = hddm_lb.synthetic_surplus(lv, 32, 405)
    n = 65536
    X = hddm_lb.synthetic_points(n, 20, 406, 0.01, 0.99, kink_level=6, kink_gap=1e-3)
    G = np.random.default_rng(407).standard_normal((n, 32))
    ctx = hddm_lb.Context(0)
    ctx.set_option(hddm_lb.OPT_AUTOTUNE, 1)
    ctx.upload_state(0, ii, jj, S)
    x = torch.from_numpy(X).cuda()
    g = torch.from_numpy(G).cuda()
    y = torch.empty((n, 32), dtype=torch.float64, device="cuda")
    dx = torch.empty((n, 20), dtype=torch.float64, device="cuda")
    _, total = ctx.surplus_offset(-1)
    ds = torch.empty((total,), dtype=torch.float64, device="cuda")
    out = {"workload": "forward + adjoint, D=20, %d nodes, dofs=32, %d points" % (ii.shape[0], n)}
    oracle = oracle_mod()
    pick = np.random.default_rng(408).choice(n, size=32, replace=False)
    dxo, _ = oracle.adjoint(ii, jj, S, X[pick], G[pick])
    yo = oracle.forward(ii, jj, S, X[pick])
    for name, (pf, pa) in (("mixed", (hddm_lb.F32, hddm_lb.F32)), ("fp64", (hddm_lb.FP64, hddm_lb.FP64))):
        def step(pf=pf, pa=pa):
            ctx.forward(x, y, None, pf)
            ctx.adjoint(x, g, dx, ds, None, pa)
        ms, kern = timed_steps(step, max(3, args.steps // 2), 2, ctx)
        out[name] = {"ms_per_step": ms, "forward_sweep_ms": kern[0]["ms_per_step"],
                     "adjoint_kernel_ms": kern[1]["ms_per_step"], "value": n * 32 / (ms * 1e-3),
                     "max_rel_err": {"y": normwise_err(y.cpu().numpy()[pick], yo),
                                     "dx": normwise_err(dx.cpu().numpy()[pick], dxo)}}
    out["mixed_over_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    ctx.close()
    return out


def legacy_per_point(reps):
    """Interpolator::interpolate through the drop-in plugin, one point per
    call (the reference host's API), timed by the stand-in host over `reps`
    calls on the reference's own fixture (surplus.plt: D=59, 118 dofs)."""
    host = os.path.join(ROOT, "tests", "plugin", "_build", "plugin_host")
    plugin = os.path.join(ROOT, "hddm-solver-postprocessors_amd", "lib", "libpostprocessor.so")
    if not (os.path.exists(host) and os.path.exists(plugin)):
        return None
    res = {}
    with tempfile.TemporaryDirectory() as d:
        import shutil
        shutil.copy(os.path.join(ROOT, "tests", "golden", "surplus.plt"), os.path.join(d, "surplus.plt"))
        for prec in ("fp64", "f16x2"):
            with open(os.path.join(d, "hddm-solver.cfg"), "w") as f:
                f.write("n_agents 59\npriorityHIP 4\nprecisionHIP %s\n" % prec)
            p = subprocess.run([host, plugin, "latency", "surplus.plt", str(reps)], cwd=d,
                               capture_output=True, text=True, timeout=300)
            us = None
            for line in p.stdout.splitlines():
                f = line.split()
                if f and f[0] == "us_per_call":
                    us = float(f[1])
            res[prec] = us
    return {"us_per_call": res, "calls": reps, "points_per_call": 1, "dofs": 118,
            "reference_cuda_us_per_state": REF_CUDA_US_PER_STATE,
            "note": "host pointers in and out, synchronous, as Interpolator::interpolate is called"}


def cpu_baseline(ii, jj, S, X, target_s, what):
    """The reference's x86 xps/chains algorithm (the oracle's port of it; the
    reference has no adjoint, so the forward) on the host cores: OpenMP over
    points on a prefix sized to about target_s seconds, plus one thread on a
    fixed 128-point subset."""
    oracle = oracle_mod()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    st = oracle.X86State(ii, jj, S)
    # rate probe long enough to be stable (~0.5 s)
    probe = 256
    while True:
        t0 = time.perf_counter()
        st.forward(X[:probe], threads)
        dt = time.perf_counter() - t0
        if dt > 0.5 or probe >= X.shape[0]:
            break
        probe = min(X.shape[0], probe * 4)
    m = int(min(X.shape[0], max(probe, probe * target_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    yc = st.forward(X[:m], threads)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    st.forward(X[:128], 1)
    dt1 = time.perf_counter() - t0
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    gold = oracle.forward(ii, jj, S, X[:16])
    return {
        "value": m * S.shape[1] / dt,
        "unit": "point·dof/s",
        "cores": threads,
        "kind": "port",
        "sample": "first %d of the %d points of %s: forward only, x86 xps/chains algorithm (FP64), OpenMP "
                  "static over points, %.1f s" % (m, X.shape[0], what, dt),
        "one_thread": {"value": 128 * S.shape[1] / dt1, "sample": "first 128 points, 1 thread, %.2f s" % dt1},
        "cpu_model": model,
        "port_vs_gold_max_rel_err": normwise_err(yc[:16], gold),
    }


def cpu_regular_d4():
    """The plain FP64 CPU case: regular grid D=4, sum(l-1) <= 4 (401 nodes),
    dofs=8, 4,096 points; the x86 port on one thread and the gold
    restatement on one thread."""
    import numpy as np
    import hddm_lb
    oracle = oracle_mod()
    ii, jj, lv, _ = hddm_lb.generate_grid(4, 4, 5, 0.0, 1, 0)
    S = hddm_lb.synthetic_surplus(lv, 8, 2)
    X = hddm_lb.synthetic_points(4096, 4, 3)
    st = oracle.X86State(ii, jj, S)
    t0 = time.perf_counter()
    for _ in range(5):
        st.forward(X, 1)
    dtx = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    oracle.forward(ii, jj, S, X, 1)
    dtg = time.perf_counter() - t0
    return {"workload": "regular grid D=4, %d nodes, dofs=8, 4096 points" % ii.shape[0],
            "x86_port_one_thread": 4096 * 8 / dtx, "gold_one_thread": 4096 * 8 / dtg, "unit": "point·dof/s"}


def roofline_of(head, work, npoints):
    """Roofline of the dominant kernel of the mixed step, from algorithmic
    FLOPs (non-zero (point, node) pairs only) over its HIP-event duration."""
    mixed = head["mixed"]
    fwd_ms = mixed["kernels"][0]["ms_per_launch"]
    adj_ms = mixed["kernels"][1]["ms_per_launch"]
    nnz = float(work["nnz"])
    dofs = 128
    # adjoint: dS (2 nnz dofs) and g.S (2 nnz dofs); forward: y (2 nnz dofs)
    adj_flops = 4.0 * nnz * dofs
    fwd_flops = 2.0 * nnz * dofs
    # executed MFMA work of the adjoint kernel: per 32-node chunk and 32-dof
    # tile (4 tiles), 24 v_mfma_f32_32x32x16_bf16 for dS (4 steps of 16 points
    # x 6 terms of the three-part splits) and 24 for g.S (2 point halves x 2
    # steps of 16 dofs x 6 terms), 32,768 FLOP each, on the bf16 units
    exec_flops = float(work["adjoint_chunks32"]) * 4 * 48 * 32768
    dominant = "adjoint_ws_kernel (f32)" if adj_ms >= fwd_ms else "forward sweep (f32)"
    if adj_ms >= fwd_ms:
        t, flops, peak = adj_ms, adj_flops, PEAK_TFLOPS["bf16x3"]
    else:
        t, flops, peak = fwd_ms, fwd_flops, PEAK_TFLOPS["f32"]
    achieved = flops / (t * 1e-3) / 1e12
    # HBM bytes per launch: the rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (calibrated on a copy kernel) of the same kernel on the same workload,
    # per point, times this launch's points
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tk = json.load(f).get("kernels", {}).get("adjoint_f32_d59x4" if adj_ms >= fwd_ms else "sweep_f32_d59x4")
        if tk:
            traffic = tk["bytes_per_point"] * npoints
    return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": traffic, "kernel": dominant, "kernel_ms": t, "flops_per_launch": flops,
            "flops_basis": "non-zero (point, node) pairs: %d; per pair and dof 2 FLOP (forward), 4 (adjoint)" % int(nnz),
            "frac_executed": exec_flops / (adj_ms * 1e-3) / 1e12 / PEAK_TFLOPS["bf16x3"],
            "flops_dense_per_launch": 4.0 * work["nno"] * dofs * npoints,
            "note": "frac: algorithmic FLOPs of the non-zero pairs only; frac_executed: the MFMA work the "
                    "adjoint kernel issues on its 32-node chunks (a 64-point block contracts every node "
                    "that is non-zero at any of its points). The F32 adjoint's products run on the bf16 "
                    "MFMA units (peak 2,500 TFLOP/s dense) with each F32 operand split into three bf16 "
                    "parts, six bf16 products per F32-accurate one (an F32-accurate peak of 417 TFLOP/s)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--points", type=int, default=1 << 18,
                    help="points per step of the headline (all four states)")
    ap.add_argument("--no-full-check", dest="full_check", action="store_false",
                    help="skip the one step over all 2^20 points")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-secondary", action="store_true", help="skip the secondary workloads")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the error checks against the oracle (profiling runs)")
    ap.add_argument("--latency-calls", type=int, default=2000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device 3 times (calibrates profiler byte counters)")
    ap.add_argument("--only", choices=["single_state_d16", "multistate_d20x12", "forward_adjoint_d20"],
                    help="run one secondary workload alone and print its JSON (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a forward shards by points; "
              "run one process per GPU on its own share). Use --gpus 1.", file=sys.stderr)
        return 2

    import torch
    import hddm_lb  # noqa: F401

    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.only:
        fn = {"single_state_d16": lambda: single_state_d16(args)[0], "multistate_d20x12": lambda: multistate_d20(args),
              "forward_adjoint_d20": lambda: forward_adjoint_d20(args)}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return 0

    head, work = headline(args)
    n = args.points
    mixed_ms = head["mixed"]["ms_per_step"]
    fp64_ms = head["fp64"]["ms_per_step"]
    out = {
        "metric": METRIC,
        "value": n * 128 / (mixed_ms * 1e-3),
        "unit": "point·dof/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded regular grids, surpluses U[-1,1]*2^-sum(l), points U[0,1)^59, gbar N(0,1))",
        "config": {"workload": "multistate forward + adjoint, 4 states, D=59, regular grids sum(l-1) <= 3 "
                               "(%d nodes each), dofs=128, 2^20 points in steps of %d" % (work["nno"], n),
                   "states": 4, "dim": 59, "nno_per_state": work["nno"], "dofs": 128, "points": n,
                   "points_total": 1 << 20,
                   "forward_precision": "f32 (FP32 MFMA, FP64 fold per chunk)",
                   "adjoint_precision": "f32 (operands split into three bf16 parts, six-term products on the "
                                        "bf16 MFMA; per-chunk F32 sums, FP64 beyond)",
                   "forward_config": head["mixed"]["forward_config"]},
        "roofline": roofline_of(head, work, n),
        "fp64_plain": {"value": n * 128 / (fp64_ms * 1e-3), "ms_per_step": fp64_ms, "steps": head["fp64"]["steps"],
                       "kernels": head["fp64"]["kernels"], "forward_config": head["fp64"]["forward_config"],
                       "max_rel_err": head["fp64"]["max_rel_err"]},
        "mixed_over_fp64": fp64_ms / mixed_ms,
        "mixed_kernels": head["mixed"]["kernels"],
        "max_rel_err": head["mixed"]["max_rel_err"],
        "executed": work,
        "full_batch_check": head["full_batch_check"],
        "forward_f16x2": head["forward_f16x2"],
    }
    say("bench: headline done")
    if not args.no_secondary:
        out["single_state_d16"], c2 = single_state_d16(args)
        say("bench: single-state done")
        out["multistate_d20x12"] = multistate_d20(args)
        say("bench: multistate done")
        out["forward_adjoint_d20"] = forward_adjoint_d20(args)
        say("bench: forward+adjoint D=20 done")
        out["legacy_per_point"] = legacy_per_point(args.latency_calls)
        say("bench: per-point latency done")
    else:
        c2 = None
    if not args.no_cpu:
        import numpy as np
        import hddm_lb as h
        # the headline workload's state 0 and its points
        ii, jj, lv, _ = h.generate_grid(59, 3, 4, 0.0, 500, 0)
        S = h.synthetic_surplus(lv, 128, 600)
        X = h.synthetic_points(n, 59, 77)
        sid = np.random.default_rng(2024).integers(0, 4, size=n).astype(np.int32)
        out["cpu_baseline"] = cpu_baseline(ii, jj, S, np.ascontiguousarray(X[sid == 0]), args.cpu_seconds,
                                           "state 0 of the D=59 workload (281,077 nodes, 128 dofs)")
        if c2 is not None:
            out["cpu_baseline"]["single_state_d16"] = cpu_baseline(*c2, max(2.0, args.cpu_seconds / 4),
                                                                   "the single-state D=16 forward")
        out["cpu_baseline"]["regular_d4"] = cpu_regular_d4()
        say("bench: CPU baseline done")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
