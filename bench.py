#!/usr/bin/env python3
"""Throughput of the Matched-TS per-time-step path on one MI355X.

Workload: Crank-Nicolson LOD with MIB-L2 on [-1,1]^3 with 257^3 nodes,
interface the ellipsoid x^2/0.49 + y^2/0.25 + z^2/0.16 = r^2 (r = 1 - 2^-20,
which moves the surface off the grid node (0, 0.5, 0)), beta- = 10 (1 + 0.5
sin x cos y), beta+ = 1, manufactured solution u- = e^-t sin x sin y sin z,
u+ = e^-t cos x cos y cos z, dt = h = 1/128, T = 0.5 (64 steps), FP64.

One step = one Solve_2nd: three line stages (x, y, z) and the boundary
faces.  The field and the MIB table are resident in HBM before timing starts.
The timed region is exactly K steps between two stream synchronizations, with
nothing but the steps' launches in it.  The same K steps then run again with
a pair of HIP events around every launch (on the instance's stream) for the
per-stage times and the roofline kernel's average duration.

--config adi257 is the Douglas-ADI target workload: MIB-L1, 257^3 nodes,
sphere r = 0.5 (pulled off the grid node by 2^-20, relative), beta 10/1,
dt = h = 1/128.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))


def load_mts():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    return ge._load_package()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--config", default="cfg1", choices=["cfg0", "cfg1", "cfg2", "adi257"],
                    help="cfg1 (default): LOD-CN MIB-L2 257^3 ellipsoid; cfg0: Douglas-ADI "
                         "MIB-L1 33^3 sphere r=0.5, beta 10/1, dt=h, T=1; cfg2: TS MIB-L2 513^3 "
                         "bumpy sphere, beta 1000/1, dt=h, T=0.25; adi257: Douglas-ADI MIB-L1 "
                         "257^3 sphere r=0.5, beta 10/1, dt=h, T=0.5")
    ap.add_argument("--n", type=int, default=None, help="intervals per axis (overrides the config)")
    ap.add_argument("--method", default=None, choices=["adi", "lod-ie", "lod-cn", "ts"])
    ap.add_argument("--mib", type=int, default=None, choices=[1, 2])
    ap.add_argument("--fp32-form", type=int, default=1,
                    help="also time the FP32 form on the same workload (0 to skip)")
    ap.add_argument("--cpu-baseline", type=int, default=1,
                    help="time the checker restatement on the host (0 to skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counter readings)")
    ap.add_argument("--workload", default="step", choices=["step", "batch", "serve"],
                    help="step: the single-solve line above (default); batch: uneven "
                         "independent solves, interleaved vs one after another; serve: "
                         "Poisson request stream through the serving loop")
    ap.add_argument("--solves", type=int, default=96, help="batch: number of solves")
    ap.add_argument("--requests", type=int, default=2000, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: arrivals per second")
    ap.add_argument("--max-inflight", type=int, default=32, help="serve: in-flight limit")
    ap.add_argument("--hbm-cap-gb", type=float, default=200.0, help="serve: admission cap")
    ap.add_argument("--n-max", type=int, default=256, help="serve: largest N drawn")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tune-steps", type=int, default=10,
                    help="steps the line-split tuner times per split")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="batch/serve: the field precision of the solves")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs the per-step path on a single GPU; "
              "multi-GPU runs are not part of it (use --gpus 1)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        return copy_calibration()
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)

    mts = load_mts()
    # the single-solve configurations; surfaces moved off grid nodes by 2^-20 (relative)
    cfgs = {
        "cfg1": dict(n=256, method="lod-cn", mib=2, T=0.5, beta=("sincos", 10.0, 1.0),
                     surf=("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20),
                     label="ellipsoid(0.7,0.5,0.4) beta-=10(1+sin x cos y/2) beta+=1"),
        "cfg0": dict(n=32, method="adi", mib=1, T=1.0, beta=("const", 10.0, 1.0),
                     surf=("ellipsoid", 1.0, 1.0, 1.0, 0.5 * (1 - 2 ** -20)),
                     label="sphere r=0.5 beta-=10 beta+=1"),
        "cfg2": dict(n=512, method="ts", mib=2, T=0.25, beta=("const", 1000.0, 1.0),
                     surf=("bumpy", 0.5 * (1 - 2 ** -20), 0.1, 5, 3),
                     label="bumpy sphere r=0.5+0.1 sin5t cos3p beta-=1000 beta+=1"),
        "adi257": dict(n=256, method="adi", mib=1, T=0.5, beta=("const", 10.0, 1.0),
                       surf=("ellipsoid", 1.0, 1.0, 1.0, 0.5 * (1 - 2 ** -20)),
                       label="sphere r=0.5 beta-=10 beta+=1"),
    }
    cf = cfgs[args.config]
    args.n = args.n or cf["n"]
    args.method = args.method or cf["method"]
    args.mib = args.mib or cf["mib"]
    n = args.n + 1
    dt = 2.0 / args.n
    T = cf["T"]
    phys = mts.physics(*cf["beta"])
    surf = mts.surface(cf["surf"][0], *cf["surf"][1:])
    t0 = time.perf_counter()
    st = mts.Setup(mts.grid(n), surf, phys, mib=args.mib)
    setup_s = time.perf_counter() - t0
    info = st.info()
    ctx = mts.Context(0)
    dev_name, _ = ctx.device_info()
    # the line split: every split that fits, timed on a scratch instance
    # (before and outside the timed region); the instance below runs the
    # fastest
    t0 = time.perf_counter()
    split = ctx.tune_lanes(st, args.method, dt, "fp64", steps=args.tune_steps)
    tune_s = time.perf_counter() - t0
    inst = mts.Instance(ctx, st, args.method, dt)

    # warmup: untimed steps from the initial condition
    inst.initialize(0.0)
    t = 0.0
    for _ in range(args.warmup):
        t += dt
        inst.step(t)
    inst.sync()

    # timed run: K steps of the reference time loop (t += dt) from t0 = 0, with
    # nothing but the steps' own launches between the two synchronizations
    inst.initialize(0.0)
    inst.sync()
    t = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t += dt
        inst.step(t)
    inst.sync()
    elapsed = time.perf_counter() - t_start
    linf, l2, _ = inst.error(t)
    diverged = not (math.isfinite(l2) and linf < 1e6)

    # the same K steps again with a pair of HIP events around every launch (on
    # the instance's stream): per-stage times and the roofline kernel's average
    inst.initialize(0.0)
    inst.profile(True)
    inst.profile_reset()
    inst.sync()
    tp = 0.0
    for _ in range(args.steps):
        tp += dt
        inst.step(tp)
    inst.sync()
    stages = inst.profile_read()
    inst.profile(False)

    # the same K steps as one captured graph (the first run captures it)
    t_graph = (args.steps + 0.5) * dt
    inst.initialize(0.0)
    inst.run(0.0, t_graph, graph=True)
    inst.initialize(0.0)
    inst.sync()
    t_start = time.perf_counter()
    inst.run(0.0, t_graph, graph=True)
    inst.sync()
    graph_ms = (time.perf_counter() - t_start) / args.steps * 1e3

    nodes = n ** 3
    value = nodes * args.steps / elapsed
    # dominant stage: largest total time; algorithmic bytes = one FP64 read and
    # one FP64 write of every node (each line stage reads its input field once
    # and writes its output once; the Douglas y/z stages also read u^n)
    per_node = {"adi_y": 24, "adi_z": 24, "adi_x": 16, "adi_rhs1": 16}
    line_stages = {k: v for k, v in stages.items() if k != "faces" and v[0] > 0}
    dom = max(line_stages, key=lambda k: line_stages[k][1])
    launches, total_ms = line_stages[dom]
    avg_s = total_ms / 1e3 / launches
    bytes_per_launch = per_node.get(dom, 16) * nodes
    achieved = bytes_per_launch / avg_s / 1e9
    peak = 8000.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        key = f"{args.method}_mib{args.mib}_n{args.n}_{dom}"
        if key in tr:
            traffic = tr[key]

    fp32 = None
    if args.fp32_form:
        fp32 = fp32_form(mts, ctx, st, args, dt, nodes, value, linf, diverged)

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(mts, st, args, cf, dt, nodes, phys)

    line = {
        "metric": "grid-pt-steps/s at fixed L-inf error (FP32+batch vs FP64); solves/s + p50/p99",
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.method.upper()} MIB-L{args.mib} {n}^3 {cf['label']}"
                        f" dt=h={dt}, {args.steps} steps from t=0 to t={t:.6g}"
                        f" (the configuration's T={T} is {int(T / dt)} steps)",
            "config": args.config,
            "diverged": diverged,
            "grid_points": nodes,
            "crossings": [int(info.n_cross[d]) for d in range(3)],
            "t_end": t,
            "linf_error": linf,
            "l2_error": l2,
            "setup_s": setup_s,
            "device": dev_name,
            "stage_ms_per_launch": {k: v[1] / max(v[0], 1) for k, v in stages.items()},
            "graph_ms_per_step": graph_ms,
            "line_split": {"min_lanes": split[0], "gen_lanes": split[1],
                           "tuner_ms_per_step": split[2], "tuner_s": tune_s},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": peak,
            "unit": "GB/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": avg_s * 1e6,
        },
        "cpu_baseline": cpu,
        "fp32_form": fp32,
    }
    print(json.dumps(line), flush=True)
    return 0


def fp32_form(mts, ctx, st, args, dt, nodes, fp64_value, fp64_linf, fp64_diverged):
    """The same K steps in the FP32 form (FP32 fields and line solve, rows
    formed in FP64), with its L-inf error against the exact solution.  When
    either form diverged (the L-inf ratio then compares overflow artefacts),
    the ratios are null and "diverged" says so."""
    split = ctx.tune_lanes(st, args.method, dt, "fp32", steps=args.tune_steps)
    inst = mts.Instance(ctx, st, args.method, dt, precision="fp32")
    inst.initialize(0.0)
    t = 0.0
    for _ in range(args.warmup):
        t += dt
        inst.step(t)
    inst.sync()
    inst.initialize(0.0)
    inst.sync()
    t = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t += dt
        inst.step(t)
    inst.sync()
    elapsed = time.perf_counter() - t_start
    linf, l2, _ = inst.error(t)
    inst.initialize(0.0)
    inst.profile(True)
    inst.profile_reset()
    inst.sync()
    tp = 0.0
    for _ in range(args.steps):
        tp += dt
        inst.step(tp)
    inst.sync()
    stages = inst.profile_read()
    inst.profile(False)
    inst.close()
    value = nodes * args.steps / elapsed
    diverged = fp64_diverged or not (math.isfinite(l2) and linf < 1e6)
    return {
        "value": value,
        "unit": "grid-pt-steps/s",
        "ms_per_step": elapsed / args.steps * 1e3,
        "dtype": "f32 fields, rows and solve; f64 MIB rows, setup and norms",
        "line_split": {"min_lanes": split[0], "gen_lanes": split[1],
                       "tuner_ms_per_step": split[2]},
        "linf_error": linf,
        "l2_error": l2,
        "diverged": diverged,
        "linf_vs_fp64": None if diverged else linf / fp64_linf,
        "speedup_vs_fp64": None if diverged else value / fp64_value,
        "stage_ms_per_launch": {k: v[1] / max(v[0], 1) for k, v in stages.items()},
    }


METRIC = "grid-pt-steps/s at fixed L-inf error (FP32+batch vs FP64); solves/s + p50/p99"


def draw_solve(rng, n, method, mib, T):
    """One solve of the uneven mix: interface drawn from sphere r ~ U[0.3, 0.7],
    ellipsoid semi-axes ~ U[0.3, 0.8] or torus R = 0.5, r ~ U[0.15, 0.25];
    beta+ = 1 and beta- log-uniform in [1e-3, 1e3]; dt = h; t from 0 to T."""
    kind = int(rng.integers(3))
    if kind == 0:
        surf = ("ellipsoid", (1.0, 1.0, 1.0, float(rng.uniform(0.3, 0.7))))
    elif kind == 1:
        surf = ("ellipsoid", tuple(float(v) for v in rng.uniform(0.3, 0.8, 3)) + (1.0,))
    else:
        surf = ("torus", (0.5, float(rng.uniform(0.15, 0.25))))
    beta_minus = float(10.0 ** rng.uniform(-3.0, 3.0))
    return {"n": n, "surface": surf, "beta": (beta_minus, 1.0), "method": method, "mib": mib,
            "dt": 2.0 / n, "T": T}


def describe(d):
    return (f"{d['method']} MIB-L{d['mib']} N={d['n']} {d['surface'][0]}"
            f"{tuple(round(v, 4) for v in d['surface'][1])} beta-={d['beta'][0]:.4g}")


def bench_batch(args):
    """Uneven independent solves (N in {32, 48, ..., 192}, Douglas-ADI or
    LOD-CN, MIB-L1, T = 0.5, FP64): every solve's time loop run one after
    another with mts_run, then all of them interleaved with mts_batch_run.
    Setup and upload happen before either timed region."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    mts = load_mts()
    rng = np.random.default_rng(args.seed)
    draws = [draw_solve(rng, int(rng.choice(np.arange(32, 193, 16))),
                        ["adi", "lod-cn"][int(rng.integers(2))], 1, 0.5)
             for _ in range(args.solves)]
    ctx = mts.Context(0)

    def make(d, precision):
        phys = mts.physics("const", *d["beta"])
        prec = precision
        try:
            st = mts.Setup(mts.grid(d["n"] + 1), mts.surface(d["surface"][0], *d["surface"][1]),
                           phys, mib=d["mib"])
            return mts.Instance(ctx, st, d["method"], d["dt"], precision=prec), None
        except mts.MtsError as e:
            return None, e.code

    t0 = time.perf_counter()
    with ThreadPoolExecutor(16) as ex:
        made = list(ex.map(lambda d: make(d, "fp64"), draws))
    setup_s = time.perf_counter() - t0
    refused = {}
    insts, ds = [], []
    for d, (inst, code) in zip(draws, made):
        if inst is None:
            refused[str(code)] = refused.get(str(code), 0) + 1
        else:
            insts.append(inst)
            ds.append(d)
    work = sum((d["n"] + 1) ** 3 * int(d["T"] / d["dt"]) for d in ds)

    for inst in insts:
        inst.initialize(0.0)
        inst.sync()
    # a loop of single solves: each one's time loop runs to completion before the next
    t0 = time.perf_counter()
    ends = []
    for inst, d in zip(insts, ds):
        ends.append(inst.run(0.0, d["T"])[1])
        inst.sync()
    loop_s = time.perf_counter() - t0
    loop_err = [inst.error(t_end)[0] for inst, t_end in zip(insts, ends)]

    for inst in insts:
        inst.initialize(0.0)
        inst.sync()
    t0 = time.perf_counter()
    out = mts.batch_run(insts, 0.0, [d["T"] for d in ds])
    batch_s = time.perf_counter() - t0
    same = True
    for inst, (steps, t_end), le, te in zip(insts, out, loop_err, ends):
        same &= t_end == te and inst.error(t_end)[0] == le
    # the same in lockstep, one launch per line sweep across the solves
    for inst in insts:
        inst.initialize(0.0)
        inst.sync()
    t0 = time.perf_counter()
    outm = mts.batch_run(insts, 0.0, [d["T"] for d in ds], multi=True)
    multi_s = time.perf_counter() - t0
    for inst, (steps, t_end), le, te in zip(insts, outm, loop_err, ends):
        same &= t_end == te and inst.error(t_end)[0] == le
    # the same solves batched in the FP32 form
    f32 = None
    if args.precision == "fp32":
        with ThreadPoolExecutor(16) as ex:
            made32 = list(ex.map(lambda d: make(d, "fp32"), ds))
        insts32 = [m_[0] for m_ in made32]
        for inst in insts32:
            inst.initialize(0.0)
            inst.sync()
        t0 = time.perf_counter()
        out32 = mts.batch_run(insts32, 0.0, [d["T"] for d in ds])
        f32_s = time.perf_counter() - t0
        # each solve's time loop as one graph: the first run captures, the second is timed
        for inst in insts32:
            inst.initialize(0.0)
        mts.batch_run(insts32, 0.0, [d["T"] for d in ds], graph=True)
        for inst in insts32:
            inst.initialize(0.0)
            inst.sync()
        t0 = time.perf_counter()
        mts.batch_run(insts32, 0.0, [d["T"] for d in ds], graph=True)
        f32g_s = time.perf_counter() - t0
        for inst in insts32:
            inst.initialize(0.0)
            inst.sync()
        t0 = time.perf_counter()
        mts.batch_run(insts32, 0.0, [d["T"] for d in ds], multi=True)
        f32m_s = time.perf_counter() - t0
        ratios = []
        for inst, (steps, t_end), le in zip(insts32, out32, loop_err):
            l32 = inst.error(t_end)[0]
            if le > 0 and le < 1.0:  # runs whose error is a discretization error
                ratios.append(l32 / le)
            inst.close()
        f32 = {"batched_s": f32_s, "value": work / f32_s, "speedup_vs_fp64_loop": loop_s / f32_s,
               "speedup_vs_fp64_batched": batch_s / f32_s,
               "graph_batched_s": f32g_s, "graph_value": work / f32g_s,
               "graph_speedup_vs_fp64_loop": loop_s / f32g_s,
               "multi_batched_s": f32m_s, "multi_value": work / f32m_s,
               "multi_speedup_vs_fp64_loop": loop_s / f32m_s,
               "linf_ratio_max": max(ratios) if ratios else None,
               "linf_ratio_median": float(np.median(ratios)) if ratios else None,
               "solves_compared": len(ratios)}
    line = {
        "metric": METRIC,
        "value": work / batch_s,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.solves} uneven solves from seed {args.seed}: N in "
                        "{32,48,...,192}, sphere/ellipsoid/torus, beta-/beta+ log-uniform "
                        "in [1e-3,1e3], Douglas-ADI or LOD-CN, MIB-L1, dt=h, T=0.5",
            "solves_run": len(insts),
            "solves_refused": refused,
            "grid_pt_steps": work,
            "host_setup_s": setup_s,
        },
        "batched_s": batch_s,
        "one_after_another_s": loop_s,
        "one_after_another_value": work / loop_s,
        "speedup": loop_s / batch_s,
        "multi_batched_s": multi_s,
        "multi_value": work / multi_s,
        "multi_speedup": loop_s / multi_s,
        "same_linf_batched_and_alone": bool(same),
        "max_linf": max(loop_err) if loop_err else None,
        "fp32_form": f32,
    }
    for inst in insts:
        inst.close()
    print(json.dumps(line), flush=True)
    return 0


def bench_serve(args):
    """Poisson request stream through the serving loop: N uniform in
    [24, n_max], T in {0.1, 0.25, 0.5, 1.0}, any method, MIB-L1 or L2, the
    interface mix of the batch workload; admission capped by HBM bytes."""
    import numpy as np

    mts = load_mts()
    rng = np.random.default_rng(args.seed)
    draws = [draw_solve(rng, int(rng.integers(24, args.n_max + 1)),
                        ["adi", "lod-ie", "lod-cn", "ts"][int(rng.integers(4))],
                        int(rng.integers(1, 3)), float(rng.choice([0.1, 0.25, 0.5, 1.0])))
             for _ in range(args.requests)]
    gaps = rng.exponential(1.0 / args.rate, args.requests)
    ctx = mts.Context(0)
    srv = mts.Server(ctx, int(args.hbm_cap_gb * (1 << 30)), max_inflight=args.max_inflight,
                     setup_threads=16)
    t0 = time.perf_counter()
    next_t = t0
    for i, d in enumerate(draws):
        next_t += gaps[i]
        now = time.perf_counter()
        if next_t > now:
            time.sleep(next_t - now)
        srv.submit(mts.request(mts.grid(d["n"] + 1),
                               mts.surface(d["surface"][0], *d["surface"][1]),
                               mts.physics("const", *d["beta"]), d["method"], d["dt"], d["T"],
                               mib=d["mib"], tag=i,
                               precision=args.precision))
        if i % 200 == 199:
            print(f"submitted {i + 1}", file=sys.stderr, flush=True)
    results = srv.drain()
    wall = time.perf_counter() - t0
    srv.close()
    ok = [r for r in results.values() if r["status"] == 0]
    # runs the scheme does not keep bounded at dt = h (TS at high contrast):
    # their L2 is not finite (FP32 overflows where FP64 reaches 1e30+; the
    # L-inf of the reference's Error skips NaN entries) or their error is huge
    import math
    diverged = [r for r in ok if not math.isfinite(r["l2"]) or r["linf"] > 1e6]
    bounded = [r for r in ok if r not in diverged]
    status = {}
    for r in results.values():
        status[str(r["status"])] = status.get(str(r["status"]), 0) + 1
    lat = np.array([r["latency_ms"] for r in ok])
    queue = np.array([r["queue_ms"] for r in ok])
    work = sum((draws[r["tag"]]["n"] + 1) ** 3 * r["steps"] for r in ok)
    worst = max(bounded, key=lambda r: r["linf"]) if bounded else None
    line = {
        "metric": METRIC,
        "value": len(ok) / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.requests} requests from seed {args.seed}, Poisson arrivals at "
                        f"{args.rate}/s, N uniform in [24,{args.n_max}], T in "
                        "{0.1,0.25,0.5,1.0}, ADI/LOD-IE/LOD-CN/TS, MIB-L1/L2, "
                        "sphere/ellipsoid/torus, beta-/beta+ log-uniform in [1e-3,1e3], dt=h",
            "hbm_cap_gb": args.hbm_cap_gb,
            "max_inflight": args.max_inflight,
            "precision": args.precision,
        },
        "status_counts": status,
        "p50_latency_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "p50_queue_ms": float(np.percentile(queue, 50)) if len(queue) else None,
        "p99_queue_ms": float(np.percentile(queue, 99)) if len(queue) else None,
        "grid_pt_steps_per_s": work / wall,
        "wall_s": wall,
        "max_linf": worst["linf"] if worst else None,
        "max_linf_solve": describe(draws[worst["tag"]]) if worst else None,
        "median_linf": float(np.median([r["linf"] for r in bounded])) if bounded else None,
        "diverged_runs": len(diverged),
    }
    print(json.dumps(line), flush=True)
    return 0


def copy_calibration():
    """dst = src for 2^26 doubles (512 MiB read, 512 MiB written), three times:
    the HBM counters of this kernel per dispatch over its known byte count give
    the scale of FETCH_SIZE and WRITE_SIZE on this device."""
    import torch

    count = 1 << 26
    src = torch.arange(count, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"copy_calibration": {"bytes_read": count * 8, "bytes_written": count * 8,
                                           "dispatches": 3, "ok": ok}}), flush=True)
    return 0 if ok else 1


REF_BIN = os.path.join(ROOT, "oracle", "_ref", "mts_ref")


def reference_run(cf, args, n_intervals, dt, steps):
    """The reference solver compiled from its own sources (oracle/_ref/mts_ref)
    for `steps` steps of the workload at n_intervals; (step-loop seconds, steps)
    or None when it is not built or stops on its own."""
    import subprocess
    import tempfile
    if not os.path.exists(REF_BIN):
        return None
    kind, *p = cf["surf"]
    bk, b0, b1 = cf["beta"]
    with tempfile.TemporaryDirectory() as out:
        cmd = [REF_BIN, "--method", args.method, "--mib", str(args.mib), "--n", str(n_intervals),
               "--lo", "-1", "--hi", "1", "--dt", repr(dt), "--T", repr(cf["T"]),
               "--surface", f"{kind}:" + ",".join(repr(float(v)) for v in p),
               "--beta", f"{bk}:{float(b0)!r},{float(b1)!r}", "--out", out,
               "--max-steps", str(steps), "--no-dump"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        path = os.path.join(out, "summary.txt")
        if r.returncode != 0 or not os.path.exists(path):
            return None
        kv = dict(l.split(None, 1) for l in open(path).read().splitlines() if l.strip())
    return float(kv["step_loop_s"]), int(kv["steps"])


def port_run(mts, st, method, dt, phys, budget_s, max_steps=4096):
    """The checker restatement (plain C, one thread): whole steps of the run
    until budget_s has passed; (seconds, steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mts_oracle

    op = mts_oracle.OracleProblem.from_setup(st, phys.beta_kind, (phys.bp[0], phys.bp[1]),
                                             -1.0, 1.0)
    u = op.init(0.0)
    t = 0.0
    steps = 0
    spent = 0.0
    while steps < 1 or (spent < budget_s and steps < max_steps):
        t0 = time.perf_counter()
        u, _ = op.run(method, dt, t, 1, u)
        spent += time.perf_counter() - t0
        t += dt
        steps += 1
    return spent, steps


def cpu_baseline(mts, st, args, cf, dt, nodes, phys):
    """The reference's CPU path on the box's host, one core (it is
    single-threaded): the reference solver itself, built from its own sources,
    on a bounded sample of the same run (its first steps) where the reference
    completes this workload; where it stops on its own (the exit(0) of its
    tangential-jump set-up), the plain-C restatement on the same run, with the
    reference/restatement speed ratio measured on the half-size grid of the
    workload beside it."""
    side = round(nodes ** (1 / 3))
    if st.reference_stops()["first_kind"] < 0:
        steps = 2 if nodes > 4e6 else max(2, min(64, int(3e7 / nodes)))
        r = reference_run(cf, args, args.n, dt, steps)
        if r is not None:
            spent, k = r
            return {
                "value": nodes * k / spent,
                "unit": "grid-pt-steps/s",
                "cores": 1,
                "kind": "reference",
                "sample": f"the reference solver (built from its sources), the first {k} step(s) "
                          f"of the same {side}^3 run, {spent:.2f} s on one host core "
                          f"(set-up excluded)",
            }
    spent, k = port_run(mts, st, args.method, dt, phys, 10.0)
    out = {
        "value": nodes * k / spent,
        "unit": "grid-pt-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{k} step(s) of the same {side}^3 run, {spent:.2f} s on one host core "
                  f"(the reference solver stops on its own at this size: "
                  f"{st.reference_stops()['message']})",
    }
    # reference / restatement speed on the half-size grid of the same workload
    nh = args.n // 2
    dth = 2.0 / nh
    sh = mts.Setup(mts.grid(nh + 1), mts.surface(cf["surf"][0], *cf["surf"][1:]), phys,
                   mib=args.mib)
    if sh.reference_stops()["first_kind"] < 0:
        r = reference_run(cf, args, nh, dth, 2)
        if r is not None:
            ps, pk = port_run(mts, sh, args.method, dth, phys, 1e9, max_steps=2)
            ratio = (r[1] / r[0]) / (pk / ps)
            out["reference_over_port"] = ratio
            out["calibration"] = (f"{nh + 1}^3 grid of the same workload: reference {r[1]} steps "
                                  f"in {r[0]:.2f} s, restatement {pk} steps in {ps:.2f} s; the "
                                  f"reference-equivalent rate here is value x reference_over_port")
    return out


if __name__ == "__main__":
    sys.exit(main())
