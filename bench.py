#!/usr/bin/env python3
"""Throughput of the Matched-TS per-time-step path on one MI355X.

Workload: Crank-Nicolson LOD with MIB-L2 on [-1,1]^3 with 257^3 nodes,
interface the ellipsoid x^2/0.49 + y^2/0.25 + z^2/0.16 = r^2 (r = 1 - 2^-20,
which moves the surface off the grid node (0, 0.5, 0)), beta- = 10 (1 + 0.5
sin x cos y), beta+ = 1, manufactured solution u- = e^-t sin x sin y sin z,
u+ = e^-t cos x cos y cos z, dt = h = 1/128, T = 0.5 (64 steps), FP64.

One step = one Solve_2nd: three line stages (x, y, z) and the boundary
faces.  The field and the MIB table are resident in HBM before timing starts.
The timed region is exactly K steps between two stream synchronizations; the
per-stage HIP events that feed the roofline figure run inside it.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
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
    ap.add_argument("--n", type=int, default=256, help="intervals per axis")
    ap.add_argument("--method", default="lod-cn", choices=["adi", "lod-ie", "lod-cn", "ts"])
    ap.add_argument("--mib", type=int, default=2, choices=[1, 2])
    ap.add_argument("--cpu-baseline", type=int, default=1,
                    help="time the checker restatement on the host (0 to skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs the per-step path on a single GPU; "
              "multi-GPU runs are not part of it (use --gpus 1)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        return copy_calibration()

    mts = load_mts()
    n = args.n + 1
    dt = 2.0 / args.n
    T = 0.5
    phys = mts.physics("sincos", 10.0, 1.0)
    surf = mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20)
    t0 = time.perf_counter()
    st = mts.Setup(mts.grid(n), surf, phys, mib=args.mib)
    setup_s = time.perf_counter() - t0
    info = st.info()
    ctx = mts.Context(0)
    dev_name, _ = ctx.device_info()
    inst = mts.Instance(ctx, st, args.method, dt)

    # warmup: untimed steps from the initial condition
    inst.initialize(0.0)
    t = 0.0
    for _ in range(args.warmup):
        t += dt
        inst.step(t)
    inst.sync()

    # timed run: K steps of the reference time loop (t += dt) from t0 = 0
    inst.initialize(0.0)
    inst.profile(True)
    inst.profile_reset()
    inst.sync()
    t = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t += dt
        inst.step(t)
    inst.sync()
    elapsed = time.perf_counter() - t_start
    stages = inst.profile_read()
    inst.profile(False)
    linf, l2, _ = inst.error(t)

    nodes = n ** 3
    value = nodes * args.steps / elapsed
    # dominant stage: largest total time; algorithmic bytes = one FP64 read and
    # one FP64 write of every node (each line stage reads its input field once
    # and writes its output once; the Douglas y/z stages also read u^n)
    per_node = {"adi_y": 24, "adi_z": 24}
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

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(mts, st, args.method, dt, nodes, phys)

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
            "workload": f"{args.method.upper()} MIB-L{args.mib} {n}^3 ellipsoid(0.7,0.5,0.4)"
                        f" beta-=10(1+sin x cos y/2) beta+=1 dt=h={dt} T={T}",
            "grid_points": nodes,
            "crossings": [int(info.n_cross[d]) for d in range(3)],
            "t_end": t,
            "linf_error": linf,
            "l2_error": l2,
            "setup_s": setup_s,
            "device": dev_name,
            "stage_ms_per_launch": {k: v[1] / max(v[0], 1) for k, v in stages.items()},
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


def cpu_baseline(mts, st, method, dt, nodes, phys):
    """The checker restatement (plain C, one thread) on a bounded sample of the
    same workload: whole steps of the same 257^3 run until ~10 s have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mts_oracle

    op = mts_oracle.OracleProblem.from_setup(st, phys.beta_kind, (phys.bp[0], phys.bp[1]),
                                             -1.0, 1.0)
    u = op.init(0.0)
    t = 0.0
    steps = 0
    spent = 0.0
    while steps < 1 or (spent < 10.0 and steps < 8):
        t0 = time.perf_counter()
        u, _ = op.run(method, dt, t, 1, u)
        spent += time.perf_counter() - t0
        t += dt
        steps += 1
    return {
        "value": nodes * steps / spent,
        "unit": "grid-pt-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} step(s) of the same {round(nodes ** (1 / 3))}^3 run, "
                  f"{spent:.1f} s on one host core",
    }


if __name__ == "__main__":
    sys.exit(main())
