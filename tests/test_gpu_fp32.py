"""The FP32 form (MTS_FP32): FP32 field storage, LDS tiles and line solve,
rows formed in FP64 and rounded once, FP64 setup and norms.

Bars, stated here because the form is not the reference's arithmetic:
* per step, the FP32 field agrees with the FP64 field to FP32 rounding (6e-8)
  amplified by the step's conditioning, 1 + 4 theta dt beta / h^2: 2e-5 of
  its magnitude on the 41^3-65^3 cases (conditioning up to 640), and the
  scaled bound on the long-line grids;
* the L-infinity error against the exact solution stays within 1.05x of the
  FP64 form's and of the oracle's (the accuracy criterion the FP32 form is
  measured at).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [
    # surface, args, beta, intervals, method, mib, dt, steps
    ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 64, "lod-cn", 2, 1 / 32, 8),
    ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 64, "lod-ie", 1, 1 / 32, 8),
    ("torus", (0.5, 0.2), ("const", 0.01, 1), 48, "ts", 2, 1 / 48, 6),
    ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), ("const", 10, 1), 40, "lod-cn", 2, 1 / 40, 6),
    # Douglas-ADI (its stage-1 right-hand side in the FP64 arithmetic of the
    # FP32 field, rounded once); the sphere has no corner pairs at this size
    ("ellipsoid", (1, 1, 1, 0.5 * (1 - 2 ** -20)), ("const", 2, 1), 64, "adi", 1, 1 / 32, 8),
]


def _pair(mts, ctx, case):
    surf, sargs, beta, n, method, mib, dt, steps = case
    st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), mts.physics(*beta), mib=mib)
    return (mts.Instance(ctx, st, method, dt, precision="fp64"),
            mts.Instance(ctx, st, method, dt, precision="fp32"), st)


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_fp32_follows_fp64_step_by_step(mts, ctx, ci):
    case = CASES[ci]
    print("fp32 case", case[0], case[3], case[4], "mib", case[5], flush=True)
    i64, i32, _ = _pair(mts, ctx, case)
    assert i32.precision == mts.FP32
    i64.initialize(0.0)
    i32.initialize(0.0)
    u0 = i64.download()
    assert rel_err(i32.download(), u0) <= 1e-7  # the initial field, rounded to FP32
    t = 0.0
    dt, steps = case[6], case[7]
    for s in range(steps):
        # restart the FP32 form from the FP64 field so each step is measured alone
        i32.upload(i64.download())
        t += dt
        i64.step(t)
        i32.step(t)
        e = rel_err(i32.download(), i64.download())
        print(f"  step {s + 1}: FP32 vs FP64 {e:.2e}", flush=True)
        assert e <= 2e-5, f"step {s + 1}"
    i64.close()
    i32.close()


@pytest.mark.parametrize("ci", [0, 1, 3, 4])
def test_fp32_error_within_5_percent_of_fp64(mts, ctx, ci):
    case = CASES[ci]
    i64, i32, _ = _pair(mts, ctx, case)
    res = []
    for inst in (i64, i32):
        inst.initialize(0.0)
        steps, t_end = inst.run(0.0, case[6] * case[7])
        res.append(inst.error(t_end))
        inst.close()
    (l64, _, _), (l32, _, _) = res
    print(f"  linf FP64 {l64:.6e} FP32 {l32:.6e}", flush=True)
    assert l32 <= 1.05 * l64


def _oracle_linf(oracle, st, beta_kind, bp, method, dt, T):
    op = oracle.OracleProblem.from_setup(st, beta_kind, bp, -1.0, 1.0)
    u, t_end = op.run(method, dt, 0.0, int(T / dt), op.init(0.0))
    return op.norms(u, t_end)[0], t_end


def test_fp32_error_within_5_percent_of_oracle(mts, oracle, ctx):
    """The FP32 form's L-infinity error against the exact solution within
    1.05x of the oracle's (FP64, the reference's algorithm) on the bench
    workload at 65^3 (LOD-CN MIB-L2, ellipsoid (0.7, 0.5, 0.4), beta- =
    10(1 + 0.5 sin x cos y), beta+ = 1, dt = h, T = 0.5: 16 steps) and on the
    smallest LOD-CN solve of the seeded uneven batch."""
    import importlib.util
    import os

    from conftest import ROOT

    spec = importlib.util.spec_from_file_location("mts_bench", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    rng = np.random.default_rng(0)
    draws = [b.draw_solve(rng, int(rng.choice(np.arange(32, 193, 16))),
                          ["adi", "lod-cn"][int(rng.integers(2))], 1, 0.5) for _ in range(96)]
    d = min((d for d in draws if d["method"] == "lod-cn"), key=lambda d: d["n"])
    cases = [
        (64, ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20)), ("sincos", 10, 1), 1, (10, 1), 2,
         1 / 32, 0.5),
        (d["n"], d["surface"], ("const",) + tuple(d["beta"]), 0, d["beta"], d["mib"], d["dt"],
         d["T"]),
    ]
    for n, surf, beta, kind, bp, mib, dt, T in cases:
        print(f"fp32 vs oracle: N={n} {surf[0]} beta {beta} MIB-L{mib}", flush=True)
        st = mts.Setup(mts.grid(n + 1), mts.surface(surf[0], *surf[1]), mts.physics(*beta),
                       mib=mib)
        lo, t_o = _oracle_linf(oracle, st, kind, bp, "lod-cn", dt, T)
        inst = mts.Instance(ctx, st, "lod-cn", dt, precision="fp32")
        inst.initialize(0.0)
        steps, t_end = inst.run(0.0, T)
        l32 = inst.error(t_end)[0]
        inst.close()
        assert t_end == t_o
        print(f"  linf oracle {lo:.6e} FP32 {l32:.6e} ({l32 / lo:.5f}x)", flush=True)
        assert l32 <= 1.05 * lo


def test_fp32_footprint_and_serving(mts, ctx):
    """The FP32 field takes half the bytes; a served FP32 solve returns the
    same norms as the same solve run alone."""
    case = CASES[0]
    i64, i32, _ = _pair(mts, ctx, case)
    b64, b32 = i64.device_bytes(), i32.device_bytes()
    # x rows padded to 128 bytes: pitch 80 doubles or 96 floats for 65 nodes;
    # field plus the node classification byte; the regular lines' solve
    # tables in the field's type (65-node lines: 16-lane groups of 4-row
    # segments, guarded and unguarded tables, the outside side only (beta
    # varies inside), three directions: 3 * 2 * (2 * 4 * 6 + 4 + 16 * 21)
    # elements: two 4-row segments of 6 factors, 2 + 2 padding, 16 lanes of
    # 21 scalars; whole 16-byte units either way); the rest the same
    tables = 3 * 2 * (2 * 4 * 6 + 4 + 16 * 21)
    assert b64 - 80 * 65 * 65 * 9 - 8 * tables == b32 - 96 * 65 * 65 * 5 - 4 * tables
    i32.initialize(0.0)
    steps, t_end = i32.run(0.0, case[6] * case[7])
    alone = i32.error(t_end)
    i64.close()
    i32.close()
    s = mts.Server(ctx, 8 << 30, max_inflight=4, setup_threads=2)
    q = mts.request(mts.grid(case[3] + 1), mts.surface(case[0], *case[1]),
                    mts.physics(*case[2]), case[4], case[6], case[6] * case[7], mib=case[5],
                    tag=7, precision="fp32")
    s.submit(q)
    r = list(s.drain().values())[0]
    s.close()
    assert r["status"] == 0, r["message"]
    assert (r["linf"], r["l2"], tuple(r["argmax"])) == alone
    assert r["footprint"] == b32


WIDE = [(600, 20, 18), (18, 300, 22), (20, 18, 1024), (130, 66, 34)]


@pytest.mark.parametrize("method", ["lod-cn", "ts"])
@pytest.mark.parametrize("shape", WIDE)
def test_fp32_wide_lines(mts, ctx, shape, method):
    """Long lines: the FP32 solve over 32- and 64-lane groups (shuffle-based
    reduction) against the FP64 form, step by step."""
    print(f"fp32 wide {shape} {method}", flush=True)
    phys = mts.physics("sincos", 10, 1)
    st = mts.Setup(mts.grid(shape), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20),
                   phys, mib=2)
    dt = 1 / 64
    # FP32 rounding (6e-8) amplified by the step's conditioning: the implicit
    # and explicit line operators have norms up to 1 + 4 theta dt beta / h^2
    # (2000 to 123000 on these grids, against 640 on the 65^3 cases above)
    h = min(2.0 / (n - 1) for n in shape)
    kappa = 1 + 4 * 0.5 * dt * 10 * 1.5 / h ** 2
    bound = 6e-8 * kappa
    i64 = mts.Instance(ctx, st, method, dt, precision="fp64")
    i32 = mts.Instance(ctx, st, method, dt, precision="fp32")
    i64.initialize(0.0)
    t = 0.0
    for s in range(2):
        i32.upload(i64.download())
        t += dt
        i64.step(t)
        i32.step(t)
        e = rel_err(i32.download(), i64.download())
        print(f"  step {s + 1}: FP32 vs FP64 {e:.2e} (bound {bound:.1e})", flush=True)
        assert e <= bound, f"step {s + 1}"
    i64.close()
    i32.close()
