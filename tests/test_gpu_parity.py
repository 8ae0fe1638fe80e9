"""GPU parity: the HIP path against the reference's golden fields and the oracle.

Tolerance: 1e-12 normwise relative (max-norm difference over the field's
max-norm), the bar for the plain FP64 form, per step and restarted from the
reference's own field of the previous step.  Where the step's line systems
are ill-conditioned (the bumpy sphere with beta 1000/1) the bound is four
times the measured spread between two backward-stable elimination orders of
the same step (the reference's Thomas and partial pivoting, in the oracle):
the GPU's partitioned line solve is one more such order.  Over a whole run the
bound is the largest per-step bound times the run's measured amplification of
a relative perturbation of its initial field.  Everything else follows the
reference's operation order.
"""
import numpy as np
import pytest

from conftest import (GOLDEN_NAMES, PARITY, beta_kind, case_setup, golden_crossings,
                      golden_oracle, load_golden, rel_err, run_amplification, step_floor,
                      step_times)

pytestmark = pytest.mark.gpu


def has_corners(g):
    return any((g[f"cr{d}_int"][:, 2] < 0).any() for d in range(3) if len(g[f"cr{d}_int"]))


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_gpu_matches_reference(mts, oracle, ctx, name):
    g = load_golden(name)
    meta = g["meta"]
    print(f"case {name}", flush=True)
    if meta["method"] == "adi" and has_corners(g):
        pytest.skip(DOUGLAS_CORNERS_OFF)
    st, _ = case_setup(mts, meta)
    op = golden_oracle(mts, oracle, g)
    method, dt = meta["method"], meta["dt"]
    us = g["u_steps"]
    ts = step_times(dt, len(us))
    inst = mts.Instance(ctx, st, method, dt)
    inst.initialize(0.0)
    floors = []
    for j in range(len(us)):
        u_in = inst.download() if j == 0 else us[j - 1]
        tol = step_floor(oracle, op, method, dt, ts[j], u_in)
        floors.append(tol)
        inst.upload(u_in)
        inst.step(ts[j + 1])
        e = rel_err(inst.download(), us[j])
        print(f"  step {j + 1}: {e:.2e} (bound {tol:.1e})", flush=True)
        assert e <= tol, f"step {j + 1}"
    inst.initialize(0.0)
    steps, t_end = inst.run(0.0, meta["T"])
    assert steps == int(g["summary"][0])
    assert t_end == g["summary"][1]
    uf = inst.download()
    tol = max(floors) * run_amplification(op, method, dt, steps)
    e = rel_err(uf, g["u_final"])
    print(f"  final: {e:.2e} (bound {tol:.1e})", flush=True)
    assert e <= tol
    linf, l2, am = inst.error(t_end)
    assert abs(linf - g["summary"][2]) <= tol * max(1.0, g["summary"][2])
    assert abs(l2 - g["summary"][3]) <= tol * max(1.0, g["summary"][3])
    inst.close()


DOUGLAS_CORNERS_OFF = (
    "Douglas-ADI with corner pairs is switched off on the GPU: two runs (ellipsoid 65^3 "
    "MIB-L1, torus 49^3 MIB-L2) raised a memory aperture violation (dispatch reported as "
    "spec_kernel) whose cause is not found; the host bounds check of those cases is clean")


def test_gpu_douglas_with_corners_is_refused(mts, ctx):
    """The switched-off combination fails loudly at instance creation (no kernel runs)."""
    phys = mts.physics("sincos", 10, 1)
    st = mts.Setup(mts.grid(65), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20), phys,
                   mib=1)
    assert sum(st.info().corner_pairs) > 0
    with pytest.raises(mts.MtsError) as e:
        mts.Instance(ctx, st, "adi", 1 / 32)
    assert e.value.code == -6
    inst = mts.Instance(ctx, st, "lod-ie", 1 / 32)  # other methods unaffected
    inst.close()


CASES = [
    # (surface, args, beta, n intervals, dt, steps)
    ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 64, 1 / 32, 3),
    ("torus", (0.5, 0.2), ("const", 0.01, 1), 48, 1 / 24, 3),
    ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), ("const", 1000, 1), 40, 1 / 20, 2),
    ("ellipsoid", (0.8, 0.55, 0.45, 0.95), ("const", 0.001, 1), 36, 1 / 64, 3),
]


@pytest.mark.parametrize("method", ["adi", "lod-ie", "lod-cn", "ts"])
@pytest.mark.parametrize("mib", [1, 2])
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_gpu_matches_oracle_per_step(mts, oracle, ctx, method, mib, ci):
    """Beyond the sizes the reference can set up: every step against the oracle
    restarted from the GPU's previous result (isolates per-step error)."""
    surf, sargs, beta, n, dt, steps = CASES[ci]
    print(f"case {surf} n={n} {method} mib{mib}", flush=True)
    phys = mts.physics(*beta)
    st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), phys, mib=mib)
    if method == "adi" and sum(st.info().corner_pairs) > 0:
        pytest.skip(DOUGLAS_CORNERS_OFF)
    op = oracle.OracleProblem.from_setup(st, phys.beta_kind, beta[1:], -1.0, 1.0)
    inst = mts.Instance(ctx, st, method, dt)
    inst.initialize(0.0)
    u = inst.download()
    np.testing.assert_array_equal(u, op.init(0.0))
    t = 0.0
    for s in range(steps):
        uo, _ = op.run(method, dt, t, 1, u)
        with oracle.line_solver("partitioned"):
            um, _ = op.run(method, dt, t, 1, u)
        tol = step_floor(oracle, op, method, dt, t, u)
        t += dt
        inst.step(t)
        ug = inst.download()
        e_ref, e_model = rel_err(ug, uo), rel_err(ug, um)
        print(f"  step {s + 1}: vs oracle (Thomas) {e_ref:.2e}, vs CPU model of the "
              f"GPU line solve {e_model:.2e} (bound {tol:.1e})", flush=True)
        assert e_ref <= tol, f"step {s + 1}"
        # the same elimination order as the kernel, so what is left is the
        # rounding of the right-hand side (the kernel applies each MIB block as
        # one precombined linear form)
        assert e_model <= tol, f"step {s + 1} (model)"
        u = ug


# x / y / z lines over 64-, 32- and 16-lane groups (1024-node x and y lines:
# the explicit stages' two LDS tiles above 64 KiB); (54, 20, 22) and
# (255, 18, 20): full splits along x with idle lanes past the last segment and
# with a last segment 3 rows short of the register bound
WIDE = [(600, 20, 18), (18, 300, 22), (20, 18, 1024), (130, 66, 34), (54, 20, 22),
        (255, 18, 20), (1024, 18, 20), (18, 1024, 20)]


@pytest.mark.parametrize("method", ["lod-ie", "lod-cn", "ts"])
@pytest.mark.parametrize("shape", WIDE)
def test_gpu_wide_lines_match_oracle(mts, oracle, ctx, shape, method):
    """Anisotropic grids: lines split over 64-, 32- and 16-lane groups."""
    print(f"case wide {shape} {method}", flush=True)
    phys = mts.physics("sincos", 10, 1)
    st = mts.Setup(mts.grid(shape), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20),
                   phys, mib=2)
    op = oracle.OracleProblem.from_setup(st, 1, (10, 1), -1.0, 1.0)
    dt = 1 / 64
    inst = mts.Instance(ctx, st, method, dt)
    inst.initialize(0.0)
    u = inst.download()
    t = 0.0
    for s in range(2):
        uo, _ = op.run(method, dt, t, 1, u)
        tol = step_floor(oracle, op, method, dt, t, u)
        t += dt
        inst.step(t)
        ug = inst.download()
        assert rel_err(ug, uo) <= tol, f"step {s + 1}"
        u = ug
    inst.close()


@pytest.mark.parametrize("n", [2, 3, 4, 5, 17, 64, 65, 128, 129, 257, 513, 1024])
def test_gpu_batched_tdma(oracle, ctx, n):
    rng = np.random.default_rng(n)
    nl = 37
    a = rng.uniform(-1, 1, (nl, n))
    c = rng.uniform(-1, 1, (nl, n))
    b = 2.2 + np.abs(a) + np.abs(c)
    r = rng.uniform(-1, 1, (nl, n))
    u = ctx.tdma_batched(a, b, c, r)
    for l in range(nl):
        ref = oracle.tdma(a[l], b[l], c[l], r[l])
        assert rel_err(u[l], ref) < 1e-13


def test_gpu_error_norms_match_oracle(mts, oracle, ctx):
    phys = mts.physics("sincos", 10, 1)
    st = mts.Setup(mts.grid(41), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20), phys,
                   mib=2)
    op = oracle.OracleProblem.from_setup(st, 1, (10, 1), -1.0, 1.0)
    inst = mts.Instance(ctx, st, "lod-cn", 0.05)
    inst.initialize(0.0)
    inst.run(0.0, 0.2)
    u = inst.download()
    li, l2, am = inst.error(0.2)
    lo, l2o, amo = op.norms(u, 0.2)
    assert abs(li - lo) <= 1e-15 * max(1.0, lo)
    assert abs(l2 - l2o) <= 1e-13 * l2o
    assert am == amo


def test_gpu_layouts_round_trip(mts, ctx):
    phys = mts.physics("const", 10, 1)
    st = mts.Setup(mts.grid((17, 19, 21)), mts.surface("ellipsoid", 1, 1, 1, 0.47), phys, mib=1)
    inst = mts.Instance(ctx, st, "lod-ie", 0.1)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(inst.shape)
    inst.upload(u)
    np.testing.assert_array_equal(inst.download(), u)
    ref = np.ascontiguousarray(np.transpose(u, (2, 1, 0)))  # [ix][iy][iz]
    np.testing.assert_array_equal(inst.download(layout=mts.LAYOUT_REF), ref)
    inst.upload(ref, layout=mts.LAYOUT_REF)
    np.testing.assert_array_equal(inst.download(), u)
