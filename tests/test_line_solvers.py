"""The line solvers of the per-step path, on the CPU.

The GPU solves each line with a partitioned algorithm (segments per lane,
fold, cyclic reduction) unless the line has a row that is not diagonally
dominant, where it runs the Thomas algorithm in the reference's order.  The
oracle holds an operation-for-operation CPU model of that choice ("partitioned")
next to the reference's Thomas algorithm and a partial-pivoting elimination
(LAPACK dgtsv's algorithm, "pivot").  Pins:

* all three solve random diagonally dominant systems to rounding (dense solve);
* the model takes the Thomas path, bit for bit, on a line with a
  non-dominant row;
* on whole time steps of the parity geometries the model stays within the
  1e-12 bar of the Thomas result, and within the spread the conditioning of
  the lines allows (a pivoting solver differs from Thomas by as much).
"""
import numpy as np
import pytest

from conftest import rel_err

KINDS = ("thomas", "pivot", "partitioned")


# 54, 66, 65, 130, 129, 255, 257, 513, 1024: full splits (every segment but the
# last holds the register bound of rows), whose last segment the model pads
# with identity rows as the kernel's unguarded solve does; 54 also leaves
# lanes past the last segment idle, 255 has a last segment 3 rows short
@pytest.mark.parametrize("n", [4, 5, 17, 41, 54, 64, 65, 66, 128, 129, 130, 255, 257, 513,
                               1024])
def test_solvers_match_dense_on_dominant_systems(oracle, n):
    rng = np.random.default_rng(100 + n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 2.2 + np.abs(a) + np.abs(c)
    b *= rng.choice([-1.0, 1.0], n)
    r = rng.uniform(-1, 1, n)
    # line systems: rows 0 and n-1 are identity rows holding boundary values
    a[0] = c[0] = a[-1] = c[-1] = 0.0
    b[0] = b[-1] = 1.0
    x = np.linalg.solve(np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1), r)
    for k in KINDS:
        assert rel_err(oracle.line_solve(k, a, b, c, r), x) < 5e-15, k


def test_partitioned_model_takes_thomas_on_non_dominant_line(oracle):
    n = 41
    rng = np.random.default_rng(7)
    a = -rng.uniform(1, 2, n)
    c = -rng.uniform(1, 2, n)
    b = 1.0 + np.abs(a) + np.abs(c)
    a[0] = c[0] = a[-1] = c[-1] = 0.0
    b[0] = b[-1] = 1.0
    a[21], b[21], c[21] = 5.6e4, 102.0, -5.6e4  # a Convert2Tri-like row
    r = rng.uniform(-1, 1, n)
    part = oracle.line_solve("partitioned", a, b, c, r)
    np.testing.assert_array_equal(part, oracle.line_solve("thomas", a, b, c, r))
    # and without that row the partitioned path is taken (differs in rounding only)
    a[21], b[21], c[21] = -1.0, 3.0, -1.0
    p2 = oracle.line_solve("partitioned", a, b, c, r)
    t2 = oracle.line_solve("thomas", a, b, c, r)
    assert rel_err(p2, t2) < 1e-14


def test_pivot_solver_interchanges_rows(oracle):
    # zero leading pivot: Thomas fails, the pivoting solver does not
    a = np.array([0.0, 1.0, 1.0, 2.0])
    b = np.array([0.0, 1.0, 3.0, 1.0])
    c = np.array([2.0, 1.0, 1.0, 0.0])
    r = np.array([1.0, 2.0, 3.0, 4.0])
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    x = oracle.line_solve("pivot", a, b, c, r)
    assert rel_err(x, np.linalg.solve(A, r)) < 1e-15
    with pytest.raises(ZeroDivisionError):
        oracle.line_solve("thomas", a, b, c, r)


STEP_CASES = [
    # surface, args, beta, n intervals, dt
    ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 32, 1 / 16),
    ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), ("const", 1000, 1), 40, 1 / 20),
    ("torus", (0.5, 0.2), ("const", 0.01, 1), 30, 1 / 15),
]


@pytest.mark.parametrize("mib", [1, 2])
@pytest.mark.parametrize("ci", range(len(STEP_CASES)))
def test_partitioned_model_step_within_bar(mts, oracle, ci, mib):
    surf, sargs, beta, n, dt = STEP_CASES[ci]
    phys = mts.physics(*beta)
    st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), phys, mib=mib)
    op = oracle.OracleProblem.from_setup(st, phys.beta_kind, beta[1:], -1.0, 1.0)
    u0 = op.init(0.0)
    for method in ("adi", "lod-ie", "lod-cn"):
        res = {}
        for k in KINDS:
            with oracle.line_solver(k):
                res[k], _ = op.run(method, dt, 0.0, 1, u0.copy())
        scale = max(1.0, float(np.max(np.abs(res["thomas"]))))
        d_part = rel_err(res["partitioned"], res["thomas"])
        d_pivot = rel_err(res["pivot"], res["thomas"])
        assert d_part <= 1e-12 * scale, (method, d_part)
        # no worse than the conditioning spread, with a margin for small spreads
        assert d_part <= max(10 * d_pivot, 1e-13), (method, d_part, d_pivot)
