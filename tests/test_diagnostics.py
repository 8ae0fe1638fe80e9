"""End-of-run interface diagnostics against the reference: per crossing, the
fictitious-point errors (Refresh_Fp) and the jump error (Refresh_Jump's
jump.err) at the final time, and the reference's own Error_Fp / Error_Jump
report lines, from the golden fixtures (written by the reference build)."""
import numpy as np
import pytest

from conftest import GOLDEN_NAMES, case_setup, load_golden


def _close(ours, ref, rtol):
    """Elementwise: infinities where the reference divided by an exact zero,
    and finite values to rtol (relative, floored at 1)."""
    ours, ref = np.asarray(ours), np.asarray(ref)
    inf_r = ~np.isfinite(ref)
    assert np.array_equal(~np.isfinite(ours), inf_r)
    d = np.abs(ours[~inf_r] - ref[~inf_r]) / np.maximum(1.0, np.abs(ref[~inf_r]))
    return float(d.max()) if d.size else 0.0


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_crossing_errors_match_reference(mts, name):
    g = load_golden(name)
    meta = g["meta"]
    st, _ = case_setup(mts, meta)
    t = float(g["summary"][1])
    worst = 0.0
    for d in range(3):
        el, er, je = st.crossing_errors(d, t)
        ref = g[f"diag{d}"]
        assert len(el) == len(ref)
        if not len(ref):
            continue
        worst = max(worst, _close(el, ref[:, 0], 0.0), _close(er, ref[:, 1], 0.0),
                    _close(je, ref[:, 2], 0.0))
    # same weights, same exact values, same operation order: bit for bit
    assert worst == 0.0, worst


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_report_lines_match_reference(mts, name):
    g = load_golden(name)
    st, _ = case_setup(mts, g["meta"])
    (fl, f2), (jl, j2) = st.diagnostics(float(g["summary"][1]))
    want = [f"{v:.3e}" for v in g["diag_report"]]
    got = [f"{v:.3e}" for v in (fl, f2, jl, j2)]
    assert got == want


def test_fp_errors_are_small_where_defined(mts):
    """The weights reproduce the exact solution's extension to the order of
    the scheme: away from the zero set of the other side's solution the
    fictitious points' relative errors are small."""
    st = mts.Setup(mts.grid(65), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20),
                   mts.physics("sincos", 10, 1), mib=2)
    el, er, _ = st.crossing_errors(2, 0.25)
    fin = np.isfinite(el) & np.isfinite(er)
    assert fin.mean() > 0.9
    assert np.median(el[fin]) < 1e-2 and np.median(er[fin]) < 1e-2
