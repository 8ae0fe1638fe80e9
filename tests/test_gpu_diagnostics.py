"""The interface diagnostics on the device (mts_instance_diagnostics,
mts_instance_crossing_errors) against the host diagnostics, which the golden
fixtures pin to the reference's Refresh_Fp / Refresh_Jump values
(tests/test_diagnostics.py): per crossing to 1e-12 of max(1, |value|), NaN
and infinity where the host has them (the device takes sin and cos of a
crossing from its jump record, which a compiler may form with one sincos
call: an ulp apart from sin alone), the maxima likewise, the L2 norms to
1e-12 relative (also another summation order of the same squares).  Geometries with irregular crossings and corner
pairs (torus, bumpy), constant and variable beta, MIB-L1 and MIB-L2, both
field precisions (the diagnostics do not read the field)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 64, 2, "lod-cn", "fp64"),
    ("torus", (0.5, 0.2), ("const", 0.01, 1), 48, 2, "ts", "fp64"),
    ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), ("const", 1000, 1), 26, 1, "lod-ie", "fp32"),
    ("ellipsoid", (1, 1, 1, 0.5 * (1 - 2 ** -20)), ("const", 10, 1), 96, 1, "adi", "fp64"),
]


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_device_diagnostics_match_host(mts, ctx, ci):
    surf, sargs, beta, n, mib, method, prec = CASES[ci]
    print(f"diagnostics {surf} N={n} MIB-L{mib} {beta}", flush=True)
    st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), mts.physics(*beta), mib=mib)
    inst = mts.Instance(ctx, st, method, 2.0 / n, precision=prec)
    info = st.info()
    for t in (0.0, 0.37):
        for d in range(3):
            hl, hr, hj = st.crossing_errors(d, t)
            dl, dr, dj = inst.crossing_errors(d, t)
            assert info.n_cross[d] == len(dl)
            for h, g in ((hl, dl), (hr, dr), (hj, dj)):
                assert _close(h, g), (d, t)
        (hf, hf2), (hjm, hj2) = st.diagnostics(t)
        (gf, gf2), (gjm, gj2) = inst.diagnostics(t)
        print(f"  t={t}: fp {gf:.6e} {gf2:.6e} jump {gjm:.6e} {gj2:.6e} "
              f"(corner pairs {tuple(info.corner_pairs)})", flush=True)
        assert _close([hf, hjm], [gf, gjm])
        assert _close_rel(hf2, gf2) and _close_rel(hj2, gj2)
    inst.close()


def _close(h, g):
    h, g = np.asarray(h, dtype=float), np.asarray(g, dtype=float)
    if not np.array_equal(np.isnan(h), np.isnan(g)):
        return False
    fin = np.isfinite(h)
    if not np.array_equal(fin, np.isfinite(g)) or not np.array_equal(h[~fin & ~np.isnan(h)],
                                                                      g[~fin & ~np.isnan(g)]):
        return False
    return bool(np.all(np.abs(h[fin] - g[fin]) <= 1e-12 * np.maximum(1.0, np.abs(h[fin]))))


def _close_rel(h, g):
    """L2 norms: infinite where the host's is (a relative error divided by an
    exact zero), else to 1e-12 relative."""
    if not np.isfinite(h):
        return g == h or (np.isnan(h) and np.isnan(g))
    return abs(g - h) <= 1e-12 * abs(h)
