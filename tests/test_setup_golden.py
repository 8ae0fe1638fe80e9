"""Host FP64 setup (Mesh + Intersections + MIB weights) against the reference.

The golden crossings were written by the reference solver built from its own
sources (tests/golden/make_golden.py).  The product's setup evaluates every
quantity in the reference's operation order, so integer data agree exactly and
floating-point data to the last bit, except where the reference's compiler
fused sin/cos of one argument into glibc's sincos (one ulp on one bumpy-sphere
normal); the test allows 4 ulp of the field's largest value.
"""
import numpy as np
import pytest

from conftest import GOLDEN_NAMES, case_setup, golden_crossings, load_golden


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_crossings_match_reference_bitwise(mts, name):
    g = load_golden(name)
    st, _ = case_setup(mts, g["meta"])
    for d in range(3):
        mine = st.crossings(d)
        ref = golden_crossings(g, d)
        assert len(mine) == len(ref), f"direction {d}: {len(mine)} vs {len(ref)} crossings"
        for f in ("j", "k", "id", "left_loc", "nw"):
            np.testing.assert_array_equal(mine[f], ref[f], err_msg=f"dir {d} field {f}")
        for f in ("gamma", "coord", "normal", "beta_in", "beta_out", "p", "wl", "wr"):
            a, b = mine[f], ref[f]
            if a.size == 0:
                continue
            tol = 4 * np.spacing(np.max(np.abs(b)))
            diff = np.max(np.abs(a - b))
            assert diff <= tol, f"dir {d} field {f}: max diff {diff} > {tol}"


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_indicator_flips_at_crossings(mts, name):
    """Along every line the classification changes sign exactly at the crossings."""
    g = load_golden(name)
    st, _ = case_setup(mts, g["meta"])
    ind = st.indicator()  # (nz, ny, nx)
    nz, ny, nx = ind.shape
    for d in range(3):
        cr = st.crossings(d)
        lines = {}
        for c in cr:
            lines.setdefault((int(c["j"]), int(c["k"])), []).append(int(c["left_loc"]))
        if d == 0:
            arr = np.transpose(ind, (1, 0, 2))  # [iy][iz][ix] -> j = iy, k = iz
            get = lambda j, k: arr[j, k, :]
            nj, nk = ny, nz
        elif d == 1:
            get = lambda j, k: ind[k, :, j]  # j = ix, k = iz
            nj, nk = nx, nz
        else:
            get = lambda j, k: ind[:, k, j]  # j = ix, k = iy
            nj, nk = nx, ny
        for k in range(1, nk - 1):
            for j in range(1, nj - 1):
                line = get(j, k).astype(int)
                flips = list(np.nonzero(line[:-1] * line[1:] < 0)[0])
                assert flips == sorted(lines.get((j, k), [])), (d, j, k)


def test_corner_pairs_present(mts):
    """The golden geometries include corner pairs (exercised by parity)."""
    total = 0
    for name in GOLDEN_NAMES:
        g = load_golden(name)
        for d in range(3):
            total += int(np.sum(g[f"cr{d}_int"][:, 2] < 0))
    assert total > 0
