"""Checks fixed by the mathematics, independent of the reference build.

* Thomas algorithm (the reference's TDMA) against a dense solve.
* MIB fictitious-point weights reproduce polynomial extensions exactly: for
  piecewise polynomials that satisfy the interface jumps, each fictitious
  value equals the other side's polynomial at that node (degree 1 for MIB-L1's
  2-point interpolation, 2 for MIB-L2 irregular, 3 for MIB-L2 corners).
* With no interface in the box, each scheme is the textbook split scheme;
  the oracle must agree with an independent matrix-form restatement.
"""
import numpy as np
import pytest

from conftest import rel_err


# ------------------------------------------------------------------ TDMA
@pytest.mark.parametrize("n", [2, 3, 5, 17, 64, 257])
def test_oracle_tdma_matches_dense(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 2.5 + np.abs(a) + np.abs(c)
    r = rng.uniform(-1, 1, n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    x = oracle.tdma(a, b, c, r)
    assert rel_err(x, np.linalg.solve(A, r)) < 1e-14
    # a[0] and c[n-1] are not part of the system
    a2, c2 = a.copy(), c.copy()
    a2[0], c2[-1] = 1e30, -1e30
    np.testing.assert_array_equal(oracle.tdma(a2, b, c2, r), x)


def test_oracle_tdma_zero_pivot(oracle):
    with pytest.raises(ZeroDivisionError):
        oracle.tdma(np.zeros(3), np.zeros(3), np.zeros(3), np.ones(3))


# ------------------------------------------------- MIB polynomial exactness
def _cases(mts):
    r0 = 0.5 * (1 - 2 ** -20)
    yield "ellipsoid", (1, 1, 1, r0), ("const", 10, 1), 16
    yield "ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 32
    yield "torus", (0.5, 0.2), ("const", 0.01, 1), 22
    yield "bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), ("const", 1000, 1), 26


def _fp_value(w, grid_vals, jumps):
    return float(np.dot(w[:len(grid_vals)], grid_vals) + np.dot(w[len(grid_vals):], jumps))


@pytest.mark.parametrize("mib", [1, 2])
@pytest.mark.parametrize("which", range(4))
def test_mib_weights_reproduce_polynomials(mts, mib, which):
    surf, sargs, beta, n = list(_cases(mts))[which]
    st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), mts.physics(*beta), mib=mib)
    ind = st.indicator()
    h = 2.0 / n
    rng = np.random.default_rng(7 + which)
    checked = 0
    worst = 0.0
    for d in range(3):
        cr = st.crossings(d)
        i = 0
        while i < len(cr):
            c = cr[i]
            j, k, ll = int(c["j"]), int(c["k"]), int(c["left_loc"])
            q = [0, 0, 0]
            axes = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[d]
            q[axes[0]], q[axes[1]] = j, k

            def side(row):
                qq = list(q)
                qq[d] = row
                return int(ind[qq[2], qq[1], qq[0]])

            if c["id"] > 0:
                deg = 1 if mib == 1 else 2
                pl = np.polynomial.Polynomial(rng.uniform(-1, 1, deg + 1))  # left side (node ll)
                pr = np.polynomial.Polynomial(rng.uniform(-1, 1, deg + 1))  # right side
                g = float(c["gamma"])
                sl = side(ll)
                up, um = (pl, pr) if sl > 0 else (pr, pl)
                ju = up(g) - um(g)
                jb = c["beta_out"] * up.deriv()(g) - c["beta_in"] * um.deriv()(g)
                x = (np.arange(-1, 3)) * h  # nodes ll-1 .. ll+2 relative to ll
                vals = np.array([pl(x[0]), pl(x[1]), pr(x[2]), pr(x[3])])
                fl = _fp_value(c["wl"][:6], vals, [ju, jb])   # right side at ll
                fr = _fp_value(c["wr"][:6], vals, [ju, jb])   # left side at ll+1
                scale = 1.0 + np.max(np.abs(c["wl"][:6])) + np.max(np.abs(c["wr"][:6]))
                e = max(abs(fl - pr(0.0)), abs(fr - pl(h))) / scale
                worst = max(worst, e)
                checked += 1
                i += 1
            else:
                R = cr[i + 1]
                deg = 1 if mib == 1 else 3
                po = np.polynomial.Polynomial(rng.uniform(-1, 1, deg + 1))  # nodes ll, ll+2
                pm = np.polynomial.Polynomial(rng.uniform(-1, 1, deg + 1))  # node ll+1
                gl, gr = float(c["gamma"]), h + float(R["gamma"])
                so = side(ll)
                upl, uml = (po, pm) if so > 0 else (pm, po)
                ju_l = upl(gl) - uml(gl)
                jb_l = c["beta_out"] * upl.deriv()(gl) - c["beta_in"] * uml.deriv()(gl)
                ju_r = upl(gr) - uml(gr)
                jb_r = R["beta_out"] * upl.deriv()(gr) - R["beta_in"] * uml.deriv()(gr)
                x = np.arange(-1, 4) * h  # nodes ll-1 .. ll+3
                vals = np.array([po(x[0]), po(x[1]), pm(x[2]), po(x[3]), po(x[4])])
                jumps = [ju_l, jb_l, ju_r, jb_r]
                f_lwl = _fp_value(c["wl"], vals, jumps)   # middle side at ll
                f_lwr = _fp_value(c["wr"], vals, jumps)   # outer side at ll+1
                f_rwl = _fp_value(R["wl"], vals, jumps)   # outer side at ll+1
                f_rwr = _fp_value(R["wr"], vals, jumps)   # middle side at ll+2
                scale = 1.0 + max(np.max(np.abs(w)) for w in (c["wl"], c["wr"], R["wl"], R["wr"]))
                e = max(abs(f_lwl - pm(0.0)), abs(f_lwr - po(h)), abs(f_rwl - po(h)),
                        abs(f_rwr - pm(2 * h))) / scale
                worst = max(worst, e)
                checked += 1
                i += 2
    assert checked > 0
    assert worst < 1e-11, worst


# ------------------------------------------- split schemes without interface
def _line_op(u, d, coef):
    """coef * beta-free second difference along axis d (array axes: z, y, x)."""
    ax = 2 - d
    out = np.zeros_like(u)
    sl = [slice(1, -1)] * 3
    c = tuple(sl)
    lo = list(sl)
    hi = list(sl)
    lo[ax] = slice(0, -2)
    hi[ax] = slice(2, None)
    out[c] = coef * (u[tuple(lo)] - 2 * u[c] + u[tuple(hi)])
    return out


def _implicit(rhs, d, lam):
    """Solve (I - lam D2) x = rhs on every interior line along d; rows 0 and
    n-1 of each line hold the known boundary values."""
    ax = 2 - d
    x = np.moveaxis(rhs.copy(), ax, -1)
    n = x.shape[-1]
    m = n - 2
    A = np.diag(np.full(m, 1 + 2 * lam)) + np.diag(np.full(m - 1, -lam), 1) + \
        np.diag(np.full(m - 1, -lam), -1)
    for idx in np.ndindex(x.shape[:-1]):
        if 0 in idx or any(idx[t] == x.shape[t] - 1 for t in range(2)):
            continue
        line = x[idx]
        r = line[1:-1].copy()
        r[0] += lam * line[0]
        r[-1] += lam * line[-1]
        line[1:-1] = np.linalg.solve(A, r)
    return np.moveaxis(x, -1, ax)


def _matrix_scheme(method, u, t1, dt, bout, X, Y, Z, h):
    def exact(t):
        return np.exp(-t) * np.cos(X) * np.cos(Y) * np.cos(Z)

    def f(t):
        return (3 * bout - 1) * exact(t)

    def bc(v, t):
        e = exact(t)
        v = v.copy()
        for ax in range(3):
            idx0 = [slice(None)] * 3
            idx0[ax] = 0
            v[tuple(idx0)] = e[tuple(idx0)]
            idx0[ax] = -1
            v[tuple(idx0)] = e[tuple(idx0)]
        return v

    L = lambda v, d, c: _line_op(v, d, c * bout / h ** 2)
    lam = lambda c: c * dt * bout / h ** 2
    inner = (slice(1, -1),) * 3
    if method == "adi":
        r = u.copy()
        r[inner] = (L(u, 1, dt) + L(u, 2, dt) + u + dt * f(t1))[inner]
        us = _implicit(bc(r, t1), 0, lam(1))
        r = us.copy()
        r[inner] = (-L(u, 1, dt) + us)[inner]
        uss = _implicit(bc(r, t1), 1, lam(1))
        r = uss.copy()
        r[inner] = (-L(u, 2, dt) + uss)[inner]
        return _implicit(bc(r, t1), 2, lam(1))
    if method in ("lod-ie", "lod-cn"):
        c = 1.0 if method == "lod-ie" else 0.5
        v = u
        for d in range(3):
            r = v.copy()
            if method == "lod-cn":
                r[inner] = (v + L(v, d, c * dt))[inner]
            v = _implicit(bc(r, t1), d, lam(c))
        return v + dt * f(t1)
    # ts
    th = t1 - 0.5 * dt
    v = u
    for d in range(3):
        r = np.zeros_like(v)
        r[inner] = (v + L(v, d, 0.5 * dt))[inner]
        v = bc(r, th)
    v = v + 0.5 * dt * f(t1 - dt)
    v = v + 0.5 * dt * f(t1)
    for d in (2, 1, 0):
        v = _implicit(bc(v, t1), d, lam(0.5))
    return v


@pytest.mark.parametrize("method", ["adi", "lod-ie", "lod-cn", "ts"])
def test_oracle_is_the_split_scheme_without_interface(mts, oracle, method):
    n = 8                      # nodes per axis; h = 2/7 so no node sits at the origin
    bout = 3.0
    st = mts.Setup(mts.grid(n), mts.surface("ellipsoid", 1, 1, 1, 0.01),
                   mts.physics("const", 10, bout), mib=1)
    assert all(st.info().n_cross[d] == 0 for d in range(3))
    op = oracle.OracleProblem.from_setup(st, 0, (10, bout), -1.0, 1.0)
    c = np.linspace(-1, 1, n)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    h = 2.0 / (n - 1)
    dt = 0.05
    u = op.init(0.0)
    v = u.copy()
    t = 0.0
    for _ in range(3):
        u, _ = op.run(method, dt, t, 1, u)
        t += dt
        v = _matrix_scheme(method, v, t, dt, bout, X, Y, Z, h)
    assert rel_err(u, v) < 1e-13
