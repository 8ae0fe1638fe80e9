"""Pin the C restatement (oracle/) against the reference build's fields.

The oracle takes the reference's own crossings and weights (golden fixtures)
and must reproduce the reference's solution to 1e-12 relative (max-norm
difference over the field's max-norm):
  * after the first step from the exact initial condition,
  * after every further step the fixture holds, restarted from the
    reference's own field of the previous step,
  * after the whole run, where the bound is 1e-12 times the run's measured
    amplification of a relative perturbation of the initial field (the
    bumpy-sphere TS case at beta 1000/1 amplifies by about 90; stable runs
    by at most a few).
Its only deliberate difference, u_dir in closed form, moves results by
rounding.  The fixtures include Douglas-ADI on interfaces with corner pairs
(torus22_adi_*, cfg1like32_adi_*), which pin the oracle's corner stencils and
corner jump terms.
"""
import pytest

from conftest import (GOLDEN_NAMES, PARITY, golden_oracle, load_golden, rel_err,
                      run_amplification, step_times)


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_oracle_reproduces_reference(mts, oracle, name):
    g = load_golden(name)
    meta = g["meta"]
    op = golden_oracle(mts, oracle, g)
    method, dt = meta["method"], meta["dt"]
    us = g["u_steps"]
    ts = step_times(dt, len(us))
    u0 = op.init(0.0)
    u1, _ = op.run(method, dt, 0.0, 1, u0)
    assert rel_err(u1, us[0]) <= PARITY
    for j in range(1, len(us)):
        uj, _ = op.run(method, dt, ts[j], 1, us[j - 1])
        assert rel_err(uj, us[j]) <= PARITY, f"step {j + 1}"
    steps = int(g["summary"][0])
    uf, t_end = op.run(method, dt, 0.0, steps, u0)
    assert t_end == g["summary"][1]
    tol = PARITY * run_amplification(op, method, dt, steps)
    assert rel_err(uf, g["u_final"]) <= tol
    linf, l2, am = op.norms(uf, t_end)
    assert abs(linf - g["summary"][2]) <= tol * max(1.0, abs(g["summary"][2]))
    assert abs(l2 - g["summary"][3]) <= tol * max(1.0, abs(g["summary"][3]))


def test_fixtures_cover_douglas_corners(mts):
    """The Douglas-ADI fixtures include interfaces with corner pairs."""
    cases = [n for n in GOLDEN_NAMES if load_golden(n)["meta"]["method"] == "adi"]
    corners = [n for n in cases if any((load_golden(n)[f"cr{d}_int"][:, 2] < 0).any()
                                       for d in range(3))]
    assert len(corners) >= 4, corners
