"""The line-split autotuner (mts_tune_lanes): it times every split that fits
the grid, keeps the fastest in the context, and later instances of the same
grid, method and precision run it.  Every split computes the same solution to
rounding: each one is checked against the oracle step by step at the golden
tests' per-step bound (1e-12, or the step's solver spread or rounding
sensitivity where larger)."""
import os

import pytest

from conftest import rel_err, step_floor, step_times

pytestmark = pytest.mark.gpu

SPLITS = [(16, 16), (16, 32), (16, 64), (32, 32), (32, 64), (64, 64)]


def _setup(mts):
    return mts.Setup(mts.grid(65), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20),
                     mts.physics("sincos", 10, 1), mib=2)


def test_tune_lanes_choice_is_used_by_later_instances(mts, ctx):
    st = _setup(mts)
    ml, gl, ms = ctx.tune_lanes(st, "lod-cn", 1 / 32, steps=3)
    print(f"tuned split {ml}/{gl}: {ms:.4f} ms per step", flush=True)
    assert (ml, gl) in SPLITS and ms > 0
    inst = mts.Instance(ctx, st, "lod-cn", 1 / 32)
    assert inst.lanes() == (ml, gl)
    inst.close()
    # another method or precision keeps the default
    other = mts.Instance(ctx, st, "lod-ie", 1 / 32)
    assert other.lanes() == (16, 16)
    other.close()


@pytest.mark.parametrize("split", [(32, 32), (16, 64), (64, 64)])
def test_every_split_matches_the_oracle(mts, oracle, ctx, split, monkeypatch):
    print(f"split {split}", flush=True)
    monkeypatch.setenv("MTS_LINE_LANES", str(split[0]))
    monkeypatch.setenv("MTS_GEN_LANES", str(split[1]))
    st = _setup(mts)
    inst = mts.Instance(mts.Context(0), st, "lod-cn", 1 / 32)
    assert inst.lanes() == split
    op = oracle.OracleProblem.from_setup(st, 1, (10, 1), -1.0, 1.0)
    u = op.init(0.0)
    ts = step_times(1 / 32, 3)
    for j in range(3):
        floor = step_floor(oracle, op, "lod-cn", 1 / 32, ts[j], u)
        inst.upload(u)
        inst.step(ts[j + 1])
        u, _ = op.run("lod-cn", 1 / 32, ts[j], 1, u)
        e = rel_err(inst.download(), u)
        assert e <= floor, (split, j + 1, e, floor)
    inst.close()
