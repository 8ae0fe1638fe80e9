"""mts_batch_run_multi: the instances of one method and precision step in
lockstep and each solve sweep runs as one launch over the tiles of all of
them.  Every solve's steps, final time and field equal the same solve run
alone with mts_run, bit for bit (the same kernels' arithmetic), over a mix of
all four methods, both MIB orders, both precisions and grids of 25-97 nodes
per axis (so the line splits and register bounds of one sweep differ)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_multi_matches_alone(mts, ctx):
    rng = np.random.default_rng(3)
    methods = ["adi", "lod-cn", "lod-ie", "ts"]
    insts, Ts = [], []
    for i in range(20):
        n = int(rng.choice([24, 32, 48, 64, 96]))
        method = methods[i % 4]
        prec = "fp32" if i % 5 == 4 else "fp64"
        r = float(rng.uniform(0.3, 0.6))
        st = mts.Setup(mts.grid(n + 1), mts.surface("ellipsoid", 1, 1, 1, r),
                       mts.physics("const", float(10 ** rng.uniform(-1, 1)), 1.0),
                       mib=int(rng.integers(1, 3)))
        try:
            inst = mts.Instance(ctx, st, method, 2.0 / n, precision=prec)
        except mts.MtsError as e:
            assert e.code == -6 and method == "adi"  # Douglas-ADI with corner pairs
            continue
        insts.append(inst)
        Ts.append(float(rng.choice([0.125, 0.25])))
    print(f"batch multi: {len(insts)} solves", flush=True)
    assert len(insts) >= 15
    for inst in insts:
        inst.initialize(0.0)
    out = mts.batch_run(insts, 0.0, Ts, multi=True)
    fields = [inst.download() for inst in insts]
    for k, inst in enumerate(insts):
        inst.initialize(0.0)
        steps, t_end = inst.run(0.0, Ts[k])
        assert (steps, t_end) == out[k], k
        assert np.array_equal(inst.download(), fields[k]), k
    for inst in insts:
        inst.close()
