"""Graph-captured time loops (mts_run_graph, mts_batch_run_graph): the same
kernels with the same arguments as the eager loop, so the fields are
bit-identical, on the first run (capture) and on replays of the kept graph
with eager work in between."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # surface, args, beta, intervals, method, mib, dt, t_finish, precision
    ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), ("sincos", 10, 1), 64, "lod-cn", 2, 1 / 32, 0.25, "fp64"),
    ("ellipsoid", (0.6, 0.5, 0.4, 1 - 2 ** -20), ("const", 10, 1), 40, "adi", 1, 1 / 40, 0.2, "fp64"),
    ("torus", (0.5, 0.2), ("const", 0.01, 1), 48, "ts", 2, 1 / 48, 0.125, "fp32"),
    ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), ("const", 1000, 1), 40, "lod-ie", 2, 1 / 20, 0.2, "fp64"),
]


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_graph_loop_matches_eager_loop(mts, ctx, ci):
    surf, sargs, beta, n, method, mib, dt, tf, prec = CASES[ci]
    print("graph case", surf, n, method, prec, flush=True)
    st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), mts.physics(*beta), mib=mib)
    inst = mts.Instance(ctx, st, method, dt, precision=prec)
    inst.initialize(0.0)
    eager = inst.run(0.0, tf)
    u_e = inst.download()
    for rep in range(3):
        inst.initialize(0.0)
        got = inst.run(0.0, tf, graph=True)
        assert got == eager
        assert np.array_equal(inst.download(), u_e), f"replay {rep}"
        # eager work between replays
        inst.step(dt)
        inst.sync()
    # a different end time captures a new graph
    inst.initialize(0.0)
    e2 = inst.run(0.0, tf / 2)
    u2 = inst.download()
    inst.initialize(0.0)
    assert inst.run(0.0, tf / 2, graph=True) == e2
    assert np.array_equal(inst.download(), u2)
    inst.close()


def test_graph_needs_profiling_off(mts, ctx):
    st = mts.Setup(mts.grid(33), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20),
                   mts.physics("const", 10, 1), mib=1)
    inst = mts.Instance(ctx, st, "lod-cn", 1 / 16)
    inst.initialize(0.0)
    inst.profile(True)
    with pytest.raises(mts.MtsError) as e:
        inst.run(0.0, 0.25, graph=True)
    assert e.value.code == -4
    inst.profile(False)
    assert inst.run(0.0, 0.25, graph=True)[0] == 4
    inst.close()


def test_batch_graph_matches_batch(mts, ctx):
    insts = []
    for surf, sargs, beta, n, method, mib, dt, tf, prec in CASES:
        st = mts.Setup(mts.grid(n + 1), mts.surface(surf, *sargs), mts.physics(*beta), mib=mib)
        insts.append(mts.Instance(ctx, st, method, dt, precision=prec))
    tfs = [c[7] for c in CASES]
    for i in insts:
        i.initialize(0.0)
    a = mts.batch_run(insts, 0.0, tfs)
    ua = [i.download() for i in insts]
    for rep in range(2):
        for i in insts:
            i.initialize(0.0)
        b = mts.batch_run(insts, 0.0, tfs, graph=True)
        assert a == b
        for i, u in zip(insts, ua):
            assert np.array_equal(i.download(), u), f"replay {rep}"
    for i in insts:
        i.close()
