"""Batched stepping (mts_batch_run) and the serving loop (mts_server_*).

Each solve runs on its own stream with its own buffers, so its arithmetic does
not depend on what else is in flight: every result must be bit-identical
whether the solve ran alone (Instance.run, or a server admitting one solve at a
time) or in a mix with others joining between its steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ELL = ("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20))
ELL_ADI = ("ellipsoid", (0.6, 0.5, 0.4, 1 - 2 ** -20))  # no corner pairs at 32 and 40 intervals
TORUS = ("torus", (0.5, 0.2))
BUMPY = ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3))

# (surface, beta, intervals, mib, method, dt, t_finish)
SOLVES = [
    (ELL, ("sincos", 10, 1), 48, 2, "lod-cn", 1 / 32, 0.25),
    (TORUS, ("const", 0.01, 1), 48, 1, "ts", 1 / 24, 0.125),
    (ELL_ADI, ("const", 10, 1), 40, 1, "adi", 1 / 40, 0.2),
    (BUMPY, ("const", 1000, 1), 40, 2, "lod-ie", 1 / 20, 0.2),
    (ELL, ("const", 0.001, 1), 64, 1, "lod-cn", 1 / 64, 0.25),
    (ELL_ADI, ("sincos", 10, 1), 32, 2, "adi", 1 / 32, 0.125),
    (TORUS, ("const", 0.01, 1), 32, 2, "lod-ie", 1 / 32, 0.25),
    (BUMPY, ("const", 1000, 1), 32, 1, "ts", 1 / 32, 0.125),
]


def _request(mts, i, solve):
    (sname, sargs), (beta, *bp), n, mib, method, dt, tf = solve
    return mts.request(mts.grid(n + 1), mts.surface(sname, *sargs), mts.physics(beta, *bp),
                       method, dt, tf, mib=mib, tag=i)


def _alone(mts, ctx, solve):
    (sname, sargs), (beta, *bp), n, mib, method, dt, tf = solve
    st = mts.Setup(mts.grid(n + 1), mts.surface(sname, *sargs), mts.physics(beta, *bp), mib=mib)
    inst = mts.Instance(ctx, st, method, dt)
    inst.initialize(0.0)
    steps, t_end = inst.run(0.0, tf)
    linf, l2, am = inst.error(t_end)
    out = {"steps": steps, "t_end": t_end, "linf": linf, "l2": l2, "argmax": am,
           "u": inst.download(), "bytes": inst.device_bytes()}
    inst.close()
    return out


@pytest.fixture(scope="module")
def alone(mts, ctx):
    res = []
    for i, s in enumerate(SOLVES):
        print(f"alone solve {i}: {s[4]} n={s[2]} mib={s[3]}", flush=True)
        res.append(_alone(mts, ctx, s))
    return res


def _same(r, a):
    assert r["status"] == 0, r["message"]
    assert r["steps"] == a["steps"]
    assert r["t_end"] == a["t_end"]
    assert r["linf"] == a["linf"]
    assert r["l2"] == a["l2"]
    assert tuple(r["argmax"]) == tuple(a["argmax"])


def test_batch_run_matches_run(mts, ctx, alone):
    insts, setups = [], []
    for (sname, sargs), (beta, *bp), n, mib, method, dt, tf in SOLVES:
        st = mts.Setup(mts.grid(n + 1), mts.surface(sname, *sargs), mts.physics(beta, *bp),
                       mib=mib)
        setups.append(st)
        inst = mts.Instance(ctx, st, method, dt)
        inst.initialize(0.0)
        insts.append(inst)
    print("batch_run", len(insts), flush=True)
    out = mts.batch_run(insts, 0.0, [s[6] for s in SOLVES])
    for i, inst in enumerate(insts):
        assert out[i] == (alone[i]["steps"], alone[i]["t_end"])
        assert np.array_equal(inst.download(), alone[i]["u"]), f"solve {i}"
        inst.close()


def test_batch_run_of_nothing(mts):
    assert mts.batch_run([], [], []) == []


def test_server_alone_and_in_the_mix_agree(mts, ctx, alone):
    reqs = [_request(mts, i, s) for i, s in enumerate(SOLVES)]
    # one solve at a time
    print("server max_inflight=1", flush=True)
    s1 = mts.Server(ctx, 64 << 30, max_inflight=1, setup_threads=1)
    for q in reqs:
        s1.submit(q)
    one = {r["tag"]: r for r in s1.drain().values()}
    s1.close()
    # everything at once, half of it submitted while the first half runs
    print("server max_inflight=8", flush=True)
    s8 = mts.Server(ctx, 64 << 30, max_inflight=8, setup_threads=4)
    for q in reqs[:4]:
        s8.submit(q)
    s8.wait(3)
    for q in reqs[4:]:
        s8.submit(q)
    mix = {r["tag"]: r for r in s8.drain().values()}
    s8.close()
    assert sorted(one) == sorted(mix) == list(range(len(SOLVES)))
    for i in range(len(SOLVES)):
        _same(one[i], alone[i])
        _same(mix[i], alone[i])
        assert mix[i]["footprint"] == alone[i]["bytes"]
        for r in (one[i], mix[i]):
            assert 0.0 <= r["queue_ms"] <= r["latency_ms"]
            assert 0.0 < r["run_ms"] <= r["latency_ms"]


def test_server_admission_cap(mts, ctx, alone):
    reqs = [_request(mts, i, s) for i, s in enumerate(SOLVES[:4])]
    # room for the largest solve only: they run one after another, results unchanged
    cap = max(a["bytes"] for a in alone[:4]) + (16 << 20)
    print("server cap", cap, flush=True)
    s = mts.Server(ctx, cap, max_inflight=8, setup_threads=2)
    for q in reqs:
        s.submit(q)
    got = {r["tag"]: r for r in s.drain().values()}
    s.close()
    for i in range(4):
        _same(got[i], alone[i])
    # a cap below any solve's footprint refuses them all with MTS_E_NOMEM
    s = mts.Server(ctx, 1 << 20, max_inflight=8, setup_threads=2)
    for q in reqs:
        s.submit(q)
    got = list(s.drain().values())
    s.close()
    assert len(got) == 4 and all(r["status"] == -5 for r in got)


def test_server_reports_failed_solves(mts, ctx):
    s = mts.Server(ctx, 64 << 30, max_inflight=4, setup_threads=2)
    # interface through a grid node: the reference exits, setup returns MTS_E_GEOMETRY
    bad = mts.request(mts.grid(25), mts.surface("ellipsoid", 0.5, 0.5, 0.5, 1.0),
                      mts.physics("const", 10, 1), "lod-cn", 1 / 24, 0.125, tag=1)
    # Douglas-ADI with corner pairs is switched off: MTS_E_DISABLED, no kernel runs
    off = mts.request(mts.grid(65), mts.surface(ELL[0], *ELL[1]), mts.physics("sincos", 10, 1),
                      "adi", 1 / 32, 0.125, mib=1, tag=2)
    good = _request(mts, 3, SOLVES[6])
    for q in (bad, off, good):
        s.submit(q)
    got = {r["tag"]: r for r in s.drain().values()}
    s.close()
    assert got[1]["status"] == -2 and "grid node" in got[1]["message"]
    assert got[2]["status"] == -6
    assert got[3]["status"] == 0 and got[3]["steps"] == 8
