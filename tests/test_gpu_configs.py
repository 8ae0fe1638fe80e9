"""The batched and served configurations at full scale, checked by property.

* The uneven batch: 96 independent solves drawn from seed 0 exactly as
  `bench.py --workload batch` draws them (N in {32, 48, ..., 192}; sphere
  r ~ U[0.3, 0.7], ellipsoid semi-axes ~ U[0.3, 0.8] or torus R = 0.5,
  r ~ U[0.15, 0.25]; beta- log-uniform in [1e-3, 1e3], beta+ = 1;
  Douglas-ADI or LOD-CN; MIB-L1; dt = h; T = 0.5).  Batched with
  mts_batch_run, every solve's final field and error norms equal the same
  solve run alone, bit for bit (its own buffers and stream).  The smallest
  solve of each method with beta- within 1e-2..1e2 is checked against the
  oracle at every step,
  restarted from the oracle's own previous field, at the golden tests'
  per-step bound (1e-12, or the step's own solver spread or rounding
  sensitivity where larger), and over the whole run at the step count times
  that bound times the run's measured amplification.  Only
  Douglas-ADI on interfaces with corner pairs is refused (MTS_E_DISABLED).
* The serving stream: requests drawn as the serving bench draws them (N
  uniform in [24, 256], T in {0.1, 0.25, 0.5, 1.0}, all four methods, MIB-L1
  or MIB-L2), submitted at once through the continuous-batching server with
  a 200 GB cap: every request finishes or is refused for a stated reason, and
  a sample finished alone reports the same steps, final time and norms.
"""
import importlib.util
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err, run_amplification, step_floor, step_times

pytestmark = pytest.mark.gpu


def _bench():
    spec = importlib.util.spec_from_file_location("mts_bench", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def _draws(count=96, seed=0):
    b = _bench()
    rng = np.random.default_rng(seed)
    return [b.draw_solve(rng, int(rng.choice(np.arange(32, 193, 16))),
                         ["adi", "lod-cn"][int(rng.integers(2))], 1, 0.5)
            for _ in range(count)]


def _setup(mts, d):
    phys = mts.physics("const", *d["beta"])
    st = mts.Setup(mts.grid(d["n"] + 1), mts.surface(d["surface"][0], *d["surface"][1]), phys,
                   mib=d["mib"])
    return st, phys


def test_gpu_uneven_batch_matches_alone(mts, oracle, ctx):
    draws = _draws()
    insts, kept, refused = [], [], 0
    for i, d in enumerate(draws):
        st, _ = _setup(mts, d)
        try:
            inst = mts.Instance(ctx, st, d["method"], d["dt"])
        except mts.MtsError as e:
            assert e.code == -6 and d["method"] == "adi" and sum(st.info().corner_pairs) > 0
            refused += 1
            continue
        inst.initialize(0.0)
        insts.append(inst)
        kept.append(i)
    print(f"batch: {len(insts)} solves, {refused} Douglas-ADI with corner pairs refused",
          flush=True)
    assert len(insts) + refused == 96 and len(insts) >= 48
    out = mts.batch_run(insts, 0.0, [draws[i]["T"] for i in kept])
    batched = []
    for inst, (steps, t_end) in zip(insts, out):
        batched.append((steps, t_end, inst.error(t_end), inst.download()))
    # each solve alone, on the same instance
    for k, (inst, i) in enumerate(zip(insts, kept)):
        inst.initialize(0.0)
        steps, t_end = inst.run(0.0, draws[i]["T"])
        assert (steps, t_end) == batched[k][:2], i
        assert inst.error(t_end) == batched[k][2], i
        assert np.array_equal(inst.download(), batched[k][3]), i
    # the smallest solve of each method with a contrast within 1e-2..1e2
    # against the oracle, per step and over the whole run (at stronger
    # contrast the rounding of the Douglas right-hand side's large terms alone
    # reaches 1e-12 of the field: the golden and per-step tests bound those
    # cases by their measured solver spread and sensitivity)
    for method in ("adi", "lod-cn"):
        cand = sorted((draws[i]["n"], k) for k, i in enumerate(kept)
                      if draws[i]["method"] == method and 1e-2 <= draws[i]["beta"][0] <= 1e2)[:1]
        for _, k in cand:
            d = draws[kept[k]]
            st, phys = _setup(mts, d)
            op = oracle.OracleProblem.from_setup(st, 0, d["beta"], -1.0, 1.0)
            steps = batched[k][0]
            inst = insts[k]
            # every step from the oracle's own previous field, at the golden
            # tests' per-step bound (1e-12, or the step's solver spread or
            # rounding sensitivity where larger)
            u_ref, floors, ts = op.init(0.0), [], step_times(d["dt"], steps)
            for j in range(steps):
                floor = step_floor(oracle, op, method, d["dt"], ts[j], u_ref)
                floors.append(floor)
                inst.upload(u_ref)
                inst.step(ts[j + 1])
                u_ref, _ = op.run(method, d["dt"], ts[j], 1, u_ref)
                e = rel_err(inst.download(), u_ref)
                assert e <= floor, (method, d["n"], j + 1, e, floor)
            # the whole run: a difference made at each step, grown by at most
            # the run's amplification
            tol = steps * max(floors) * run_amplification(op, method, d["dt"], steps)
            e = rel_err(batched[k][3], u_ref)
            print(f"  oracle {method} N={d['n']} {steps} steps: per step <= {max(floors):.1e}, "
                  f"whole run {e:.2e} (bound {tol:.1e})", flush=True)
            assert e <= tol
    for inst in insts:
        inst.close()


def test_gpu_serving_stream_completes(mts, ctx):
    b = _bench()
    rng = np.random.default_rng(1)
    reqs = []
    for i in range(200):
        d = b.draw_solve(rng, int(rng.integers(24, 257)),
                         ["adi", "lod-ie", "lod-cn", "ts"][int(rng.integers(4))],
                         int(rng.integers(1, 3)), float(rng.choice([0.1, 0.25, 0.5, 1.0])))
        reqs.append(d)
    srv = mts.Server(ctx, 200 << 30, 32, 8)
    for i, d in enumerate(reqs):
        srv.submit(mts.request(mts.grid(d["n"] + 1), mts.surface(d["surface"][0],
                                                                 *d["surface"][1]),
                               mts.physics("const", *d["beta"]), d["method"], d["dt"], d["T"],
                               mib=d["mib"], tag=i))
    res = {r["tag"]: r for r in srv.drain().values()}
    srv.close()
    assert sorted(res) == list(range(200))
    ok = [t for t, r in res.items() if r["status"] == 0]
    bad = {t: (r["status"], r["message"]) for t, r in res.items() if r["status"] != 0}
    print(f"serving: {len(ok)} finished, {len(bad)} refused", flush=True)
    for t, (code, msg) in bad.items():
        # refusals the library states: Douglas-ADI with corner pairs, or a
        # geometry the reference itself cannot set up
        assert code in (-6, -2), (t, code, msg)
        if code == -6:
            assert reqs[t]["method"] == "adi"
    assert len(ok) >= 150
    # a sample alone, on fresh instances
    for t in ok[:: max(1, len(ok) // 8)]:
        d = reqs[t]
        st, _ = _setup(mts, d)
        inst = mts.Instance(ctx, st, d["method"], d["dt"])
        inst.initialize(0.0)
        steps, t_end = inst.run(0.0, d["T"])
        linf, l2, am = inst.error(t_end)
        r = res[t]
        assert (r["steps"], r["t_end"]) == (steps, t_end)
        assert (r["linf"], r["l2"], tuple(r["argmax"])) == (linf, l2, tuple(am))
        inst.close()
