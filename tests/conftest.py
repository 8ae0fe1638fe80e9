"""Shared fixtures: the product binding, the oracle binding, golden vectors."""
import glob
import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "matched-time-splitting-3d_amd")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X")


def load_mts():
    mod = sys.modules.get("mts_amd")
    if mod is None:
        spec = importlib.util.spec_from_file_location(
            "mts_amd", os.path.join(PKG_DIR, "__init__.py"),
            submodule_search_locations=[PKG_DIR])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["mts_amd"] = mod
        spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mts():
    return load_mts()


@pytest.fixture(scope="session")
def oracle():
    import mts_oracle
    return mts_oracle


@pytest.fixture(scope="session")
def ctx(mts):
    # The per-step path runs on the GPU only: no device is an error, not a skip.
    return mts.Context(0)


def rel_err(a, b):
    """Normwise relative difference max|a - b| / max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


GOLDEN_NAMES = sorted(os.path.splitext(os.path.basename(p))[0]
                      for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False)
    g = {k: z[k] for k in z.files}
    g["meta"] = json.loads(str(g["meta"]))
    g["u_step1"] = g["u_steps"][0]
    return g


def golden_crossings(g, d):
    """Golden crossings of direction d as a structured array (oracle field names)."""
    ii, dd = g[f"cr{d}_int"], g[f"cr{d}_dbl"]
    dt = np.dtype([("j", "i4"), ("k", "i4"), ("id", "i4"), ("left_loc", "i4"), ("nw", "i4"),
                   ("gamma", "f8"), ("coord", "f8", 3), ("normal", "f8", 3),
                   ("beta_in", "f8"), ("beta_out", "f8"), ("p", "f8", 9), ("wl", "f8", 9),
                   ("wr", "f8", 9)])
    out = np.zeros(len(ii), dtype=dt)
    if len(ii):
        out["j"], out["k"], out["id"], out["left_loc"], out["nw"] = ii.T
        out["gamma"] = dd[:, 0]
        out["coord"] = dd[:, 1:4]
        out["normal"] = dd[:, 4:7]
        out["beta_in"] = dd[:, 7]
        out["beta_out"] = dd[:, 8]
        out["p"] = dd[:, 9:18]
        out["wl"] = dd[:, 18:27]
        out["wr"] = dd[:, 27:36]
    return out


def case_setup(mts, meta):
    """Product Setup for a golden case's geometry."""
    n = meta["n"] + 1
    s = meta["surface"]
    b = meta["beta"]
    phys = mts.physics(b[0], *b[1:])
    st = mts.Setup(mts.grid(n), mts.surface(s[0], *s[1:]), phys, mib=meta["mib"])
    return st, phys


def beta_kind(meta):
    return {"const": 0, "sincos": 1}[meta["beta"][0]]


PARITY = 1e-12  # relative bar of the plain FP64 form against the reference


# field passes of one step, each a place where an implementation can round
# differently from the reference: Douglas and LOD three line sweeps, TS three
# explicit and three implicit ones
STEP_PASSES = {"adi": 3, "lod-ie": 3, "lod-cn": 3, "ts": 6}


def step_floor(oracle, op, method, dt, t, u_in, seeds=(0, 1)):
    """Per-step parity bound for one step from u_in at time t: 1e-12, or,
    where the step itself makes rounding matter more, the larger of
      * four times the spread between two backward-stable elimination orders
        of the step's line systems (the reference's Thomas and partial
        pivoting, both in the oracle): any stable line solver other than the
        reference's own Thomas order can differ from it by that much;
      * the step's measured sensitivity to one rounding (the change of its
        output when its input is perturbed by one ulp, a random sign per node,
        worst of two patterns) once per field pass of the step: what an
        implementation that rounds each pass's values differently can move it
        by.  This is large only where explicit half-steps with a strong
        contrast amplify (TS with beta 1000/1 at dt = h)."""
    u_t, _ = op.run(method, dt, t, 1, u_in)
    with oracle.line_solver("pivot"):
        u_p, _ = op.run(method, dt, t, 1, u_in)
    sens = 0.0
    for seed in seeds:
        rng = np.random.default_rng(seed)
        u_e = u_in * (1.0 + 2.0 ** -52 * rng.choice([-1.0, 1.0], np.shape(u_in)))
        u_s, _ = op.run(method, dt, t, 1, u_e)
        sens = max(sens, rel_err(u_s, u_t))
    return max(PARITY, 4.0 * rel_err(u_p, u_t), STEP_PASSES[method] * sens)


def run_amplification(op, method, dt, steps, seed=0):
    """Measured amplification of a relative perturbation of the initial field
    over the whole run (a random sign pattern of relative size 1e-9, mapped
    through the oracle): how much a difference made at the first step can grow
    by the end, whatever the solution's own growth."""
    u0 = op.init(0.0)
    uf, _ = op.run(method, dt, 0.0, steps, u0)
    rng = np.random.default_rng(seed)
    up = u0 * (1.0 + 1e-9 * rng.choice([-1.0, 1.0], u0.shape))
    ufp, _ = op.run(method, dt, 0.0, steps, up)
    return max(1.0, rel_err(ufp, uf) / rel_err(up, u0))


def golden_oracle(mts, oracle, g):
    """Oracle problem on a golden case's own crossings and weights."""
    meta = g["meta"]
    n = meta["n"] + 1
    s, b = meta["surface"], meta["beta"]
    phys = mts.physics(b[0], *b[1:])
    st = mts.Setup(mts.grid(n), mts.surface(s[0], *s[1:]), phys, mib=meta["mib"])
    return oracle.OracleProblem((n, n, n), -1.0, 1.0, st.indicator(), beta_kind(meta), b[1:],
                                [golden_crossings(g, d) for d in range(3)])


def step_times(dt, count):
    """t^0 .. t^count accumulated as the reference's loop does (t += dt)."""
    ts = [0.0]
    for _ in range(count):
        ts.append(ts[-1] + dt)
    return ts
