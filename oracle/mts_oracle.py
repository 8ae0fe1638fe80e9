"""ORACLE binding (test infrastructure only).

ctypes wrapper of oracle/_build/libmts_oracle.so, the plain-C restatement of
the reference's per-step path, and of oracle/_ref/mts_ref, the reference
solver built from its own sources.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use this module, to check results against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmts_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "mts_ref")

ADI, LOD_IE, LOD_CN, TS = 0, 1, 2, 3
METHODS = {"adi": ADI, "lod-ie": LOD_IE, "lod-cn": LOD_CN, "ts": TS}


class _Problem(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("ind", C.c_void_p), ("beta_kind", C.c_int), ("bp", C.c_double * 2),
                ("ncross", C.c_int * 3), ("cr_int", C.c_void_p * 3),
                ("cr_dbl", C.c_void_p * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; run make -C oracle port")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(_Problem)
        L.mts_oracle_init.argtypes = [P, C.c_double, C.c_void_p]
        L.mts_oracle_init.restype = C.c_int
        L.mts_oracle_run.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                     C.POINTER(C.c_double)]
        L.mts_oracle_run.restype = C.c_int
        L.mts_oracle_adi_stages.argtypes = [P, C.c_double, C.c_double] + [C.c_void_p] * 4
        L.mts_oracle_adi_stages.restype = C.c_int
        L.mts_oracle_norms.argtypes = [P, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.mts_oracle_norms.restype = None
        L.mts_oracle_tdma.argtypes = [C.c_int] + [C.c_void_p] * 5
        L.mts_oracle_tdma.restype = C.c_int
        L.mts_oracle_set_line_solver.argtypes = [C.c_int]
        L.mts_oracle_set_line_solver.restype = None
        L.mts_oracle_line_solve.argtypes = [C.c_int] + [C.c_void_p] * 5
        L.mts_oracle_line_solve.restype = C.c_int
        _lib = L
    return _lib


class OracleProblem:
    """Geometry + crossings for the restatement.

    ``crossings[d]`` is a structured array with fields j, k, id, left_loc,
    coord(3), normal(3), wl(9), wr(9), sorted by line id (k * nj + j); the
    product's Setup.crossings() and the reference dump both provide that.
    """

    def __init__(self, n, lo, hi, indicator, beta_kind, bp, crossings):
        self.n = tuple(int(v) for v in n)
        self.ind = np.ascontiguousarray(indicator, dtype=np.int8)
        assert self.ind.size == self.n[0] * self.n[1] * self.n[2]
        self._keep = [self.ind]
        p = _Problem()
        for a in range(3):
            p.n[a] = self.n[a]
            p.lo[a] = float(lo[a] if np.ndim(lo) else lo)
            p.hi[a] = float(hi[a] if np.ndim(hi) else hi)
        p.ind = self.ind.ctypes.data
        p.beta_kind = int(beta_kind)
        p.bp[0], p.bp[1] = float(bp[0]), float(bp[1])
        for d in range(3):
            cr = crossings[d]
            ii = np.zeros((max(len(cr), 1), 4), dtype=np.int32)
            dd = np.zeros((max(len(cr), 1), 24), dtype=np.float64)
            if len(cr):
                ii[:len(cr), 0] = cr["j"]
                ii[:len(cr), 1] = cr["k"]
                ii[:len(cr), 2] = cr["id"]
                ii[:len(cr), 3] = cr["left_loc"]
                dd[:len(cr), 0:3] = cr["coord"]
                dd[:len(cr), 3:6] = cr["normal"]
                dd[:len(cr), 6:15] = cr["wl"]
                dd[:len(cr), 15:24] = cr["wr"]
            self._keep += [ii, dd]
            p.ncross[d] = len(cr)
            p.cr_int[d] = ii.ctypes.data
            p.cr_dbl[d] = dd.ctypes.data
        self._p = p

    @classmethod
    def from_setup(cls, setup, beta_kind, bp, lo, hi):
        info = setup.info()
        n = (info.n[0], info.n[1], info.n[2])
        return cls(n, lo, hi, setup.indicator(), beta_kind, bp,
                   [setup.crossings(d) for d in range(3)])

    @property
    def shape(self):
        return (self.n[2], self.n[1], self.n[0])

    def init(self, t0=0.0) -> np.ndarray:
        u = np.empty(self.shape)
        if lib().mts_oracle_init(C.byref(self._p), t0, u.ctypes.data):
            raise RuntimeError("oracle init failed")
        return u

    def run(self, method, dt, t0, nsteps, u):
        m = METHODS[method] if isinstance(method, str) else int(method)
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        t_end = C.c_double()
        if lib().mts_oracle_run(C.byref(self._p), m, dt, t0, nsteps, u.ctypes.data,
                                C.byref(t_end)):
            raise RuntimeError("oracle run failed (zero pivot or bad crossings)")
        return u, t_end.value

    def adi_stages(self, dt, t0, u):
        """One Douglas-ADI step from t0: (stage-1 right-hand side with boundary
        values, u* after the x sweep, u** after the y sweep, u^{n+1})."""
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        r1, us, us2 = (np.empty(self.shape) for _ in range(3))
        if lib().mts_oracle_adi_stages(C.byref(self._p), dt, t0, u.ctypes.data, r1.ctypes.data,
                                       us.ctypes.data, us2.ctypes.data):
            raise RuntimeError("oracle Douglas step failed")
        return r1, us, us2, u

    def norms(self, u, t):
        u = np.ascontiguousarray(u, dtype=np.float64)
        li, l2 = C.c_double(), C.c_double()
        am = (C.c_int * 3)()
        lib().mts_oracle_norms(C.byref(self._p), u.ctypes.data, t, C.byref(li), C.byref(l2), am)
        return li.value, l2.value, (am[0], am[1], am[2])


def tdma(a, b, c, r):
    a, b, c, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, r))
    u = np.empty_like(r)
    if lib().mts_oracle_tdma(len(r), a.ctypes.data, b.ctypes.data, c.ctypes.data, r.ctypes.data,
                             u.ctypes.data):
        raise ZeroDivisionError("zero pivot in tridiagonal solve")
    return u


SOLVERS = {"thomas": 0, "pivot": 1, "partitioned": 2}


class line_solver:
    """Context manager selecting the oracle's line solver ("thomas" is the
    reference's TDMA; "pivot" partial-pivoting elimination; "partitioned" the
    CPU model of the GPU's partitioned solve).  Restores Thomas on exit."""

    def __init__(self, kind):
        self.kind = SOLVERS[kind]

    def __enter__(self):
        lib().mts_oracle_set_line_solver(self.kind)
        return self

    def __exit__(self, *exc):
        lib().mts_oracle_set_line_solver(0)
        return False


def line_solve(kind, a, b, c, r):
    a, b, c, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, r))
    u = np.empty_like(r)
    with line_solver(kind):
        rc = lib().mts_oracle_line_solve(len(r), a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                         r.ctypes.data, u.ctypes.data)
    if rc:
        raise ZeroDivisionError("zero pivot or unsupported size in line solve")
    return u


# --------------------------------------------------------------- reference
def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def run_reference(outdir, method, mib, n, dt, T, surface, beta, lo=-1.0, hi=1.0,
                  max_steps=None, dump=True, dump_steps=0):
    """Run oracle/_ref/mts_ref and parse its dumps (see oracle/ref_harness.cc).
    dump_steps: also return the fields after steps 1..dump_steps as u_steps."""
    os.makedirs(outdir, exist_ok=True)
    cmd = [REF_BIN, "--method", method, "--mib", str(mib), "--n", str(n), "--lo", repr(lo),
           "--hi", repr(hi), "--dt", repr(dt), "--T", repr(T), "--surface", surface,
           "--beta", beta, "--out", outdir]
    if max_steps is not None:
        cmd += ["--max-steps", str(max_steps)]
    if not dump:
        cmd += ["--no-dump"]
    if dump_steps:
        cmd += ["--dump-steps", str(dump_steps)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    summ_path = os.path.join(outdir, "summary.txt")
    if proc.returncode != 0 or not os.path.exists(summ_path):
        raise RuntimeError(f"reference exited: {proc.stdout.strip()} {proc.stderr.strip()}")
    summary = {}
    with open(summ_path) as f:
        for line in f:
            parts = line.split()
            summary[parts[0]] = parts[1:] if len(parts) > 2 else parts[1]
    out = {"summary": summary}
    nn = n + 1
    if dump:
        for name in ("u_step1", "u_final"):
            p = os.path.join(outdir, name + ".bin")
            if os.path.exists(p):
                out[name] = np.fromfile(p, dtype=np.float64).reshape(nn, nn, nn)
        if dump_steps:
            out["u_steps"] = np.stack([
                np.fromfile(os.path.join(outdir, f"u_step{j}.bin"), dtype=np.float64)
                .reshape(nn, nn, nn) for j in range(1, dump_steps + 1)])
        out["crossings"] = parse_crossings(os.path.join(outdir, "crossings.txt"))
        jp = os.path.join(outdir, "jumps.txt")
        if os.path.exists(jp) and os.path.getsize(jp):
            out["jumps"] = np.loadtxt(jp, ndmin=2)
    return out


REF_CROSSING_DTYPE = np.dtype([
    ("dir", "i4"), ("j", "i4"), ("k", "i4"), ("id", "i4"), ("left_loc", "i4"),
    ("gamma", "f8"), ("coord", "f8", 3), ("normal", "f8", 3), ("beta_in", "f8"),
    ("beta_out", "f8"), ("p", "f8", 9), ("nw", "i4"), ("wl", "f8", 9), ("wr", "f8", 9)])


def parse_crossings(path):
    """Parse crossings.txt into one structured array per direction (line order)."""
    rows = {0: [], 1: [], 2: []}
    with open(path) as f:
        for line in f:
            t = line.split()
            d = "xyz".index(t[0])
            rec = np.zeros((), dtype=REF_CROSSING_DTYPE)
            rec["dir"] = d
            rec["j"], rec["k"], rec["id"], rec["left_loc"] = (int(v) for v in t[1:5])
            vals = [float(v) for v in t[5:]]
            rec["gamma"] = vals[0]
            rec["coord"] = vals[1:4]
            rec["normal"] = vals[4:7]
            rec["beta_in"], rec["beta_out"] = vals[7], vals[8]
            rec["p"] = vals[9:18]
            nw = int(vals[18])
            rec["nw"] = nw
            rec["wl"][:nw] = vals[19:19 + nw]
            rec["wr"][:nw] = vals[19 + nw:19 + 2 * nw]
            rows[d].append(rec)
    out = []
    for d in range(3):
        arr = np.array(rows[d], dtype=REF_CROSSING_DTYPE)
        if len(arr):
            nj = None
            arr = arr  # sorted below by line id once the grid is known by the caller
        out.append(arr)
    return out


def sort_by_line(cr, nj):
    """Order reference crossings by the product's line id k * nj + j (stable)."""
    if len(cr) == 0:
        return cr
    key = cr["k"].astype(np.int64) * nj + cr["j"]
    return cr[np.argsort(key, kind="stable")]
