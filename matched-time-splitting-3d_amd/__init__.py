"""Python binding of the MI355X Matched-TS per-step path (ctypes over libmts.so).

The product is the C ABI in ``include/mts.h`` (HIP kernels for gfx950, host
FP64 setup, C++ drop-in classes).  This module is a thin ctypes layer used by
the tests and the benchmark; it mirrors the reference's objects:

* :class:`Setup`     -- ``Mesh`` + ``Intersections`` (reference mesh/mesh.cc,
  spatial/intersection.cc): node classification, crossings, MIB weights.
* :class:`Instance`  -- ``Douglas_ADI`` / ``LOD`` / ``TS`` (reference
  temporal/*.cc): a device-resident solver with ``initialize``, ``step``
  (``Solve_2nd``), ``run`` (the driver loops of helper.h) and ``error``.

Nothing here computes anything on the CPU: every numerical call goes to the
library, and the library runs the per-step path on the GPU only.  Loading
fails loudly when the library has not been built.

The directory name is not a Python identifier, so load it with
``importlib.util.spec_from_file_location("mts_amd", <dir>/__init__.py)``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmts.so")

ADI, LOD_IE, LOD_CN, TS = 0, 1, 2, 3
METHODS = {"adi": ADI, "lod-ie": LOD_IE, "lod-cn": LOD_CN, "ts": TS}
SURF_ELLIPSOID, SURF_TORUS, SURF_BUMPY, SURF_CALLBACK = 0, 1, 2, 3
BETA_CONST, BETA_SINCOS = 0, 1
LAYOUT_XFAST, LAYOUT_REF = 0, 1
FP64, FP32 = 0, 1
PRECISIONS = {"fp64": FP64, "fp32": FP32}


class MtsError(RuntimeError):
    """A failed call of the C ABI; ``code`` is the MTS_E_* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"mts error {code}: {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("n", C.c_int * 3)]


PHI_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_double, C.c_double)
GRAD_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_double, C.c_double,
                      C.POINTER(C.c_double))


class SurfaceDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_double * 8), ("phi", PHI_FN), ("grad", GRAD_FN),
                ("user", C.c_void_p)]


class PhysicsDesc(C.Structure):
    _fields_ = [("beta_kind", C.c_int), ("bp", C.c_double * 4)]


class Crossing(C.Structure):
    _fields_ = [("dir", C.c_int), ("j", C.c_int), ("k", C.c_int), ("id", C.c_int),
                ("left_loc", C.c_int), ("nw", C.c_int), ("gamma", C.c_double),
                ("coord", C.c_double * 3), ("normal", C.c_double * 3), ("p", C.c_double * 9),
                ("pinv_d0", C.c_double), ("beta_in", C.c_double), ("beta_out", C.c_double),
                ("wl", C.c_double * 9), ("wr", C.c_double * 9)]


class StageTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int), ("total_ms", C.c_double)]


class ReferenceStops(C.Structure):
    _fields_ = [("count", C.c_int * 4), ("first_kind", C.c_int), ("first_dir", C.c_int),
                ("first_index", C.c_int), ("message", C.c_char * 160)]


class SetupInfo(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("n_cross", C.c_longlong * 3),
                ("corner_pairs", C.c_int * 3), ("singular_warnings", C.c_int), ("mib", C.c_int)]


class Request(C.Structure):
    _fields_ = [("grid", GridDesc), ("surface", SurfaceDesc), ("physics", PhysicsDesc),
                ("mib", C.c_int), ("method", C.c_int), ("dt", C.c_double), ("t0", C.c_double),
                ("t_finish", C.c_double), ("tag", C.c_longlong), ("precision", C.c_int)]


class Result(C.Structure):
    _fields_ = [("id", C.c_longlong), ("tag", C.c_longlong), ("status", C.c_int),
                ("message", C.c_char * 160), ("steps", C.c_int), ("t_end", C.c_double),
                ("linf", C.c_double), ("l2", C.c_double), ("argmax", C.c_int * 3),
                ("setup_ms", C.c_double), ("queue_ms", C.c_double), ("run_ms", C.c_double),
                ("latency_ms", C.c_double), ("footprint", C.c_size_t)]

    def as_dict(self) -> dict:
        return {"id": self.id, "tag": self.tag, "status": self.status,
                "message": self.message.decode(), "steps": self.steps, "t_end": self.t_end,
                "linf": self.linf, "l2": self.l2, "argmax": tuple(self.argmax),
                "setup_ms": self.setup_ms, "queue_ms": self.queue_ms, "run_ms": self.run_ms,
                "latency_ms": self.latency_ms, "footprint": self.footprint}


_lib_handle = None


def lib() -> C.CDLL:
    """The loaded libmts.so (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        P, I, D, VP = C.POINTER, C.c_int, C.c_double, C.c_void_p
        sig = {
            "mts_last_error": ([], C.c_char_p),
            "mts_version": ([], C.c_char_p),
            "mts_setup_create": ([P(GridDesc), P(SurfaceDesc), P(PhysicsDesc), I, I, P(VP)], I),
            "mts_setup_destroy": ([VP], None),
            "mts_setup_get_info": ([VP, P(SetupInfo)], I),
            "mts_setup_reference_stops": ([VP, P(ReferenceStops)], I),
            "mts_setup_indicator": ([VP, VP, C.c_size_t], I),
            "mts_setup_coords": ([VP, I, VP, C.c_size_t], I),
            "mts_setup_crossings": ([VP, I, VP, C.c_size_t], I),
            "mts_setup_diagnostics": ([VP, D, VP, VP], I),
            "mts_setup_crossing_errors": ([VP, I, D, VP, VP, VP, C.c_size_t], I),
            "mts_ctx_create": ([I, P(VP)], I),
            "mts_ctx_destroy": ([VP], None),
            "mts_ctx_device_info": ([VP, C.c_char_p, C.c_size_t, P(C.c_size_t)], I),
            "mts_instance_create": ([VP, VP, P(PhysicsDesc), I, D, P(VP)], I),
            "mts_instance_destroy": ([VP], None),
            "mts_field_upload": ([VP, VP, I], I),
            "mts_field_download": ([VP, VP, I], I),
            "mts_initialize": ([VP, D], I),
            "mts_step": ([VP, D], I),
            "mts_run": ([VP, D, D, P(I), P(D)], I),
            "mts_sync": ([VP], I),
            "mts_error_norms": ([VP, D, P(D), P(D), P(I)], I),
            "mts_error_norms_noalloc": ([VP, D, P(D), P(D), P(I)], I),
            "mts_field_device": ([VP, P(VP), P(I), P(VP)], I),
            "mts_tdma_batched": ([VP, I, I, VP, VP, VP, VP, VP], I),
            "mts_profile_enable": ([VP, I], I),
            "mts_profile_reset": ([VP], I),
            "mts_profile_read": ([VP, VP, I, P(I)], I),
            "mts_instance_dt": ([VP, P(D)], I),
            "mts_instance_create_ex": ([VP, VP, P(PhysicsDesc), I, D, I, P(VP)], I),
            "mts_instance_precision": ([VP, P(I)], I),
            "mts_tune_lanes": ([VP, VP, P(PhysicsDesc), I, D, I, I, P(I), P(I), P(D)], I),
            "mts_instance_diagnostics": ([VP, D, VP, VP], I),
            "mts_instance_crossing_errors": ([VP, I, D, VP, VP, VP, C.c_size_t], I),
            "mts_instance_lanes": ([VP, P(I), P(I)], I),
            "mts_instance_device_bytes": ([VP, P(C.c_size_t)], I),
            "mts_batch_run": ([VP, I, VP, VP, VP, VP], I),
            "mts_batch_run_graph": ([VP, I, VP, VP, VP, VP], I),
            "mts_batch_run_multi": ([VP, I, VP, VP, VP, VP], I),
            "mts_run_graph": ([VP, D, D, P(I), P(D)], I),
            "mts_server_create": ([VP, C.c_size_t, I, I, P(VP)], I),
            "mts_server_submit": ([VP, P(Request), P(C.c_longlong)], I),
            "mts_server_poll": ([VP, VP, I, P(I)], I),
            "mts_server_wait": ([VP, C.c_longlong], I),
            "mts_server_destroy": ([VP], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib_handle = L
    return _lib_handle


def _check(rc: int):
    if rc != 0:
        raise MtsError(rc, lib().mts_last_error().decode())


def _dptr(a: np.ndarray) -> int:
    return a.ctypes.data


def surface(kind: str, *params: float) -> SurfaceDesc:
    """Surface descriptor: ``ellipsoid`` (a, b, c, r), ``torus`` (R, r),
    ``bumpy`` (R0, A, m, k)."""
    s = SurfaceDesc()
    s.kind = {"ellipsoid": SURF_ELLIPSOID, "torus": SURF_TORUS, "bumpy": SURF_BUMPY}[kind]
    for i, v in enumerate(params):
        s.p[i] = float(v)
    return s


def physics(beta: str, *params: float) -> PhysicsDesc:
    """Diffusion coefficient: ``const`` (beta-, beta+) or ``sincos`` (c, beta+)."""
    p = PhysicsDesc()
    p.beta_kind = {"const": BETA_CONST, "sincos": BETA_SINCOS}[beta]
    for i, v in enumerate(params):
        p.bp[i] = float(v)
    return p


def grid(n, lo=-1.0, hi=1.0) -> GridDesc:
    """Grid of n[a] nodes per axis on [lo, hi]^3 (n may be an int)."""
    g = GridDesc()
    ns = (n, n, n) if isinstance(n, int) else tuple(n)
    los = (lo,) * 3 if np.isscalar(lo) else tuple(lo)
    his = (hi,) * 3 if np.isscalar(hi) else tuple(hi)
    for a in range(3):
        g.n[a] = int(ns[a])
        g.lo[a] = float(los[a])
        g.hi[a] = float(his[a])
    return g


CROSSING_DTYPE = np.dtype([
    ("dir", "i4"), ("j", "i4"), ("k", "i4"), ("id", "i4"), ("left_loc", "i4"), ("nw", "i4"),
    ("gamma", "f8"), ("coord", "f8", 3), ("normal", "f8", 3), ("p", "f8", 9),
    ("pinv_d0", "f8"), ("beta_in", "f8"), ("beta_out", "f8"), ("wl", "f8", 9),
    ("wr", "f8", 9)], align=True)
assert CROSSING_DTYPE.itemsize == C.sizeof(Crossing)


class Setup:
    """Mesh + Intersections: host FP64 setup of one geometry."""

    def __init__(self, g: GridDesc, surf: SurfaceDesc, phys: PhysicsDesc, mib: int = 1,
                 accuracy: int = 2):
        self._h = C.c_void_p()
        self.grid, self.surface, self.physics = g, surf, phys
        _check(lib().mts_setup_create(C.byref(g), C.byref(surf), C.byref(phys), mib, accuracy,
                                      C.byref(self._h)))

    def close(self):
        if self._h:
            lib().mts_setup_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self) -> SetupInfo:
        i = SetupInfo()
        _check(lib().mts_setup_get_info(self._h, C.byref(i)))
        return i

    def reference_stops(self) -> dict:
        """Where the reference's own run would have stopped (exit(0) in its
        tangential-jump set-up or first step); first_kind -1: it would run."""
        r = ReferenceStops()
        _check(lib().mts_setup_reference_stops(self._h, C.byref(r)))
        return {"count": list(r.count), "first_kind": r.first_kind, "first_dir": r.first_dir,
                "first_index": r.first_index, "message": r.message.decode()}

    @property
    def shape(self):
        i = self.info()
        return (i.n[2], i.n[1], i.n[0])  # (nz, ny, nx) of the x-fastest arrays

    def indicator(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.int8)
        _check(lib().mts_setup_indicator(self._h, _dptr(out), out.size))
        return out

    def coords(self, axis: int) -> np.ndarray:
        n = self.info().n[axis]
        out = np.empty(n, dtype=np.float64)
        _check(lib().mts_setup_coords(self._h, axis, _dptr(out), n))
        return out

    def diagnostics(self, t: float):
        """((FP Lmax, FP L2), (jump Lmax, jump L2)) at time t (Error_Fp, Error_Jump)."""
        fp = np.zeros(2)
        jp = np.zeros(2)
        _check(lib().mts_setup_diagnostics(self.handle, t, _dptr(fp), _dptr(jp)))
        return (float(fp[0]), float(fp[1])), (float(jp[0]), float(jp[1]))

    def crossing_errors(self, d: int, t: float):
        """Per crossing of direction d: (errl, errr, jump error) arrays."""
        n = int(self.info().n_cross[d])
        el, er, je = np.zeros(n), np.zeros(n), np.zeros(n)
        if n:
            _check(lib().mts_setup_crossing_errors(self.handle, d, t, _dptr(el), _dptr(er),
                                                   _dptr(je), n))
        return el, er, je

    def crossings(self, d: int) -> np.ndarray:
        n = int(self.info().n_cross[d])
        out = np.zeros(max(n, 1), dtype=CROSSING_DTYPE)
        _check(lib().mts_setup_crossings(self._h, d, _dptr(out), n))
        return out[:n]


class Context:
    """A GPU (one per process here)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().mts_ctx_create(device, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().mts_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_info(self):
        name = C.create_string_buffer(256)
        hbm = C.c_size_t()
        _check(lib().mts_ctx_device_info(self._h, name, 256, C.byref(hbm)))
        return name.value.decode(), hbm.value

    def tune_lanes(self, setup: "Setup", method, dt: float, precision="fp64", steps: int = 5):
        """Time every line split that fits the grid on a scratch instance and
        keep the fastest for later instances of this grid, method and
        precision in this context: (min_lanes, gen_lanes, ms_per_step)."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        ml, gl, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().mts_tune_lanes(self._h, setup.handle, C.byref(setup.physics), m, float(dt),
                                    prec, int(steps), C.byref(ml), C.byref(gl), C.byref(ms)))
        return ml.value, gl.value, ms.value

    def tdma_batched(self, a, b, c, r) -> np.ndarray:
        """Solve nlines tridiagonal systems (arrays shaped (nlines, n))."""
        a, b, c, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, r))
        nlines, n = a.shape
        u = np.empty_like(a)
        _check(lib().mts_tdma_batched(self._h, n, nlines, _dptr(a), _dptr(b), _dptr(c),
                                      _dptr(r), _dptr(u)))
        return u


class Instance:
    """Douglas_ADI / LOD / TS solver resident on the GPU."""

    def __init__(self, ctx: Context, setup: Setup, method, dt: float, precision="fp64"):
        self._h = C.c_void_p()
        self.ctx, self.setup = ctx, setup
        self.method = METHODS[method] if isinstance(method, str) else int(method)
        self.dt = float(dt)
        self.precision = (PRECISIONS[precision] if isinstance(precision, str)
                          else int(precision))
        _check(lib().mts_instance_create_ex(ctx.handle, setup.handle, C.byref(setup.physics),
                                            self.method, self.dt, self.precision,
                                            C.byref(self._h)))
        self.shape = setup.shape

    def close(self):
        if self._h:
            lib().mts_instance_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, u: np.ndarray, layout: int = LAYOUT_XFAST):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _check(lib().mts_field_upload(self._h, _dptr(u), layout))

    def download(self, layout: int = LAYOUT_XFAST) -> np.ndarray:
        nz, ny, nx = self.shape
        out = np.empty((nz, ny, nx) if layout == LAYOUT_XFAST else (nx, ny, nz))
        _check(lib().mts_field_download(self._h, _dptr(out), layout))
        return out

    def initialize(self, t0: float = 0.0):
        _check(lib().mts_initialize(self._h, t0))

    def step(self, t_next: float):
        _check(lib().mts_step(self._h, t_next))

    def run(self, t0: float, t_finish: float, graph: bool = False):
        """The reference's time loop; graph=True captures it as one HIP graph
        (kept and replayed for the same t0 and t_finish)."""
        steps, t_end = C.c_int(), C.c_double()
        f = lib().mts_run_graph if graph else lib().mts_run
        _check(f(self._h, t0, t_finish, C.byref(steps), C.byref(t_end)))
        return steps.value, t_end.value

    def sync(self):
        _check(lib().mts_sync(self._h))

    def error(self, t: float):
        """(L-inf, L2, argmax (ix, iy, iz)) of u - exact at time t."""
        linf, l2 = C.c_double(), C.c_double()
        am = (C.c_int * 3)()
        _check(lib().mts_error_norms(self._h, t, C.byref(linf), C.byref(l2), am))
        return linf.value, l2.value, (am[0], am[1], am[2])

    def profile(self, on: bool = True):
        """Bracket every stage launch with HIP events (per-stage timing)."""
        _check(lib().mts_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        _check(lib().mts_profile_reset(self._h))

    def profile_read(self):
        """{stage: (launches, total_ms)} of the timed launches so far."""
        buf = (StageTime * 64)()
        cnt = C.c_int()
        _check(lib().mts_profile_read(self._h, buf, 64, C.byref(cnt)))
        return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms)
                for i in range(min(cnt.value, 64))}

    def diagnostics(self, t: float):
        """The interface diagnostics on the device: ((FP Lmax, FP L2),
        (jump Lmax, jump L2)) at time t."""
        fp = np.zeros(2)
        jp = np.zeros(2)
        _check(lib().mts_instance_diagnostics(self._h, t, _dptr(fp), _dptr(jp)))
        return (float(fp[0]), float(fp[1])), (float(jp[0]), float(jp[1]))

    def crossing_errors(self, d: int, t: float):
        """Per crossing of direction d, on the device: (errl, errr, jump error)."""
        n = int(self.setup.info().n_cross[d])
        el, er, je = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().mts_instance_crossing_errors(self._h, d, t, _dptr(el), _dptr(er),
                                                  _dptr(je), n))
        return el, er, je

    def lanes(self):
        """The line split the instance runs: (min_lanes, gen_lanes)."""
        ml, gl = C.c_int(), C.c_int()
        _check(lib().mts_instance_lanes(self._h, C.byref(ml), C.byref(gl)))
        return ml.value, gl.value

    def device_bytes(self) -> int:
        """Device memory the instance holds."""
        b = C.c_size_t()
        _check(lib().mts_instance_device_bytes(self._h, C.byref(b)))
        return b.value

    def device_field(self):
        p, pitch, st = C.c_void_p(), C.c_int(), C.c_void_p()
        _check(lib().mts_field_device(self._h, C.byref(p), C.byref(pitch), C.byref(st)))
        return p.value, pitch.value, st.value


def version() -> str:
    return lib().mts_version().decode()


def batch_run(instances, t0, t_finish, graph: bool = False, multi: bool = False):
    """Runs the time loops of several instances with their steps interleaved
    (``mts_batch_run``), each loop as one graph (``mts_batch_run_graph``), or
    in lockstep with one launch per line sweep across instances
    (``mts_batch_run_multi``); returns [(steps, t_end)] per instance."""
    n = len(instances)
    t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, np.float64), (n,)))
    tf = np.ascontiguousarray(np.broadcast_to(np.asarray(t_finish, np.float64), (n,)))
    hs = (C.c_void_p * n)(*[i._h.value for i in instances])
    steps = np.zeros(n, np.int32)
    t_end = np.zeros(n, np.float64)
    f = (lib().mts_batch_run_multi if multi else
         lib().mts_batch_run_graph if graph else lib().mts_batch_run)
    _check(f(hs, n, _dptr(t0), _dptr(tf), _dptr(steps), _dptr(t_end)))
    return [(int(steps[i]), float(t_end[i])) for i in range(n)]


def request(g: GridDesc, surf: SurfaceDesc, phys: PhysicsDesc, method, dt: float,
            t_finish: float, mib: int = 1, t0: float = 0.0, tag: int = 0,
            precision="fp64") -> Request:
    """An ``mts_request`` for :meth:`Server.submit`."""
    m = METHODS[method] if isinstance(method, str) else int(method)
    pr = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    return Request(g, surf, phys, mib, m, dt, t0, t_finish, tag, pr)


class Server:
    """The continuous-batching request loop (``mts_server_*``): solves are set
    up on host threads, admitted under an HBM cap and an in-flight limit, and
    advanced one step per round each on their own streams."""

    def __init__(self, ctx: Context, hbm_cap_bytes: int, max_inflight: int = 8,
                 setup_threads: int = 4):
        self._h = C.c_void_p()
        self.ctx = ctx
        self._keep = {}  # id -> request, kept until its result is polled (callback surfaces)
        _check(lib().mts_server_create(ctx.handle, int(hbm_cap_bytes), int(max_inflight),
                                       int(setup_threads), C.byref(self._h)))

    def submit(self, req: Request) -> int:
        rid = C.c_longlong()
        _check(lib().mts_server_submit(self._h, C.byref(req), C.byref(rid)))
        self._keep[rid.value] = req
        return rid.value

    def poll(self, capacity: int = 256):
        buf = (Result * capacity)()
        cnt = C.c_int()
        _check(lib().mts_server_poll(self._h, buf, capacity, C.byref(cnt)))
        out = [buf[i].as_dict() for i in range(cnt.value)]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def wait(self, max_outstanding: int = 0):
        _check(lib().mts_server_wait(self._h, int(max_outstanding)))

    def drain(self):
        """Waits for every submitted solve and returns all results by id."""
        self.wait(0)
        out = {}
        while True:
            got = self.poll()
            if not got:
                return out
            for r in got:
                out[r["id"]] = r

    def close(self):
        if self._h:
            lib().mts_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
