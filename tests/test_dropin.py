"""The drop-in layer on the host: the LU class (decomposition/lu.cc), Mesh
classification (mesh/mesh.cc:20-70), the C++ classes' error modes and the
config.txt driver's parsing (configuration/read_config.cc), with no GPU.

LU is pinned by brute force (numpy's LAPACK solve, inverse and determinant
on random well-conditioned systems), and the classification by the setup's
own indicator, which the setup golden tests pin against the reference."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR

BUILD = os.path.join(PKG_DIR, "_build")


def _lu(mts, a):
    h = C.c_void_p()
    L = mts.lib()
    L.mts_lu_create.argtypes = [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]
    rc = L.mts_lu_create(a.shape[0], np.ascontiguousarray(a).ctypes.data, C.byref(h))
    return rc, h


def _sig(mts):
    L = mts.lib()
    VP = C.c_void_p
    L.mts_lu_solve.argtypes = [VP, VP, VP]
    L.mts_lu_solve_matrix.argtypes = [VP, C.c_int, VP, VP]
    L.mts_lu_inverse.argtypes = [VP, VP]
    L.mts_lu_det.argtypes = [VP, C.POINTER(C.c_double)]
    L.mts_lu_mprove.argtypes = [VP, VP, VP]
    L.mts_lu_destroy.argtypes = [VP]
    L.mts_lu_destroy.restype = None
    L.mts_mesh_classify.argtypes = [VP, VP, VP, C.c_size_t]
    return L


@pytest.mark.parametrize("n", [1, 2, 6, 12, 36])
def test_lu_against_dense_solve(mts, n):
    L = _sig(mts)
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + n * np.eye(n) * rng.choice([-1, 1], n)
    rng.shuffle(a)  # rows out of order: pivoting must find them
    rc, h = _lu(mts, a)
    assert rc == 0
    b = rng.standard_normal(n)
    x = np.zeros(n)
    assert L.mts_lu_solve(h, b.ctypes.data, x.ctypes.data) == 0
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-12, atol=1e-13)
    bm = rng.standard_normal((n, 3))
    xm = np.zeros((n, 3))
    assert L.mts_lu_solve_matrix(h, 3, bm.ctypes.data, xm.ctypes.data) == 0
    np.testing.assert_allclose(xm, np.linalg.solve(a, bm), rtol=1e-12, atol=1e-13)
    inv = np.zeros((n, n))
    assert L.mts_lu_inverse(h, inv.ctypes.data) == 0
    np.testing.assert_allclose(inv, np.linalg.inv(a), rtol=1e-11, atol=1e-13)
    d = C.c_double()
    assert L.mts_lu_det(h, C.byref(d)) == 0
    assert d.value == pytest.approx(np.linalg.det(a), rel=1e-11)
    # mprove: one refinement step removes a perturbation of the solution
    xp = x + 1e-7 * rng.standard_normal(n)
    assert L.mts_lu_mprove(h, b.ctypes.data, xp.ctypes.data) == 0
    assert np.max(np.abs(xp - np.linalg.solve(a, b))) < 1e-12
    L.mts_lu_destroy(h)


def test_lu_zero_row_is_refused(mts):
    _sig(mts)
    rc, h = _lu(mts, np.array([[1.0, 2.0], [0.0, 0.0]]))
    assert rc == -1 and b"Singular" in mts.lib().mts_last_error()
    assert not h.value


def test_lu_singular_matrix_follows_the_textbook_pivot_quirk(mts):
    """A singular matrix without a zero row decomposes: lu.cc:22-67 keeps the
    pivot row of the previous column (imax) when a column below the diagonal
    is all zero, and swaps it back in.  By hand for [[1, 2], [2, 4]]: k = 0
    keeps row 0 (scaled tie), leaving [[1, 2], [2, 0]]; k = 1 finds no pivot,
    reuses imax = 0 and swaps, giving [[2, 0], [1, 2]] with one interchange,
    so det = -(2 * 2) = -4.  A triangular singular matrix takes the 1e-40
    pivot instead (lu.cc:60-62)."""
    L = _sig(mts)
    rc, h = _lu(mts, np.array([[1.0, 2.0], [2.0, 4.0]]))
    assert rc == 0
    d = C.c_double()
    L.mts_lu_det(h, C.byref(d))
    assert d.value == -4.0
    L.mts_lu_destroy(h)
    rc, h = _lu(mts, np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]))
    assert rc == 0
    L.mts_lu_det(h, C.byref(d))
    assert abs(d.value) <= 1e-40
    L.mts_lu_destroy(h)


@pytest.mark.parametrize("surf,args,n", [("ellipsoid", (0.7, 0.5, 0.4, 1 - 2 ** -20), 33),
                                         ("torus", (0.5, 0.2), 41),
                                         ("bumpy", (0.5 * (1 - 2 ** -20), 0.1, 5, 3), 29)])
def test_mesh_classify_matches_setup(mts, surf, args, n):
    L = _sig(mts)
    g, s = mts.grid(n), mts.surface(surf, *args)
    out = np.zeros(n ** 3, np.int8)
    assert L.mts_mesh_classify(C.byref(g), C.byref(s), out.ctypes.data, out.size) == 0
    st = mts.Setup(g, s, mts.physics("const", 10, 1), mib=1)
    assert np.array_equal(out, st.indicator().ravel())


def test_mesh_classify_interface_on_node(mts):
    L = _sig(mts)
    g, s = mts.grid(33), mts.surface("ellipsoid", 1, 1, 1, 0.5)
    out = np.zeros(33 ** 3, np.int8)
    assert L.mts_mesh_classify(C.byref(g), C.byref(s), out.ctypes.data, out.size) == -2
    assert b"grid node" in mts.lib().mts_last_error()
    assert L.mts_mesh_classify(C.byref(g), C.byref(s), out.ctypes.data, 10) == -1


def test_cpp_classes_on_the_host():
    exe = os.path.join(BUILD, "dropin_check")
    p = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    checks = [json.loads(l) for l in p.stdout.splitlines()]
    assert len(checks) == 8
    bad = [c for c in checks if not c["ok"]]
    assert not bad, bad


CFG = """# cube [-1,1]^3, 33^3 nodes
x-max = 1
x-min = -1
y-max = 1
y-min = -1
z-max = 1
z-min = -1
x-grid = 33
y-grid = 33
z-grid = 33
time-start = 0
time-finish = 1
time-step = 0.0625
surface = ellipsoid(1, 1, 1, 0.4999995231628418)
temporal-method = adi
equation = exp-sin-cos
beta = const(10, 1)
spatial-method = mib-v1
spatial-accuracy = 2
"""


def _driver(tmp_path, text, *extra):
    cfg = tmp_path / "config.txt"
    cfg.write_text(text)
    out = tmp_path / "out.txt"
    p = subprocess.run([os.path.join(BUILD, "mts_driver"), "-i", str(cfg), "-o", str(out), *extra],
                       capture_output=True, text=True, timeout=120)
    return p, (out.read_text() if out.exists() else "")


def test_driver_dry_run_writes_configuration(tmp_path):
    p, out = _driver(tmp_path, CFG, "--dry-run")
    assert p.returncode == 0, p.stderr
    assert "Mesh Size                          :   NX = 33; NY = 33; NZ = 33" in out
    assert "Time Step                          :   6.2e-02" in out
    assert "Temporal Method                    :   ADI" in out
    assert "Spatial Method                     :   MIB-V1" in out
    assert "crossings x" in p.stdout


def test_driver_reads_domain_bounds_as_integers(tmp_path):
    """The reference reads x-max etc. with stoi (configuration/read_config.cc:39-56), so
    a fractional bound is truncated: 1.9 runs the box of 1."""
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    p1, out1 = _driver(tmp_path / "a", CFG, "--dry-run")
    p2, out2 = _driver(tmp_path / "b", CFG.replace("x-max = 1\n", "x-max = 1.9\n"), "--dry-run")
    assert p1.returncode == 0 and p2.returncode == 0, p2.stderr
    assert out1 == out2

    def crossings(stdout):
        return [l for l in stdout.splitlines() if "crossings" in l]
    assert crossings(p1.stdout) == crossings(p2.stdout) != []


@pytest.mark.parametrize("key,val,msg", [
    ("equation", "3", "not available in this build"),
    ("beta", "3", "not available in this build"),
    ("surface", "cone", "not available in this build"),
    ("surface", "blob", "Surface type not found"),
    ("temporal-method", "rk4", "Temporal-method type not found"),
    ("spatial-method", "mib-v3", "Spatial method type not found"),
    ("spatial-accuracy", "4", "only 2nd-order"),
])
def test_driver_refuses_what_this_build_does_not_evaluate(tmp_path, key, val, msg):
    lines = [l if not l.startswith(key + " ") else f"{key} = {val}" for l in CFG.splitlines()]
    p, _ = _driver(tmp_path, "\n".join(lines) + "\n", "--dry-run")
    assert p.returncode == 1
    assert msg in p.stderr


def test_driver_interface_on_node_exits_like_the_reference(tmp_path):
    p, _ = _driver(tmp_path, CFG.replace("0.4999995231628418", "0.5"), "--dry-run")
    assert p.returncode == 1 and "[ERROR]" in p.stderr and "grid node" in p.stderr


def test_driver_needs_input_and_output():
    p = subprocess.run([os.path.join(BUILD, "mts_driver")], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 1 and "Usage" in p.stdout
