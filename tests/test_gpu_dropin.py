"""The C++ drop-in classes and the config.txt driver on the GPU.

The driver runs the 33^3 Douglas-ADI MIB-L1 case whose reference output is
held in tests/golden/cfg0_adi_mib1.npz (steps, final time, L-inf, L2,
argmax): its report must carry the same numbers at the precision it prints.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, load_golden

pytestmark = pytest.mark.gpu
BUILD = os.path.join(PKG_DIR, "_build")


def test_cpp_solver_classes():
    print("dropin_check gpu", flush=True)
    p = subprocess.run([os.path.join(BUILD, "dropin_check"), "gpu"], capture_output=True,
                       text=True, timeout=240)
    assert p.returncode == 0, p.stdout + p.stderr
    checks = {c["check"]: c for c in map(json.loads, p.stdout.splitlines())}
    assert set(checks) == {"lod_host_sync_same_field", "lod_same_as_c_abi", "ts_same_as_c_abi",
                           "adi_corners_refused"}
    bad = [c for c in checks.values() if not c["ok"]]
    assert not bad, bad


def check_write_txt(tmp_path, g):
    """The result files the reference's main writes after every run (Write_txt,
    helper.h:442-575): per crossing, z lines then y then x, each in the
    reference's loop order (the line's first fixed index outer), the node its
    ID's parity picks, its value and its error against the outer exact
    solution at the configuration's finish time; "%5d%5d%5d%20.6f" lines.
    The values are the reference's own final field (the GPU's agrees to far
    below the printed 6 decimals)."""
    meta = g["meta"]
    n1 = meta["n"] + 1
    xs = -1.0 + np.arange(n1) * (2.0 / (n1 - 1.0))
    e = np.exp(-meta["T"])
    want = []
    for d in (2, 1, 0):
        ii = g[f"cr{d}_int"]
        order = sorted(range(len(ii)), key=lambda q: (ii[q][0], ii[q][1]))
        for q in order:
            j, k, cid, ll = (int(v) for v in ii[q][:4])
            node = ll if abs(cid) % 2 == 1 else ll + 1
            ix, iy, iz = (node, j, k) if d == 0 else (j, node, k) if d == 1 else (j, k, node)
            u = float(g["u_final"][iz, iy, ix])
            ex = e * np.cos(xs[ix]) * np.cos(xs[iy]) * np.cos(xs[iz])
            want.append((ix, iy, iz, u, abs(u - ex)))
    for fname, col in (("Numerical Solutions.txt", 3), ("Numerical Error.txt", 4)):
        lines = (tmp_path / fname).read_text().splitlines()
        assert len(lines) == len(want)
        for line, w in zip(lines, want):
            assert len(line) == 35 and line[15:].startswith(" ")
            got = (int(line[0:5]), int(line[5:10]), int(line[10:15]))
            assert got == w[:3]
            assert abs(float(line[15:]) - w[col]) <= 1.01e-6


def test_driver_reproduces_the_reference_report(tmp_path):
    g = load_golden("cfg0_adi_mib1")
    meta = g["meta"]
    s = meta["surface"]
    cfg = tmp_path / "config.txt"
    cfg.write_text("\n".join([
        "x-max = 1", "x-min = -1", "y-max = 1", "y-min = -1", "z-max = 1", "z-min = -1",
        f"x-grid = {meta['n'] + 1}", f"y-grid = {meta['n'] + 1}", f"z-grid = {meta['n'] + 1}",
        "time-start = 0", f"time-finish = {meta['T']!r}", f"time-step = {meta['dt']!r}",
        f"surface = ellipsoid({s[1]!r},{s[2]!r},{s[3]!r},{s[4]!r})",
        "temporal-method = adi", "equation = exp-sin-cos",
        f"beta = const({meta['beta'][1]},{meta['beta'][2]})",
        "spatial-method = mib-v1", "spatial-accuracy = 2", ""]))
    out = tmp_path / "out.txt"
    print("mts_driver cfg0", flush=True)
    p = subprocess.run([os.path.join(BUILD, "mts_driver"), "-i", str(cfg), "-o", str(out)],
                       capture_output=True, text=True, timeout=240, cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    text = out.read_text()
    assert "Error of Douglas-ADI with MIB" in text
    check_write_txt(tmp_path, g)
    # the interface diagnostics, as the reference's report lines
    rep = [f"{v:.3e}" for v in g["diag_report"]]
    assert f"Error of FPs ------------------------\nLmax = {rep[0]}\nL2   = {rep[1]}" in text
    assert f"Error of Jumps ------------------------\nLmax = {rep[2]}\nL2   = {rep[3]}" in text
    steps, t_end, linf, l2, ax, ay, az = g["summary"][:7]
    m = re.search(r"T = (\S+)", text)
    assert float(m.group(1)) == float(f"{t_end:.1e}")
    # The sphere problem is symmetric under every pair of axis reflections
    # (sin x sin y sin z and cos x cos y cos z are both even under them): the
    # errors at a node and at its images agree to rounding, so which one is
    # the first maximum follows the last bit of each.
    m = re.search(r"\[Xmax,Ymax,Zmax\] = \[(\d+),(\d+),(\d+)\]", text)
    n1 = meta["n"]
    p = (int(ax), int(ay), int(az))
    orbit = {p}
    for flip in ((1, 1, 0), (1, 0, 1), (0, 1, 1)):
        orbit.add(tuple(n1 - v if f else v for v, f in zip(p, flip)))
    assert tuple(map(int, m.groups())) in orbit
    m = re.search(r"Lmax = (\S+);  L2 = (\S+)", text)
    assert m.group(1) == f"{linf:.3e}" and m.group(2) == f"{l2:.3e}"
    assert "CPU time cost" in text


@pytest.mark.parametrize("name", ["sphere16_lod-cn_mib2", "sphere16_lod-ie_mib1", "sphere16_ts_mib1",
                                  "torus22_lodcn_mib2"])
def test_driver_other_methods_match_the_reference(tmp_path, name):
    """The driver's LOD and TS branches on golden cases: the reported L-inf
    error and the interface diagnostics at the reference's printed precision."""
    g = load_golden(name)
    meta = g["meta"]
    s = meta["surface"]
    beta = meta["beta"]
    cfg = tmp_path / "config.txt"
    surf = (f"ellipsoid({s[1]!r},{s[2]!r},{s[3]!r},{s[4]!r})" if s[0] == "ellipsoid"
            else f"torus({s[1]!r},{s[2]!r})")
    cfg.write_text("\n".join([
        "x-max = 1", "x-min = -1", "y-max = 1", "y-min = -1", "z-max = 1", "z-min = -1",
        f"x-grid = {meta['n'] + 1}", f"y-grid = {meta['n'] + 1}", f"z-grid = {meta['n'] + 1}",
        "time-start = 0", f"time-finish = {meta['T']!r}", f"time-step = {meta['dt']!r}",
        f"surface = {surf}", f"temporal-method = {meta['method']}", "equation = exp-sin-cos",
        f"beta = {beta[0]}({beta[1]},{beta[2]})", f"spatial-method = mib-v{meta['mib']}",
        "spatial-accuracy = 2", ""]))
    out = tmp_path / "out.txt"
    print("mts_driver", name, flush=True)
    p = subprocess.run([os.path.join(BUILD, "mts_driver"), "-i", str(cfg), "-o", str(out)],
                       capture_output=True, text=True, timeout=240, cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    text = out.read_text()
    linf = float(g["summary"][2])
    m = re.search(r"Lmax = (\S+);  L2 = (\S+)", text)
    assert abs(float(m.group(1)) - linf) <= 5e-4 * abs(linf)  # printed with 3 digits
    rep = [f"{v:.3e}" for v in g["diag_report"]]
    assert f"Error of FPs ------------------------\nLmax = {rep[0]}\nL2   = {rep[1]}" in text
    assert f"Error of Jumps ------------------------\nLmax = {rep[2]}\nL2   = {rep[3]}" in text