"""Every global-memory and LDS index the per-step kernels form, and the values
they produce, checked on the host: mts_cpucheck runs the kernels' own
per-node / per-lane functions (device_core.hpp, compiled for the host under
AddressSanitizer) over buffers of exactly the device sizes for one step of
each method, over grids whose lines split across 16-, 32- and 64-lane
groups, and can dump each stage's field for a comparison with the oracle."""
import os
import subprocess
import time

import numpy as np
import pytest

from conftest import PKG_DIR

TOOL = os.path.join(PKG_DIR, "_build", "mts_cpucheck")
E1 = repr(1 - 2 ** -20)
R0 = repr(0.5 * (1 - 2 ** -20))
GEOMS = {
    "sphere": ["17", "ellipsoid", "1", "1", "1", R0, "const", "10", "1"],
    "cfg1like": ["65", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    "torus": ["49", "torus", "0.5", "0.2", "const", "0.01", "1"],
    "bumpy": ["41", "bumpy", R0, "0.1", "5", "3", "const", "1000", "1"],
    # anisotropic grids: line splits over 16-, 32- and 64-lane groups
    "wide_x": ["600x20x18", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    "wide_y": ["18x300x22", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    "wide_z": ["20x18x1024", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    # 1024-node x and y lines: 64-lane groups, the two LDS tiles of the
    # explicit stages above 64 KiB
    "wide_x1024": ["1024x18x20", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    "wide_y1024": ["18x1024x20", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    # full splits along x: idle lanes past the last segment; a short last segment
    "full_idle": ["54x20x22", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    "full_short": ["255x18x20", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
    # x pitch equal to nx (a multiple of 16): no padding after the x rows
    "pitch16": ["32x33x34", "ellipsoid", "0.7", "0.5", "0.4", E1, "sincos", "10", "1"],
}


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("mib", ["1", "2"])
@pytest.mark.parametrize("method", ["adi", "lod-ie", "lod-cn", "ts"])
def test_kernel_indices_in_bounds(method, mib, geom):
    assert os.path.exists(TOOL), "build first (make -C matched-time-splitting-3d_amd)"
    args = GEOMS[geom]
    t0 = time.time()
    p = subprocess.run([TOOL, method, mib] + args, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    # the tile lists are checked first (one line per direction and element
    # size: coverage, classes and the grouping by 128-byte line), then the
    # kernels
    lines = p.stdout.splitlines()
    assert sum(l.startswith("dir ") and " tiles (" in l for l in lines) == 6, p.stdout
    assert lines[-1].startswith("ok"), p.stdout
    assert time.time() - t0 < 300


VALUE_CASES = [
    # the two Douglas-ADI geometries with corner pairs whose GPU runs faulted
    ("adi", "2", "torus"), ("adi", "1", "cfg1like"),
    ("adi", "1", "torus"), ("adi", "2", "cfg1like"), ("adi", "2", "sphere"),
    ("lod-ie", "1", "cfg1like"), ("lod-cn", "2", "torus"), ("ts", "2", "bumpy"),
    ("lod-cn", "2", "wide_y"), ("ts", "1", "full_short"),
]


def _geom_objects(mts, geom):
    args = GEOMS[geom]
    nn = [int(v) for v in args[0].split("x")] if "x" in args[0] else [int(args[0])] * 3
    kind = args[1]
    nsurf = {"ellipsoid": 4, "torus": 2, "bumpy": 4}[kind]
    sp = [float(v) for v in args[2:2 + nsurf]]
    b = args[2 + nsurf:]
    phys = mts.physics(b[0], float(b[1]), float(b[2]))
    return nn, mts.surface(kind, *sp), phys, (float(b[1]), float(b[2]))


@pytest.mark.parametrize("method,mib,geom", VALUE_CASES)
def test_kernel_values_match_oracle(mts, oracle, tmp_path, method, mib, geom):
    """Values, not only indices: one step emulated on the host with the
    kernels' own row, block and write-back functions (a host Thomas solve in
    place of the lane-parallel one) equals the oracle's step stage by stage,
    to the per-step parity bound (1e-12, or the measured spread of two stable
    line solvers where the systems are ill-conditioned)."""
    from conftest import rel_err, step_floor
    p = subprocess.run([TOOL, method, mib] + GEOMS[geom] + ["--dump", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    nn, surf, phys, bp = _geom_objects(mts, geom)
    st = mts.Setup(mts.grid(nn), surf, phys, mib=int(mib))
    op = oracle.OracleProblem.from_setup(st, phys.beta_kind, bp, -1.0, 1.0)
    shape = (nn[2], nn[1], nn[0])

    def ld(name):
        return np.fromfile(os.path.join(tmp_path, name), dtype=np.float64).reshape(shape)

    u0 = op.init(0.0)
    np.testing.assert_array_equal(ld("u0.bin"), u0)
    dt = 2.0 / (nn[0] - 1)
    tol = step_floor(oracle, op, method, dt, 0.0, u0)
    if method == "adi":
        r1, us, us2, u1 = op.adi_stages(dt, 0.0, u0)
        assert rel_err(ld("rhs1.bin"), r1) <= 1e-14
        assert rel_err(ld("ustar.bin"), us) <= tol
        assert rel_err(ld("ustar2.bin"), us2) <= tol
    else:
        u1, _ = op.run(method, dt, 0.0, 1, u0)
    assert rel_err(ld("u1.bin"), u1) <= tol
