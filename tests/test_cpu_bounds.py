"""Every global-memory and LDS index the per-step kernels form, checked on the
host: mts_cpucheck runs the kernels' own per-node / per-lane functions
(device_core.hpp, compiled for the host under AddressSanitizer) over buffers
of exactly the device sizes for one step of each method, over grids whose
lines split across 16-, 32- and 64-lane groups."""
import os
import subprocess
import time

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
    assert p.stdout.startswith("ok"), p.stdout
    assert time.time() - t0 < 300
