"""Setup refusals: the cases where the reference exits (or reads out of
bounds) become error statuses with the reason, and the process survives."""
import ctypes

import numpy as np
import pytest


def test_interface_on_grid_node(mts):
    # sphere of radius 0.5 through the node (0.5, 0, 0) of a 33^3 grid
    # (Mesh::Mesh exits "Interface on grid", mesh/mesh.cc:64-66)
    with pytest.raises(mts.MtsError, match="interface on grid node"):
        mts.Setup(mts.grid(33), mts.surface("ellipsoid", 1, 1, 1, 0.5), mts.physics("const", 10, 1))


def test_open_surface(mts):
    # a sphere sticking out of the box crosses the first cell of lines
    with pytest.raises(mts.MtsError) as e:
        mts.Setup(mts.grid(17), mts.surface("ellipsoid", 1, 1, 1, 0.97), mts.physics("const", 10, 1))
    assert e.value.code == -2


def test_bad_arguments(mts):
    with pytest.raises(mts.MtsError):
        mts.Setup(mts.grid(4), mts.surface("ellipsoid", 1, 1, 1, 0.3), mts.physics("const", 1, 1))
    with pytest.raises(mts.MtsError, match="MIB method"):
        mts.Setup(mts.grid(17), mts.surface("ellipsoid", 1, 1, 1, 0.3), mts.physics("const", 1, 1),
                  mib=3)
    with pytest.raises(mts.MtsError, match="2nd-order"):
        mts.Setup(mts.grid(17), mts.surface("ellipsoid", 1, 1, 1, 0.3), mts.physics("const", 1, 1),
                  accuracy=4)


def test_process_survives_and_next_setup_works(mts):
    for _ in range(3):
        with pytest.raises(mts.MtsError):
            mts.Setup(mts.grid(33), mts.surface("ellipsoid", 1, 1, 1, 0.5),
                      mts.physics("const", 10, 1))
    st = mts.Setup(mts.grid(33), mts.surface("ellipsoid", 1, 1, 1, 0.47), mts.physics("const", 10, 1))
    assert st.info().n_cross[0] > 0


def test_callback_surface_matches_builtin(mts):
    """A user level set through the C callback gives the same crossings as the
    built-in ellipsoid when it evaluates the same expression."""
    def phi(_, x, y, z):
        return x * x / (0.7 * 0.7) + y * y / (0.5 * 0.5) + z * z / (0.4 * 0.4) - 0.9 * 0.9

    def grad(_, x, y, z, g):
        g[0] = 2 * x / (0.7 * 0.7)
        g[1] = 2 * y / (0.5 * 0.5)
        g[2] = 2 * z / (0.4 * 0.4)

    s = mts.SurfaceDesc()
    s.kind = mts.SURF_CALLBACK
    s.phi = mts.PHI_FN(phi)
    s.grad = mts.GRAD_FN(grad)
    cb = mts.Setup(mts.grid(21), s, mts.physics("const", 10, 1), mib=2)
    ref = mts.Setup(mts.grid(21), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 0.9),
                    mts.physics("const", 10, 1), mib=2)
    np.testing.assert_array_equal(cb.indicator(), ref.indicator())
    for d in range(3):
        a, b = cb.crossings(d), ref.crossings(d)
        np.testing.assert_array_equal(a["left_loc"], b["left_loc"])
        np.testing.assert_allclose(a["coord"], b["coord"], rtol=0, atol=1e-14)
        np.testing.assert_allclose(a["wl"], b["wl"], rtol=1e-10, atol=1e-12)
