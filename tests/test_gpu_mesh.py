"""The device mesh -> SDF build (sdr_sdf_from_mesh) against the oracle's
`Sdf::from_mesh` on the same meshes: the distances bit for bit (the same f64 arithmetic
in the same order), the signs equal on every node off the surface (the device atan2
may move the winding sum in its last bits, which can only flip a node on the surface).
The box's faces lie on grid planes, which puts nodes at distance exactly 0."""
import numpy as np
import pytest

from sdr_amd import scenes
from sdr_amd.solver import Sdf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", ["box", "icosphere"])
def test_mesh_sdf_matches_oracle(dev, orc, mesh):
    if mesh == "box":
        v, f = scenes.box_mesh(0.7)
        dx, pad = 0.05, 0.4
    else:
        v, f = scenes.icosphere(1.0, 4, center=(0.013, -0.021, 0.007))
        dx, pad = 0.06, 0.3
    a = Sdf.from_mesh(v, f, dx, padding=pad, lib=orc)
    b = Sdf.from_mesh(v, f, dx, padding=pad, lib=dev)
    assert (a.nx, a.ny, a.nz, a.origin, a.dx) == (b.nx, b.ny, b.nz, b.origin, b.dx)
    assert np.array_equal(np.abs(a.data).view(np.uint64), np.abs(b.data).view(np.uint64))
    # The box's faces fall on grid planes, so about 5% of its nodes lie on the surface
    # (distance exactly 0) and carry no sign to compare.
    off = np.abs(a.data) > 1e-9
    assert off.sum() > 0.9 * off.size
    assert np.array_equal(np.sign(a.data[off]), np.sign(b.data[off]))
    assert (b.data < 0).sum() > 1000
