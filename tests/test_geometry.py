"""Geometry preparation: the semantics of Python/tigre/utilities/geometry.py."""
import numpy as np
import pytest

from tigre_amd import _native
from tigre_amd.geometry import ConeGeometryDefault, Geometry, cone_geometry, geometry, prepared


def test_check_geo_expands_and_defaults():
    g = cone_geometry((8, 9, 10), (8.0, 9.0, 10.0), (6, 7), (1.0, 1.0), 100.0, 150.0)
    for a in ("offOrigin", "offDetector", "rotDetector"):
        delattr(g, a)
    g.check_geo(np.linspace(0, 1, 5))
    assert g.angles.shape == (5, 3) and np.all(g.angles[:, 1:] == 0)
    assert g.DSO.shape == (5,) and g.DSD.shape == (5,)
    assert g.offOrigin.shape == (5, 3) and np.all(g.offOrigin == 0)
    assert g.offDetector.shape == (5, 2) and g.rotDetector.shape == (5, 3)
    assert g.COR.shape == (5,) and np.all(g.COR == 0)
    assert g.mode == "cone" and g.accuracy == 0.5


def test_check_geo_rejects_inconsistent_sizes():
    g = cone_geometry((8, 9, 10), (8.0, 9.0, 10.0), (6, 7), (1.0, 1.0), 100.0, 150.0)
    g.sVoxel = np.array([8.0, 9.0, 11.0])
    with pytest.raises(AttributeError):
        g.check_geo(np.zeros(3))
    g = cone_geometry((8, 9, 10), (8.0, 9.0, 10.0), (6, 7), (1.0, 1.0), 100.0, 150.0)
    with pytest.raises(BufferError):
        g.check_geo(np.zeros((3, 2)))
    g.offOrigin = np.zeros((4, 3))
    with pytest.raises(AttributeError):
        g.check_geo(np.zeros(3))


def test_per_view_arrays_pass_through():
    g = cone_geometry((8, 9, 10), (8.0, 9.0, 10.0), (6, 7), (1.0, 1.0), 100.0, 150.0)
    g.DSO = np.array([100.0, 101.0, 102.0])
    g.offDetector = np.arange(6.0).reshape(3, 2)
    g.check_geo(np.zeros((3, 3)))
    assert np.all(g.DSO == [100, 101, 102]) and g.offDetector.shape == (3, 2)


def test_cast_to_single_keeps_integers():
    g = prepared(cone_geometry((8, 9, 10), (8.0, 9.0, 10.0), (6, 7), (1.0, 1.0), 100.0, 150.0), np.arange(4.0))
    assert g.sVoxel.dtype == np.float32 and g.DSO.dtype == np.float32 and g.angles.dtype == np.float32
    assert np.issubdtype(np.asarray(g.nVoxel).dtype, np.integer)


def test_c_geometry_axis_order():
    """Python (z, y, x) / (v, u) -> C (X, Y, Z) / (U, V); rotDetector row = (yaw, pitch, roll)
    (_types.pxd:73-144)."""
    g = cone_geometry((8, 9, 10), (8.0, 18.0, 30.0), (6, 7), (2.0, 3.0), 100.0, 150.0)
    g.offOrigin = np.array([1.0, 2.0, 3.0])
    g.offDetector = np.array([4.0, 5.0])
    g.rotDetector = np.array([0.1, 0.2, 0.3])
    gp = prepared(g, np.zeros(2))
    cg = _native.CGeometry(gp, 2)  # owns the arrays the struct points into
    c = cg.struct
    assert (c.nVoxelX, c.nVoxelY, c.nVoxelZ) == (10, 9, 8)
    assert (c.dVoxelX, c.dVoxelY, c.dVoxelZ) == (3.0, 2.0, 1.0)
    assert (c.nDetecU, c.nDetecV) == (7, 6) and (c.dDetecU, c.dDetecV) == (3.0, 2.0)
    assert (c.offOrigX[1], c.offOrigY[1], c.offOrigZ[1]) == (3.0, 2.0, 1.0)
    assert (c.offDetecU[0], c.offDetecV[0]) == (5.0, 4.0)
    assert c.dYaw[0] == pytest.approx(0.1) and c.dPitch[0] == pytest.approx(0.2) and c.dRoll[0] == pytest.approx(0.3)


def test_default_geometries():
    g = ConeGeometryDefault(high_resolution=True)
    assert list(g.nVoxel) == [512, 512, 512] and list(g.nDetector) == [512, 512] and g.DSD == 1536.0
    g = geometry(mode="cone", default=True, high_resolution=False)
    assert list(g.nVoxel) == [64, 64, 64]
    assert isinstance(geometry(mode="cone"), Geometry)
    # parallel beam needs nVoxel (geometry.py:274-278) and follows ParallelGeo's defaults
    with pytest.raises(ValueError):
        geometry(mode="parallel")
    p = geometry(mode="parallel", nVoxel=np.array([8, 10, 12]))
    assert p.mode == "parallel" and list(p.nDetector) == [8, 12] and float(p.DSD) == 16.0
    with pytest.raises(NotImplementedError):
        geometry(mode="fan")
