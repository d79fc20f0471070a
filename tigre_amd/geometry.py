"""Cone-beam geometry object with TIGRE's field names and checks.

Behaviour follows Python/tigre/utilities/geometry.py:
  - ``check_geo`` (:24-111): angles become an (n, 3) array of ZYZ Euler
    angles; mandatory fields are checked; per-view fields are expanded to one
    value (or row) per projection; defaults offOrigin 0, offDetector 0,
    rotDetector 0, COR 0, accuracy 0.5, mode 'cone';
  - ``cast_to_single`` (:121-140): every non-integer value except nVoxel,
    nDetector, n_proj, mode and filter is rounded to float32;
  - ``geometry()`` / ``ConeGeometryDefault`` (geometry_default.py:7-62).
Python arrays are ordered (z, y, x) for the volume and (v, u) for the detector.
"""
from __future__ import annotations

import copy

import numpy as np

_PER_VIEW_SCALARS = ("DSD", "DSO", "COR")
_NO_CAST = ("nVoxel", "nDetector", "n_proj", "mode", "filter")


class Geometry(object):
    def __init__(self):
        self.mode = None
        self.n_proj = None
        self.angles = None
        self.filter = None
        self.nVoxel = np.zeros(3)
        self.dVoxel = np.zeros(3)
        self.sVoxel = np.zeros(3)
        self.nDetector = np.zeros(2)
        self.dDetector = np.zeros(2)
        self.sDetector = np.zeros(2)

    # -- validation and expansion ------------------------------------------------
    def check_geo(self, angles, verbose=False):
        angles = np.atleast_1d(np.asarray(angles))
        if angles.ndim == 1:
            n = angles.shape[0]
            self.angles = np.stack([angles.astype(np.float64, copy=False), np.zeros(n), np.zeros(n)], axis=1)
            self.angles = self.angles.astype(np.result_type(angles.dtype, np.float32))
        elif angles.ndim == 2:
            if angles.shape[1] != 3:
                raise BufferError("Expected angles of dimensions (n, 3), got: " + str(angles.shape))
            self.angles = angles.copy()
        else:
            raise BufferError("Unexpected angles shape: " + str(angles.shape))
        self.n_proj = self.angles.shape[0]
        if self.mode is None:
            self.mode = "cone"

        missing = [a for a in ("nVoxel", "sVoxel", "dVoxel", "nDetector", "sDetector", "dDetector", "DSO", "DSD")
                   if not hasattr(self, a)]
        if missing:
            raise AttributeError("following mandatory fields missing from geometry:" + str(missing))

        for name, shape in (("nVoxel", (3,)), ("sVoxel", (3,)), ("dVoxel", (3,))):
            if np.asarray(getattr(self, name)).shape != shape:
                raise AttributeError("geo.%s.shape should be %s" % (name, shape))
        if not np.sum(np.abs(np.asarray(self.dVoxel) * np.asarray(self.nVoxel) - np.asarray(self.sVoxel))) < 1e-6:
            raise AttributeError("nVoxel*dVoxel is not equal to sVoxel. Check fields.")
        for name, shape in (("nDetector", (2,)), ("sDetector", (2,)), ("dDetector", (2,))):
            if np.asarray(getattr(self, name)).shape != shape:
                raise AttributeError("geo.%s.shape should be %s" % (name, shape))
        if not np.sum(np.abs(np.asarray(self.dDetector) * np.asarray(self.nDetector)
                             - np.asarray(self.sDetector))) < 1e-6:
            raise AttributeError("nDetector*dDetector is not equal to sDetector. Check fields.")

        n = self.n_proj
        for attr in ("DSD", "DSO"):
            self._expand(attr, n)
        defaults = {"offOrigin": np.array([0, 0, 0]), "offDetector": np.array([0, 0]),
                    "rotDetector": np.array([0, 0, 0]), "COR": np.zeros(n)}
        for attr, default in defaults.items():
            if not hasattr(self, attr):
                setattr(self, attr, default)
            self._expand(attr, n)
        if not hasattr(self, "accuracy"):
            self.accuracy = 0.5
        if verbose:
            print(self)

    def _expand(self, attr, n):
        """One value (scalar fields) or one row (vector fields) per projection,
        geometry.py:142-195."""
        val = getattr(self, attr)
        if isinstance(val, (float, int, np.floating, np.integer)):
            setattr(self, attr, np.tile(val, n))
            return
        if not isinstance(val, np.ndarray):
            raise TypeError("Data type not understood for: geo." + attr + " with type = " + str(type(val)))
        if val.ndim == 1:
            if val.shape in ((3,), (2,)) and attr not in _PER_VIEW_SCALARS:
                setattr(self, attr, np.tile(val, (n, 1)))
            elif val.shape == (1,):
                setattr(self, attr, np.tile(val, n) if attr in _PER_VIEW_SCALARS else np.tile(val, (n, 1)))
            elif val.shape == (n,):
                pass
            return
        if val.shape != (n, val.shape[1]):
            raise AttributeError(
                attr + " with shape: " + str(val.shape) + " not compatible with shapes: "
                + str([(n,), (n, val.shape[1]), (3,), (2,), (1,)]))

    def cast_to_single(self):
        """Round every non-integer value to float32 (integers stay integers)."""
        for attr, val in list(self.__dict__.items()):
            if val is None or attr in _NO_CAST:
                continue
            try:
                arr = np.array(val)
            except ValueError:
                continue
            if np.issubdtype(arr.dtype, np.integer) or arr.dtype.kind not in "fiub":
                continue
            setattr(self, attr, np.float32(val))

    def checknans(self):
        for attr, val in self.__dict__.items():
            if isinstance(val, np.ndarray) and val.dtype.kind == "f" and np.isnan(val).all():
                raise ValueError("Nan found in Geometry attribute:" + attr)

    def __str__(self):
        lines = ["TIGRE parameters", "-----", "Geometry parameters",
                 "Distance from source to detector (DSD) = %s mm" % (self.DSD,),
                 "Distance from source to origin (DSO)= %s mm" % (self.DSO,), "-----", "Detector parameters",
                 "Number of pixels (nDetector) = %s" % (self.nDetector,),
                 "Size of each pixel (dDetector) = %s mm" % (self.dDetector,),
                 "Total size of the detector (sDetector) = %s mm" % (self.sDetector,), "-----",
                 "Image parameters", "Number of voxels (nVoxel) = %s" % (self.nVoxel,),
                 "Total size of the image (sVoxel) = %s mm" % (self.sVoxel,),
                 "Size of each voxel (dVoxel) = %s mm" % (self.dVoxel,)]
        if hasattr(self, "accuracy"):
            lines += ["-----", "Samples per pixel of forward projection (accuracy) = %s" % (self.accuracy,)]
        return "\n".join(lines)

    def copy(self):
        return copy.deepcopy(self)


class ConeGeometryDefault(Geometry):
    """TIGRE's preset cone geometries (geometry_default.py:7-62)."""

    def __init__(self, high_resolution=True, nVoxel=None):
        Geometry.__init__(self)
        self.DSD = 1536.0
        self.DSO = 1000.0
        if nVoxel is not None:
            nVoxel = np.asarray(nVoxel)
            self.nDetector = np.array((nVoxel[1], nVoxel[2]))
            self.dDetector = np.array([0.8, 0.8])
            self.nVoxel = nVoxel
        elif high_resolution:
            self.nDetector = np.array((512, 512))
            self.dDetector = np.array((0.8, 0.8))
            self.nVoxel = np.array((512, 512, 512))
        else:
            self.nDetector = np.array((128, 128))
            self.dDetector = np.array((0.8, 0.8)) * 4
            self.nVoxel = np.array((64, 64, 64))
        self.sDetector = self.nDetector * self.dDetector
        self.sVoxel = np.array((256, 256, 256))
        self.dVoxel = self.sVoxel / self.nVoxel
        self.offOrigin = np.array((0, 0, 0))
        self.offDetector = np.array((0, 0))
        self.rotDetector = np.array((0, 0, 0))
        self.accuracy = 0.5
        self.mode = "cone"
        self.filter = None


class ParallelGeo(Geometry):
    """Parallel-beam geometry of geometry.py:274-300: 1 mm voxels, detector pixels
    of 1 mm, nDetector = (nVoxel[0], max(nVoxel[1], nVoxel[2])), DSO = nVoxel[0],
    DSD = 2 nVoxel[0]."""

    def __init__(self, nVoxel):  # noqa: N803
        if nVoxel is None:
            raise ValueError("nVoxel needs to be given for initialisation of parallel beam")
        Geometry.__init__(self)
        self.mode = "parallel"
        self.nVoxel = np.asarray(nVoxel)
        self.dVoxel = np.array([1, 1, 1])
        self.sVoxel = self.nVoxel
        self.DSO = np.float32(self.nVoxel[0])
        self.DSD = np.float32(self.nVoxel[0] * 2)
        self.dDetector = np.array([1, 1])
        self.nDetector = np.zeros((2,), dtype=np.int64)
        self.nDetector[0] = self.nVoxel[0]
        self.nDetector[1] = max(self.nVoxel[1], self.nVoxel[2])
        self.sDetector = self.nDetector
        self.accuracy = 0.5
        self.offOrigin = np.array([0, 0, 0])
        self.offDetector = np.array([0, 0])
        self.rotDetector = np.array([0, 0, 0])


def geometry(mode="cone", nVoxel=None, default=False, high_resolution=True):
    """Geometry factory with the signature of tigre.geometry (geometry.py:302-352)."""
    if mode == "cone":
        if default:
            return ConeGeometryDefault(high_resolution=high_resolution, nVoxel=nVoxel)
        return Geometry()
    if mode == "parallel":
        return ParallelGeo(nVoxel)
    raise NotImplementedError("tigre_amd implements cone and parallel beams; mode=%r is not available" % (mode,))


def cone_geometry(n_voxel, s_voxel, n_detector, d_detector, DSO, DSD, accuracy=0.5):
    """Convenience constructor: sizes in (z, y, x) / (v, u) order, mm units."""
    g = Geometry()
    g.mode = "cone"
    g.nVoxel = np.asarray(n_voxel, dtype=np.int64)
    g.sVoxel = np.asarray(s_voxel, dtype=np.float64)
    g.dVoxel = g.sVoxel / g.nVoxel
    g.nDetector = np.asarray(n_detector, dtype=np.int64)
    g.dDetector = np.asarray(d_detector, dtype=np.float64)
    g.sDetector = g.nDetector * g.dDetector
    g.DSO = float(DSO)
    g.DSD = float(DSD)
    g.offOrigin = np.array([0.0, 0.0, 0.0])
    g.offDetector = np.array([0.0, 0.0])
    g.rotDetector = np.array([0.0, 0.0, 0.0])
    g.accuracy = accuracy
    return g


def prepared(geo, angles):
    """Deep copy, check_geo and cast_to_single, as Ax.py:42-49 / Atb.py:15-21 do."""
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    g.cast_to_single()
    return g
