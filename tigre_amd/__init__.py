"""tigre_amd: TIGRE's cone-beam projector pair and TV step on MI355X (gfx950).

Public names mirror the reference package (Python/tigre/__init__.py):
``Ax``, ``Atb``, ``geometry``, ``geometry_default``, ``minTV``,
``im3ddenoise`` and the algorithms in ``tigre_amd.algorithms``.
"""
from .geometry import ConeGeometryDefault, Geometry, cone_geometry, geometry  # noqa: F401
from .geometry import ConeGeometryDefault as geometry_default  # noqa: F401
from .operators import Atb, Ax  # noqa: F401
from .tv import AwminTV, im3ddenoise, minimizeTV, minPICCS, minTV, tvdenoise  # noqa: F401

__version__ = "0.1.0"
