"""stl2sdf on MI355X: host-side mirror of the reference Julia module
``stl2sdf`` (src/stl2sdf.jl) whose compute runs in libsdf_mi355x.so."""
from .MeshGrid import Grid, TriangularMesh, generateGridPoints, getMesh_AABB  # noqa: F401
from .MeshGrid import noninteractive_sdf_grid_setup  # noqa: F401
from .SignedDistances import evalDistancesOnTriMesh, raycast_sign_detection  # noqa: F401
