"""stl2sdf on MI355X: host-side mirror of the reference Julia module
``stl2sdf`` (src/stl2sdf.jl) whose compute runs in libsdf_mi355x.so."""
from .MeshGrid import Grid, TriangularMesh, generateGridPoints, getMesh_AABB  # noqa: F401
from .MeshGrid import noninteractive_sdf_grid_setup  # noqa: F401
from .SignedDistances import analyze_sdf_components, evalDistancesOnTriMesh  # noqa: F401
from .SignedDistances import raycast_sign_detection, remove_sdf_artifacts_  # noqa: F401
from .SdfSmoothing import RBFs_smoothing  # noqa: F401
from .stl_to_sdf import Options, stl_to_sdf  # noqa: F401
