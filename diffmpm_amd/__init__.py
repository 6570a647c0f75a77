"""diffmpm_amd — MI355X-native explicit MPM time step with diffmpm's API.

The classes mirror geoelements/diffmpm (Nodes, Particles, Linear1D,
Quadrilateral4Node, LinearElastic, SimpleMaterial, Constraint, NodalForce,
ParticleTraction, Unit, Linear, Mesh1D, Mesh2D, USF, USL, MPMExplicit).  State
lives in device SoA blocks; every step method runs hand-written HIP kernels for
gfx950 through the C ABI declared in include/diffmpm_amd.h.
"""
from .constraint import Constraint
from .element import Linear1D, Quadrilateral4Node, _Element
from .forces import NodalForce, ParticleTraction
from .functions import Linear, Unit
from .material import LinearElastic, SimpleMaterial
from .mesh import Mesh1D, Mesh2D
from .node import Nodes
from .particle import Particles
from .scheme import USF, USL
from .solver import MPMExplicit

__version__ = "0.1.0"

__all__ = ["Constraint", "Linear1D", "Quadrilateral4Node", "NodalForce", "ParticleTraction",
           "Linear", "Unit", "LinearElastic", "SimpleMaterial", "Mesh1D", "Mesh2D", "Nodes",
           "Particles", "USF", "USL", "MPMExplicit", "__version__"]
