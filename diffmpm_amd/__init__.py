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
           "Particles", "USF", "USL", "MPMExplicit", "MPM", "__version__"]


class MPM:
    """__init__.py:13-47 of the reference: a simulation from a TOML file
    (``io.Config``), solved with ``MPMExplicit.solve_jit``.  Output goes to
    ``output.folder/meta.title`` as NPZ files when ``output.format`` is
    "npz".  ``meta.scheme`` ("usf" or "usl") selects the scheme; the
    reference reads it but does not pass it on, so it always runs USF (every
    input file it ships says "usf", so both agree on those)."""

    def __init__(self, filepath, *, device=None, dtype=None):
        from pathlib import Path

        from . import writers
        from .io import Config
        self._config = Config(filepath, device=device, dtype=dtype)
        mesh = self._config.parse()
        pc = self._config.parsed_config
        out_dir = Path(pc["output"]["folder"]).joinpath(pc["meta"]["title"])
        fmt = pc["output"].get("format", None)
        if fmt is None or fmt.lower() == "none":
            writer_func = None
        elif fmt == "npz":
            writer_func = writers.NPZWriter().write
        else:
            raise ValueError(f"Specified output format not supported: {fmt}")
        if pc["meta"]["type"] != "MPMExplicit":
            raise ValueError("Wrong type of solver specified.")
        self.solver = MPMExplicit(mesh, pc["meta"]["dt"], pc["meta"].get("scheme", "usf"),
                                  velocity_update=pc["meta"]["velocity_update"],
                                  sim_steps=pc["meta"]["nsteps"],
                                  out_steps=pc["output"]["step_frequency"], out_dir=out_dir,
                                  writer_func=writer_func)

    def solve(self):
        """Runs the simulation (solve_jit with the file's gravity) and
        returns the final particle arrays."""
        return self.solver.solve_jit(self._config.parsed_config["external_loading"]["gravity"])
