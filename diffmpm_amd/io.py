"""TOML/JSON input front-end (io.py of the reference).

``Config(filepath)`` reads a simulation file in the reference's format
(diffmpm/io.py:16-170): ``[meta]``, ``[output]``, ``[mesh]`` with
``[[mesh.constraints]]``, ``[[materials]]``, ``[[particles]]`` (particle
positions from a JSON file of shape (N, 1, ndim)), ``[external_loading]``
with ``[[external_loading.concentrated_nodal_forces]]`` and
``[[external_loading.particle_surface_traction]]``, and
``[[math_functions]]``; node and particle sets come from the JSON file named
by ``mesh.entity_sets``.  ``parse()`` builds the mesh of this package's GPU
classes (the same constructors the reference calls).

Relative file names (entity sets, particle files) are opened from the
current directory as in the reference, or, when absent there, from the
directory of the TOML file.
"""
import json
import os

try:
    import tomllib as _toml
except ImportError:  # Python < 3.11
    import tomli as _toml

import numpy as np
import torch

from . import element as _element
from . import material as _material
from . import mesh as _mesh
from .constraint import Constraint
from .forces import NodalForce, ParticleTraction
from .functions import Linear, Unit
from .particle import Particles

__all__ = ["Config"]


class Config:
    """io.py:16-170."""

    def __init__(self, filepath, *, device=None, dtype=None):
        self._filepath = filepath
        self._device = device
        self._dtype = dtype
        self.parsed_config = {}
        self.parse()

    def _placement(self):
        from ._lib import default_device
        kw = {"device": self._device if self._device is not None else default_device()}
        if self._dtype is not None:
            kw["dtype"] = self._dtype
        return kw

    def _path(self, name):
        if os.path.isabs(name) or os.path.exists(name):
            return name
        alt = os.path.join(os.path.dirname(os.path.abspath(self._filepath)), name)
        return alt if os.path.exists(alt) else name

    def parse(self):
        """io.py:22-36: reads the file and returns the mesh."""
        with open(self._filepath, "rb") as f:
            self._fileconfig = _toml.load(f)
        with open(self._path(self._fileconfig["mesh"]["entity_sets"])) as f:
            self.entity_sets = json.load(f)
        self._parse_meta(self._fileconfig)
        self._parse_output(self._fileconfig)
        self._parse_materials(self._fileconfig)
        self._parse_particles(self._fileconfig)
        if "math_functions" in self._fileconfig:
            self._parse_math_functions(self._fileconfig)
        self._parse_external_loading(self._fileconfig)
        return self._parse_mesh(self._fileconfig)

    def _get_node_set_ids(self, set_ids):
        """io.py:38-43: concatenated node sets."""
        ids = []
        for s in set_ids:
            ids.extend(self.entity_sets["node_sets"][str(s)])
        return torch.as_tensor(ids, dtype=torch.int64)

    def _get_particle_set_ids(self, set_type, set_nos, set_ids):
        """io.py:45-51: concatenated particle sets of particle-set numbers."""
        ids = []
        for n in set_nos:
            for s in set_ids:
                ids.extend(self.entity_sets["particle_sets"][n][str(s)])
        return torch.as_tensor(ids, dtype=torch.int64)

    def _parse_meta(self, config):
        self.parsed_config["meta"] = config["meta"]

    def _parse_output(self, config):
        self.parsed_config["output"] = config["output"]

    def _parse_materials(self, config):
        """io.py:59-67: material classes by name, constructed from the
        remaining keys."""
        materials = []
        for mc in config["materials"]:
            mc = dict(mc)
            cls = getattr(_material, mc.pop("type"))
            materials.append(cls(mc))
        self.parsed_config["materials"] = materials

    def _parse_particles(self, config):
        """io.py:69-79: positions from JSON, element ids 0, initial
        velocity broadcast to every particle."""
        sets = []
        for pc in config["particles"]:
            mat = self.parsed_config["materials"][pc["material_id"]]
            with open(self._path(pc["file"])) as f:
                loc = np.asarray(json.load(f), dtype=np.float64)
            ps = Particles(torch.from_numpy(loc), mat,
                           torch.zeros(loc.shape[0], dtype=torch.int32), **self._placement())
            v = torch.as_tensor(pc["init_velocity"], dtype=torch.float64)
            ps.velocity = torch.broadcast_to(v.to(ps.dtype).to(ps.device),
                                             (len(ps), 1, loc.shape[-1])).clone()
            sets.append(ps)
        self.parsed_config["particles"] = sets

    def _parse_math_functions(self, config):
        """io.py:81-95: Linear functions by position; other types are not
        supported (as in the reference)."""
        fns = []
        for i, fc in enumerate(config["math_functions"]):
            if fc["type"] != "Linear":
                raise NotImplementedError("Function type other than `Linear` not yet supported")
            fns.append(Linear(i, np.asarray(fc["xvalues"]), np.asarray(fc["fxvalues"])))
        self.parsed_config["math_functions"] = fns

    def _parse_external_loading(self, config):
        """io.py:97-145: gravity, concentrated nodal forces on node sets,
        particle surface tractions on particle sets (Unit function when no
        math_function_id is given)."""
        el = config["external_loading"]
        out = {"gravity": torch.as_tensor(el["gravity"], dtype=torch.float64),
               "concentrated_nodal_forces": []}
        for c in el.get("concentrated_nodal_forces", []):
            fn = (self.parsed_config["math_functions"][c["math_function_id"]]
                  if "math_function_id" in c else Unit(-1))
            out["concentrated_nodal_forces"].append(
                NodalForce(node_ids=self._get_node_set_ids(c["nset_ids"]), function=fn,
                           dir=c["dir"], force=c["force"]))
        tractions = []
        for c in el.get("particle_surface_traction", []):
            fn = (self.parsed_config["math_functions"][c["math_function_id"]]
                  if "math_function_id" in c else Unit(-1))
            tractions.append(
                ParticleTraction(pset=c["pset"],
                                 pids=self._get_particle_set_ids("particle", c["pset"],
                                                                 c["pset_ids"]),
                                 function=fn, dir=c["dir"], traction=c["traction"]))
        self.parsed_config["external_loading"] = out
        self.parsed_config["particle_surface_traction"] = tractions

    def _parse_mesh(self, config):
        """io.py:147-170: a generated structured mesh of the named element
        class with the constraints and nodal forces; Mesh1D or Mesh2D by
        meta.dimension."""
        mc = config["mesh"]
        if mc["type"] != "generator":
            raise NotImplementedError("Mesh type other than `generator` not yet supported.")
        element_cls = getattr(_element, mc["element"])
        mesh_cls = getattr(_mesh, f"Mesh{config['meta']['dimension']}D")
        constraints = [(self._get_node_set_ids(c["nset_ids"]), Constraint(c["dir"], c["velocity"]))
                       for c in mc["constraints"]]
        ne = mc["nelements"]
        el_len = mc["element_length"]
        if element_cls is _element.Linear1D:  # one-element lists name the 1D sizes too
            ne = int(np.asarray(ne).reshape(-1)[0])
            el_len = float(np.asarray(el_len).reshape(-1)[0])
        elements = element_cls(ne, int(np.prod(np.asarray(ne))), el_len, constraints,
                               concentrated_nodal_forces=self.parsed_config["external_loading"][
                                   "concentrated_nodal_forces"], **self._placement())
        self.parsed_config["elements"] = elements
        return mesh_cls(self.parsed_config)
