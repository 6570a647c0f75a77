"""TOML/JSON input front-end: ``Config`` reads a simulation file in the
reference's format and builds this package's GPU objects from it.

The file format is the reference's interface (diffmpm/io.py:16-170; its
benchmark input files under benchmarks/2d/* run unmodified):

* ``[meta]`` title, type ("MPMExplicit"), dimension, scheme, dt, nsteps,
  velocity_update; ``[output]`` format, folder, step_frequency;
* ``[mesh]`` type ("generator"), element (class name), nelements,
  element_length, entity_sets (a JSON file of ``node_sets`` {id: [nodes]}
  and ``particle_sets`` [{id: [particles]}, ...]) and ``[[mesh.constraints]]``
  (nset_ids, dir, velocity);
* ``[[materials]]`` (type = class name, the other keys are its properties);
* ``[[particles]]`` (file: JSON positions of shape (N, 1, ndim), material_id,
  init_velocity);
* ``[[math_functions]]`` (type "Linear", xvalues, fxvalues), referred to by
  position;
* ``[external_loading]`` gravity, ``[[...concentrated_nodal_forces]]``
  (nset_ids, dir, force, optional math_function_id) and
  ``[[...particle_surface_traction]]`` (pset, pset_ids, dir, traction,
  optional math_function_id); without a function id a load is constant
  (``Unit``).

``parsed_config`` holds what the reference's ``Config`` holds (meta,
output, materials, particles, math_functions, external_loading,
particle_surface_traction, elements), so ``Mesh1D``/``Mesh2D`` and ``MPM``
take it unchanged.  Relative file names are opened from the current
directory, as the reference does, or else from the TOML file's directory.
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


def _function(index, rec):
    kind = rec["type"]
    if kind != "Linear":
        raise NotImplementedError(f"math function type {kind!r}: only 'Linear' is implemented")
    return Linear(index, np.asarray(rec["xvalues"]), np.asarray(rec["fxvalues"]))


def _material_from(rec):
    props = {k: v for k, v in rec.items() if k != "type"}
    return getattr(_material, rec["type"])(props)


class Config:
    """A simulation read from a TOML file; ``parse()`` returns its mesh."""

    def __init__(self, filepath, *, device=None, dtype=None):
        self._filepath = filepath
        self._device = device
        self._dtype = dtype
        self.parsed_config = {}
        self.parse()

    # -- files and sets ---------------------------------------------------------
    def _locate(self, name):
        if os.path.isabs(name) or os.path.exists(name):
            return name
        beside = os.path.join(os.path.dirname(os.path.abspath(self._filepath)), name)
        return beside if os.path.exists(beside) else name

    def _read_json(self, name):
        with open(self._locate(name)) as f:
            return json.load(f)

    def _node_ids(self, set_ids):
        """The nodes of the named node sets, set after set."""
        table = self.entity_sets["node_sets"]
        return torch.tensor([n for s in set_ids for n in table[str(s)]], dtype=torch.int64)

    def _particle_ids(self, psets, set_ids):
        """The particles of the named sets of each listed particle-set group."""
        groups = self.entity_sets["particle_sets"]
        ids = [p for g in psets for s in set_ids for p in groups[g][str(s)]]
        return torch.tensor(ids, dtype=torch.int64)

    def _on_device(self):
        from ._lib import default_device
        kw = {"device": default_device() if self._device is None else self._device}
        if self._dtype is not None:
            kw["dtype"] = self._dtype
        return kw

    # -- objects ---------------------------------------------------------------
    def _particle_set(self, rec, materials):
        loc = np.asarray(self._read_json(rec["file"]), dtype=np.float64)
        ps = Particles(torch.from_numpy(loc), materials[rec["material_id"]],
                       torch.zeros(loc.shape[0], dtype=torch.int32), **self._on_device())
        v0 = torch.as_tensor(rec["init_velocity"], dtype=torch.float64).to(ps.dtype)
        ps.velocity = v0.to(ps.device).expand(len(ps), 1, loc.shape[-1]).clone()
        return ps

    def _elements(self, mesh, nodal_forces):
        if mesh["type"] != "generator":
            raise NotImplementedError(
                f"mesh type {mesh['type']!r}: only 'generator' is implemented")
        cls = getattr(_element, mesh["element"])
        cons = [(self._node_ids(c["nset_ids"]), Constraint(c["dir"], c["velocity"]))
                for c in mesh["constraints"]]
        counts, length = mesh["nelements"], mesh["element_length"]
        if cls is _element.Linear1D:  # a one-element list names a 1D size
            counts = int(np.ravel(counts)[0])
            length = float(np.ravel(length)[0])
        total = int(np.prod(np.ravel(counts)))
        return cls(counts, total, length, cons, concentrated_nodal_forces=nodal_forces,
                   **self._on_device())

    def parse(self):
        """Reads the file and builds the mesh (the mesh class named by
        meta.dimension) with its particle sets, loads and constraints."""
        with open(self._filepath, "rb") as f:
            doc = _toml.load(f)
        self._fileconfig = doc
        self.entity_sets = self._read_json(doc["mesh"]["entity_sets"])
        pc = {"meta": doc["meta"], "output": doc["output"]}
        pc["materials"] = [_material_from(rec) for rec in doc["materials"]]
        pc["particles"] = [self._particle_set(rec, pc["materials"]) for rec in doc["particles"]]
        fns = [_function(k, rec) for k, rec in enumerate(doc.get("math_functions", []))]
        if "math_functions" in doc:
            pc["math_functions"] = fns

        def fn_of(rec):
            return fns[rec["math_function_id"]] if "math_function_id" in rec else Unit(-1)

        loads = doc["external_loading"]
        nodal = [NodalForce(node_ids=self._node_ids(r["nset_ids"]), function=fn_of(r),
                            dir=r["dir"], force=r["force"])
                 for r in loads.get("concentrated_nodal_forces", [])]
        pc["external_loading"] = {"gravity": torch.as_tensor(loads["gravity"],
                                                             dtype=torch.float64),
                                  "concentrated_nodal_forces": nodal}
        pc["particle_surface_traction"] = [
            ParticleTraction(pset=r["pset"], pids=self._particle_ids(r["pset"], r["pset_ids"]),
                             function=fn_of(r), dir=r["dir"], traction=r["traction"])
            for r in loads.get("particle_surface_traction", [])]
        pc["elements"] = self._elements(doc["mesh"], nodal)
        self.parsed_config = pc
        return getattr(_mesh, f"Mesh{doc['meta']['dimension']}D")(pc)
