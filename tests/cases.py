"""Builders shared by the tests: the reference's end-to-end golden cases and
seeded synthetic meshes, for both the oracle and the HIP package."""
import json
import math
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


def load_goldens():
    with open(os.path.join(GOLDEN_DIR, "reference_goldens.json")) as f:
        d = json.load(f)
    d.pop("_source", None)
    return d


def load_kats():
    with open(os.path.join(GOLDEN_DIR, "reference_kats.json")) as f:
        return json.load(f)


def oracle_function(spec):
    from oracle import mpm as om
    if spec is None or spec["type"] == "Unit":
        return om.Unit(-1)
    return om.Linear(0, spec["xvalues"], spec["fxvalues"])


def oracle_material(spec):
    from oracle import mpm as om
    spec = dict(spec)
    kind = spec.pop("type")
    return getattr(om, kind)(spec)


def build_oracle_golden(case, dtype=np.float32):
    """Restates the TOML front-end (io.py:52-170) for one golden case."""
    from oracle import mpm as om
    cons = [(np.asarray(c["nodes"]), om.Constraint(c["dir"], c["velocity"]))
            for c in case["constraints"]]
    cnfs = [om.NodalForce(np.asarray(f["nodes"]), oracle_function(f.get("function")),
                          f["dir"], f["force"]) for f in case["nodal_forces"]]
    el = om.Elements("quad4", case["nelements"], case["el_len"], cons, cnfs, dtype)
    ps = om.Particles(np.asarray(case["particles"]), oracle_material(case["material"]),
                      dtype=dtype)
    ps.velocity[:] = np.asarray(case["init_velocity"], dtype=np.float64).astype(dtype)
    trs = [om.ParticleTraction(t["pset"], np.asarray(t["pids"]),
                               oracle_function(t.get("function")), t["dir"], t["traction"])
           for t in case["tractions"]]
    mesh = om.Mesh(el, [ps], trs)
    return om.MPMExplicit(mesh, case["dt"], case["scheme"], case["velocity_update"],
                          case["nsteps"])


def check_value(stress, chk):
    """The reference's assertion: round(reduce(stress[:, comp]) - value, d) == 0."""
    col = np.asarray(stress)[:, chk["component"]].astype(np.float64)
    got = col.max() if chk["reduce"] == "max" else col.min()
    return got, round(got - chk["value"], chk["decimals"]) == 0.0
