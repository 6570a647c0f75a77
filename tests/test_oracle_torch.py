"""The torch CPU restatement of the 2D step (``oracle/torch_step.py``, the
multi-threaded CPU baseline of bench.py) agrees with the NumPy oracle on
seeded cases: same element ids and particle state within FP32 rounding."""
import numpy as np
import pytest
import torch

import cases
from oracle import mpm as om
from oracle.torch_step import TorchQ4


def build(spec):
    nx, ny = spec["nelements"]
    h = spec["el_len"]
    mat = spec["material"]
    de = om.LinearElastic({k: mat[k] for k in ("density", "youngs_modulus",
                                               "poisson_ratio")}).de64
    cons = [(c["nodes"], c["dir"], c["velocity"]) for c in spec["constraints"]]
    pts = np.asarray(spec["particles"], np.float64)
    t = TorchQ4(nx, ny, h[0], h[1], torch.from_numpy(pts.astype(np.float32)), de,
                mat["density"], cons, spec["dt"], spec["scheme"], spec["velocity_update"])
    t.velocity = torch.as_tensor(np.asarray(spec["init_velocity"]), dtype=torch.float32)
    return t


@pytest.mark.parametrize("scheme,vu", [("usf", False), ("usl", False), ("usf", True)])
def test_torch_step_matches_numpy_oracle(scheme, vu):
    spec = cases.synthetic_spec("quad4", nx=12, ny=10, rows=7, seed=31, scheme=scheme,
                                velocity_update=vu, traction=False, nodal_force=False,
                                jitter=0.3, nsteps=10)
    osim = cases.build_oracle(spec, np.float32)
    t = build(spec)
    ops = osim.mesh.particles[0]
    assert np.array_equal(t.eid.numpy(), ops.element_ids)
    assert np.allclose(t.volume.numpy(), ops.volume, rtol=1e-6)
    g = np.asarray(spec["gravity"])
    for i in range(spec["nsteps"]):
        osim.step(i, g)
        t.step(i, g)
        assert np.array_equal(t.eid.numpy(), ops.element_ids), i
        for name, got in (("loc", t.loc), ("velocity", t.velocity), ("stress", t.stress),
                          ("strain", t.strain), ("volume", t.volume),
                          ("density", t.density),
                          ("volumetric_strain_centroid", t.evc)):
            want = cases.oracle_particle_field(ops, name)
            err = cases.normwise(got.double().numpy().reshape(want.shape), want)
            assert err < 1e-5, (i, name, err)
