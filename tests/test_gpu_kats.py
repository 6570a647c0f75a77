"""The reference's unit tests (tests/test_element.py, test_material.py,
test_particle.py), run through the HIP path, in FP32 and FP64.  Expected
values come from tests/golden/reference_kats.json."""
import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu
KATS = cases.load_kats()
DTYPES = [torch.float32, torch.float64]


def d():
    import diffmpm_amd
    return diffmpm_amd


def elements(dtype, constraints=(), forces=()):
    return d().Quadrilateral4Node((1, 1), 1, (1.0, 1.0), list(constraints),
                                  concentrated_nodal_forces=list(forces), device="cuda",
                                  dtype=dtype)


def particles(dtype):
    loc = torch.tensor([[0.5, 0.5], [0.5, 0.5]]).reshape(2, 1, 2)
    return d().Particles(loc, d().SimpleMaterial({"E": 1, "density": 1}),
                         torch.tensor([0, 0]), device="cuda", dtype=dtype)


def close(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(np.asarray(b, np.float64)).reshape(a.shape) if np.size(b) == a.numel() \
        else torch.as_tensor(np.asarray(b, np.float64))
    return torch.allclose(a, b.double().expand_as(a))


@pytest.mark.parametrize("dtype", DTYPES)
def test_node_ids_loc_vel(dtype):
    el = elements(dtype)
    assert el.id_to_node_ids(0).reshape(-1).tolist() == KATS["element_node_ids"]["ids"]
    assert close(el.id_to_node_loc(0), KATS["element_node_loc"]["loc"])
    el.nodes.velocity += torch.tensor(KATS["element_node_vel"]["add_velocity"], device="cuda")
    assert close(el.id_to_node_vel(0), KATS["element_node_vel"]["vel"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_nodal_mass_momentum_forces(dtype):
    el, ps = elements(dtype), particles(dtype)
    ps.mass += 1
    el.compute_nodal_mass(ps)
    assert close(el.nodes.mass, KATS["nodal_mass"]["mass"])
    el, ps = elements(dtype), particles(dtype)
    ps.velocity += 1
    el.compute_nodal_momentum(ps)
    assert close(el.nodes.momentum, KATS["nodal_momentum"]["momentum"])
    el, ps = elements(dtype), particles(dtype)
    ps.f_ext += 1
    el.compute_external_force(ps)
    assert close(el.nodes.f_ext, KATS["external_force"]["f_ext"])
    for c in KATS["body_force"]["cases"]:
        el, ps = elements(dtype), particles(dtype)
        ps.mass += 1
        el.compute_body_force(ps, torch.tensor(c["gravity"]))
        assert close(el.nodes.f_ext, c["f_ext"])
    el, ps = elements(dtype), particles(dtype)
    el.compute_body_force(ps, 0)  # scalar gravity, tests/test_element.py:128
    assert close(el.nodes.f_ext, np.zeros((4, 2)))
    el, ps = elements(dtype), particles(dtype)
    ps.traction += torch.tensor(KATS["traction_forces"]["add_traction"], device="cuda")
    el.apply_particle_traction_forces(ps)
    assert close(el.nodes.f_ext, KATS["traction_forces"]["f_ext"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_grid_methods(dtype):
    k = KATS["concentrated_nodal_force"]
    forces = [d().NodalForce(torch.tensor(f["nodes"]), d().Unit(-1), f["dir"], f["force"])
              for f in k["forces"]]
    el = d().Quadrilateral4Node((1, 1), 1, 1, [], concentrated_nodal_forces=forces,
                                device="cuda", dtype=dtype)
    el.apply_concentrated_nodal_forces(particles(dtype), k["time"])
    assert close(el.nodes.f_ext, k["f_ext"])
    k = KATS["boundary_constraints"]
    cons = [(torch.tensor(c["nodes"]), d().Constraint(c["dir"], c["velocity"]))
            for c in k["constraints"]]
    el = elements(dtype, constraints=cons)
    el.nodes.velocity += k["add_velocity"]
    el.apply_boundary_constraints()
    assert close(el.nodes.velocity, k["velocity"])
    k = KATS["nodal_acceleration_velocity"]
    el = elements(dtype)
    el.nodes.f_ext += torch.tensor(k["add_fext"], device="cuda")
    el.nodes.mass = k["set_mass"]
    el.update_nodal_acceleration_velocity(particles(dtype), k["dt"])
    assert close(el.nodes.acceleration, k["acceleration"])
    assert close(el.nodes.velocity, k["velocity"])
    assert close(el.nodes.momentum, k["momentum"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_element_ids_internal_force_volume(dtype):
    el, ps = elements(dtype), particles(dtype)
    ps.element_ids = torch.tensor(KATS["particle_element_ids"]["set_ids"])
    el.set_particle_element_ids(ps)
    assert ps.element_ids.tolist() == KATS["particle_element_ids"]["ids"]
    el, ps = elements(dtype), particles(dtype)
    ps.compute_volume(el, 1)
    ps.stress += KATS["internal_force"]["add_stress"]
    el.compute_internal_force(ps)
    assert close(el.nodes.f_int, KATS["internal_force"]["f_int"])
    el = elements(dtype)
    el.compute_volume()
    assert close(el.volume, KATS["element_volume"]["volume"])


@pytest.mark.parametrize("case", KATS["material_stress"]["cases"])
def test_material_compute_stress(case):
    mat = cases.gpu_material(case["material"])
    ds = torch.tensor(case["dstrain"], device="cuda").reshape(1, 6, 1)
    assert close(mat.compute_stress(ds), case["stress"])


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["material_stress"]["cases"])
def test_material_stress_kernel(dtype, case):
    """The stress kernel (dmpm_compute_stress) against the same KATs."""
    mat = cases.gpu_material(case["material"])
    ps = d().Particles(torch.tensor([[[0.5, 0.5]]]), mat, torch.tensor([0]), device="cuda",
                       dtype=dtype)
    ps.dstrain = torch.tensor(case["dstrain"]).reshape(1, 6, 1)
    ps.compute_stress()
    assert close(ps.stress, case["stress"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_particle_methods(dtype):
    k = KATS["update_velocity"]
    for c in k["cases"]:
        el, ps = elements(dtype), particles(dtype)
        ps.update_natural_coords(el)
        el.nodes.acceleration += k["add_node_acceleration"]
        el.nodes.velocity += k["add_node_velocity"]
        ps.update_position_velocity(el, k["dt"], c["velocity_update"])
        assert close(ps.velocity, c["velocity"])
    k = KATS["compute_strain"]
    el, ps = elements(dtype), particles(dtype)
    el.nodes.velocity = torch.tensor(k["node_velocity"], dtype=dtype).reshape(4, 1, 2)
    ps.update_natural_coords(el)
    ps.compute_strain(el, k["dt"])
    assert close(ps.strain, k["strain"])
    assert close(ps.volumetric_strain_centroid, k["volumetric_strain_centroid"])
    el, ps = elements(dtype), particles(dtype)
    ps.compute_volume(el, el.total_elements)
    assert close(ps.volume, KATS["particle_volume"]["volume"])
    k = KATS["assign_traction"]
    ps.assign_traction(torch.tensor(k["pids"]), k["dir"], k["traction"])
    assert close(ps.traction, k["result"])
    ps.traction += KATS["zero_traction"]["add_traction"]
    ps.zero_traction()
    assert close(ps.traction, KATS["zero_traction"]["result"])


def test_native_library_loaded():
    """The HIP library is what runs: it is loaded in this process."""
    import diffmpm_amd._lib as L
    L.lib()
    with open("/proc/self/maps") as f:
        assert "libdiffmpm_amd.so" in f.read()
