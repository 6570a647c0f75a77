"""Pins the oracle (CPU restatement) against the reference's own known-answer
unit tests and its three end-to-end golden benchmarks (fixtures in
tests/golden/, restated from the reference's tests/ and benchmarks/)."""
import numpy as np
import pytest

import cases
from oracle import mpm as om

KATS = cases.load_kats()
DTYPES = [np.float32, np.float64]


def q4(dtype, constraints=(), forces=()):
    return om.Elements("quad4", (1, 1), (1.0, 1.0), constraints, forces, dtype)


def two_particles(dtype):
    return om.Particles(np.array([[0.5, 0.5], [0.5, 0.5]]),
                        om.SimpleMaterial({"E": 1, "density": 1}), np.array([0, 0]), dtype)


def close(a, b):
    return np.allclose(np.asarray(a, np.float64), np.asarray(b, np.float64))


@pytest.mark.parametrize("dtype", DTYPES)
def test_shapefn(dtype):
    el = q4(dtype)
    for c in KATS["shapefn"]["cases"]:
        assert close(el.shapefn(np.array([c["xi"]], dtype)), [c["N"]])


@pytest.mark.parametrize("dtype", DTYPES)
def test_shapefn_natural_grad(dtype):
    el = q4(dtype)
    for c in KATS["shapefn_natural_grad"]["cases"]:
        assert close(el.shapefn_natural_grad(np.array([c["xi"]], dtype))[0], c["dN"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_shapefn_grad(dtype):
    el = q4(dtype)
    k = KATS["shapefn_grad"]
    coords = np.array([k["coords"]], dtype)
    for c in k["cases"]:
        xi = np.array([c["xi"]], dtype)
        assert close(el.shapefn_grad(xi, coords), el.shapefn_natural_grad(xi) * k["scale"])


def test_element_node_ids_loc_vel():
    el = q4(np.float32)
    assert el.id_to_node_ids(0).tolist() == KATS["element_node_ids"]["ids"]
    assert close(el.nodes.loc[el.id_to_node_ids(0)], KATS["element_node_loc"]["loc"])
    el.nodes.velocity += np.array(KATS["element_node_vel"]["add_velocity"], np.float32)
    assert close(el.nodes.velocity[el.id_to_node_ids(0)], KATS["element_node_vel"]["vel"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_p2g_kats(dtype):
    el, ps = q4(dtype), two_particles(dtype)
    ps.mass += 1
    el.compute_nodal_mass(ps)
    assert close(el.nodes.mass, KATS["nodal_mass"]["mass"])
    el, ps = q4(dtype), two_particles(dtype)
    ps.velocity += 1
    el.compute_nodal_momentum(ps)
    assert close(el.nodes.momentum, KATS["nodal_momentum"]["momentum"])
    el, ps = q4(dtype), two_particles(dtype)
    ps.f_ext += 1
    el.compute_external_force(ps)
    assert close(el.nodes.f_ext, KATS["external_force"]["f_ext"])
    for c in KATS["body_force"]["cases"]:
        el, ps = q4(dtype), two_particles(dtype)
        ps.mass += 1
        el.compute_body_force(ps, np.array(c["gravity"]))
        assert close(el.nodes.f_ext, c["f_ext"])
    el, ps = q4(dtype), two_particles(dtype)
    ps.traction += np.array(KATS["traction_forces"]["add_traction"], dtype)
    el.apply_particle_traction_forces(ps)
    assert close(el.nodes.f_ext, KATS["traction_forces"]["f_ext"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_grid_kats(dtype):
    k = KATS["concentrated_nodal_force"]
    forces = [om.NodalForce(np.array(f["nodes"]), om.Unit(-1), f["dir"], f["force"])
              for f in k["forces"]]
    el = q4(dtype, forces=forces)
    el.apply_concentrated_nodal_forces(two_particles(dtype), k["time"])
    assert close(el.nodes.f_ext, k["f_ext"])
    k = KATS["boundary_constraints"]
    cons = [(np.array(c["nodes"]), om.Constraint(c["dir"], c["velocity"]))
            for c in k["constraints"]]
    el = q4(dtype, constraints=cons)
    el.nodes.velocity += k["add_velocity"]
    el.apply_boundary_constraints()
    assert close(el.nodes.velocity, k["velocity"])
    k = KATS["nodal_acceleration_velocity"]
    el = q4(dtype)
    el.nodes.f_ext += np.array(k["add_fext"], dtype)
    el.nodes.mass[:] = k["set_mass"]
    el.update_nodal_acceleration_velocity(two_particles(dtype), k["dt"])
    assert close(el.nodes.acceleration, k["acceleration"])
    assert close(el.nodes.velocity, k["velocity"])
    assert close(el.nodes.momentum, k["momentum"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_element_ids_internal_force_volume(dtype):
    el, ps = q4(dtype), two_particles(dtype)
    ps.element_ids = np.array(KATS["particle_element_ids"]["set_ids"])
    ps.set_particle_element_ids(el)
    assert ps.element_ids.tolist() == KATS["particle_element_ids"]["ids"]
    el, ps = q4(dtype), two_particles(dtype)
    ps.compute_volume(el, 1)
    ps.stress += KATS["internal_force"]["add_stress"]
    el.compute_internal_force(ps)
    assert close(el.nodes.f_int, KATS["internal_force"]["f_int"])
    el = q4(dtype)
    el.compute_volume()
    assert close(el.volume, KATS["element_volume"]["volume"])


@pytest.mark.parametrize("case", KATS["material_stress"]["cases"])
def test_material_stress(case):
    mat = cases.oracle_material(case["material"])
    got = mat.compute_stress(np.array([case["dstrain"]], np.float32))
    assert close(got, [case["stress"]])


@pytest.mark.parametrize("dtype", DTYPES)
def test_particle_kats(dtype):
    k = KATS["update_velocity"]
    for c in k["cases"]:
        el, ps = q4(dtype), two_particles(dtype)
        ps.update_natural_coords(el)
        el.nodes.acceleration += k["add_node_acceleration"]
        el.nodes.velocity += k["add_node_velocity"]
        ps.update_position_velocity(el, k["dt"], c["velocity_update"])
        assert close(ps.velocity, c["velocity"])
    k = KATS["compute_strain"]
    el, ps = q4(dtype), two_particles(dtype)
    el.nodes.velocity = np.array(k["node_velocity"], dtype)
    ps.update_natural_coords(el)
    ps.compute_strain(el, k["dt"])
    assert close(ps.strain, k["strain"])
    assert close(ps.volumetric_strain_centroid, k["volumetric_strain_centroid"])
    el, ps = q4(dtype), two_particles(dtype)
    ps.compute_volume(el, el.total_elements)
    assert close(ps.volume, KATS["particle_volume"]["volume"])
    k = KATS["assign_traction"]
    ps.assign_traction(np.array(k["pids"]), k["dir"], k["traction"])
    assert close(ps.traction, k["result"])
    ps.traction += KATS["zero_traction"]["add_traction"]
    ps.zero_traction()
    assert close(ps.traction, KATS["zero_traction"]["result"])


GOLDENS = cases.load_goldens()


@pytest.mark.parametrize("name", sorted(GOLDENS))
@pytest.mark.parametrize("dtype", DTYPES)
def test_goldens(name, dtype):
    case = GOLDENS[name]
    sim = cases.build_oracle_golden(case, dtype)
    by_step = {}
    for c in case["checks"]:
        by_step.setdefault(c["after_step"], []).append(c)
    failures = []

    def cb(i, s):
        for c in by_step.get(i, []):
            got, ok = cases.check_value(s.mesh.particles[0].stress, c)
            if not ok:
                failures.append((i, c, got))

    sim.sim_steps = max(by_step) + 1
    sim.solve(np.asarray(case["gravity"]), cb)
    assert not failures, failures


def test_goldens_fp64_tight():
    """The FP64 oracle reproduces the reference goldens (which carry 16
    digits) to 1e-12 up to step 750 of the nodal-force bar."""
    case = GOLDENS["uniaxial_nodal_forces"]
    sim = cases.build_oracle_golden(case, np.float64)
    errs = []

    def cb(i, s):
        for c in case["checks"]:
            if c["after_step"] == i and i <= 750:
                got, _ = cases.check_value(s.mesh.particles[0].stress, c)
                errs.append(abs(got - c["value"]))

    sim.sim_steps = 751
    sim.solve(np.zeros(2), cb)
    assert len(errs) == 6 and max(errs) < 1e-12, errs
