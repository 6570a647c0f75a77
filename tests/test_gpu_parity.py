"""Parity of the HIP step with the oracle (CPU restatement).

* reference goldens end to end (FP32 and FP64);
* per-step parity on seeded synthetic cases covering USF/USL, velocity_update,
  Linear1D and Quadrilateral4Node, constraints, ramped nodal forces, particle
  tractions, gravity and per-particle E: the GPU state is re-synchronised to
  the oracle's before every step, and after the step every particle and nodal
  field must agree normwise within RTOL_F32 = 1e-5 (FP32 vs FP32 oracle) and
  RTOL_F64 = 1e-12 (FP64 vs FP64 oracle);
* rollout drift without re-synchronisation;
* bitwise determinism, and bitwise equality of the fused run with the
  per-method path.
"""
import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu

RTOL_F32 = 1e-5
RTOL_F64 = 1e-12
GOLDENS = cases.load_goldens()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("name", sorted(GOLDENS))
def test_reference_goldens(name, dtype):
    case = GOLDENS[name]
    steps = max(c["after_step"] for c in case["checks"]) + 1
    want = {c["after_step"] for c in case["checks"]}
    seen = {}

    def writer(args, transforms, **kw):
        arrays, i = args
        if i in want:
            seen[i] = arrays["stress"][0].reshape(-1, 6).copy()

    sim = cases.build_gpu(case, dtype)
    sim.sim_steps = steps
    sim.out_steps = 10 if steps > 100 else 1
    sim.writer_func = writer
    sim.solve_jit(torch.tensor(case["gravity"], dtype=torch.float64))
    for c in case["checks"]:
        got, ok = cases.check_value(seen[c["after_step"]], c)
        assert ok, (c, got)


VARIANTS = [
    dict(element="quad4", scheme="usf", velocity_update=False),
    dict(element="quad4", scheme="usl", velocity_update=False),
    dict(element="quad4", scheme="usf", velocity_update=True),
    dict(element="quad4", scheme="usl", velocity_update=True, per_particle_E=True),
    dict(element="quad4", scheme="usf", velocity_update=False, per_particle_E=True, seed=3),
    dict(element="linear1d", nx=16, scheme="usf", velocity_update=False),
    dict(element="linear1d", nx=16, scheme="usl", velocity_update=True),
    dict(element="quad4", scheme="usf", velocity_update=False, traction=False),
    dict(element="quad4", scheme="usl", velocity_update=True, traction=False, per_particle_E=True),
]


def _compare(osim, gsim, rtol, where):
    ops, gps = osim.mesh.particles[0], gsim.mesh.particles[0]
    assert np.array_equal(ops.element_ids, gps.element_ids.cpu().numpy()), where
    errs = {}
    for name in cases.PARTICLE_FIELDS:
        errs["p." + name] = cases.normwise(cases.gpu_particle_field(gps, name),
                                           cases.oracle_particle_field(ops, name))
    for name in cases.NODE_FIELDS:
        errs["n." + name] = cases.normwise(cases.gpu_node_field(gsim.mesh.elements.nodes, name),
                                           cases.oracle_node_field(osim.mesh.elements.nodes, name))
    bad = {k: v for k, v in errs.items() if not v <= rtol}
    assert not bad, (where, bad)
    return max(errs.values())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: "-".join(f"{k}={v[k]}" for k in v))
def test_per_step_parity(variant, dtype):
    spec = cases.synthetic_spec(nsteps=8, **variant)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    rtol = RTOL_F64 if dtype == torch.float64 else RTOL_F32
    osim = cases.build_oracle(spec, npdt)
    gsim = cases.build_gpu(spec, dtype)
    g = np.asarray(spec["gravity"])
    _compare(osim, gsim, rtol, "init")
    for i in range(spec["nsteps"]):
        cases.load_oracle_state_into_gpu(osim, gsim)
        osim.step(i, g)
        gsim._step(i, torch.tensor(spec["gravity"], dtype=torch.float64))
        _compare(osim, gsim, rtol, f"step {i}")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_simple_material_per_step_parity(scheme, dtype):
    """SimpleMaterial (material.py:119-131: sigma += E * dstrain) through the
    fused phases, per step against the oracle."""
    spec = cases.synthetic_spec(nsteps=6, scheme=scheme, seed=17)
    spec["material"] = {"type": "SimpleMaterial", "E": 1e6, "density": 1000.0}
    npdt = np.float64 if dtype == torch.float64 else np.float32
    rtol = RTOL_F64 if dtype == torch.float64 else RTOL_F32
    osim = cases.build_oracle(spec, npdt)
    gsim = cases.build_gpu(spec, dtype)
    g = np.asarray(spec["gravity"])
    for i in range(spec["nsteps"]):
        cases.load_oracle_state_into_gpu(osim, gsim)
        osim.step(i, g)
        gsim._step(i, torch.tensor(spec["gravity"], dtype=torch.float64))
        _compare(osim, gsim, rtol, f"step {i}")


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_simple_material_one_launch_step_equals_phase_launches(scheme):
    spec = _moving_spec(scheme, drift=(0.6, 0.3))
    spec["material"] = {"type": "SimpleMaterial", "E": 1e3, "density": 1000.0}
    _run_pair(spec, scheme, 8, 4)


def _two_set_spec(scheme):
    """Two particle sets of the same size on one 8x6 mesh, both over every
    cell: E = 1e6 and E = 3e6 (nu = 0.25), own jitter and velocities.  The
    reference maps each set to the nodes after zeroing the nodal arrays, so
    the nodal quantities of a step are the last set's (element.py:143, 175,
    225, 586, 887) and the first set moves with them; the second set covers
    every node, so no node is massless.  Both sides run the sets in list
    order (solve_jit stacks the sets' arrays, so they have one size, as in
    the reference's solver.py:214-222)."""
    from oracle.mpm import gauss_points_2d
    spec = cases.synthetic_spec(nx=8, ny=6, rows=6, seed=13, scheme=scheme, traction=False,
                                nsteps=6)
    h = spec["el_len"][0]
    second = gauss_points_2d(8, 6, h, h, rows=6, jitter=0.2, seed=14, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(15))
    spec["extra_sets"] = [dict(particles=second.reshape(len(second), -1).tolist(),
                               init_velocity=(0.3 * rng.standard_normal((len(second), 2))).tolist(),
                               material={"type": "LinearElastic", "density": 1500.0,
                                         "youngs_modulus": 3e6, "poisson_ratio": 0.25})]
    return spec


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_two_particle_sets_per_step_parity(scheme, dtype):
    """Every particle field of both sets and every nodal field after each step
    (state re-synchronised from the oracle before each step)."""
    spec = _two_set_spec(scheme)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    rtol = RTOL_F64 if dtype == torch.float64 else RTOL_F32
    osim = cases.build_oracle(spec, npdt)
    gsim = cases.build_gpu(spec, dtype)
    assert len(gsim.mesh.particles) == 2
    g = np.asarray(spec["gravity"])
    for i in range(spec["nsteps"]):
        cases.load_oracle_state_into_gpu(osim, gsim)
        osim.step(i, g)
        gsim._step(i, torch.tensor(spec["gravity"], dtype=torch.float64))
        for k, (ops, gps) in enumerate(zip(osim.mesh.particles, gsim.mesh.particles)):
            assert np.array_equal(ops.element_ids, gps.element_ids.cpu().numpy()), (i, k)
            for name in cases.PARTICLE_FIELDS:
                err = cases.normwise(cases.gpu_particle_field(gps, name),
                                     cases.oracle_particle_field(ops, name))
                assert err <= rtol, (i, k, name, err)
        for name in cases.NODE_FIELDS:
            err = cases.normwise(cases.gpu_node_field(gsim.mesh.elements.nodes, name),
                                 cases.oracle_node_field(osim.mesh.elements.nodes, name))
            assert err <= rtol, (i, name, err)


def test_two_particle_sets_solve_jit_equals_steps():
    """solve_jit with two sets (the per-method step, one set after the
    other) equals the same steps through MPMExplicit._step."""
    spec = _two_set_spec("usf")
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    a.solve_jit(g)
    for i in range(spec["nsteps"]):
        b._step(i, g)
    for pa, pb in zip(a.mesh.particles, b.mesh.particles):
        assert torch.equal(pa._buf.view(torch.int32), pb._buf.view(torch.int32))


@pytest.mark.parametrize("variant", VARIANTS[:2] + VARIANTS[5:6],
                         ids=lambda v: "-".join(f"{k}={v[k]}" for k in v))
def test_rollout_drift(variant):
    """FP32 GPU rollout of 50 steps against the FP64 oracle: the drift stays
    at FP32 rounding level (normwise 1e-4 on every field)."""
    spec = cases.synthetic_spec(nsteps=50, **variant)
    osim = cases.build_oracle(spec, np.float64)
    gsim = cases.build_gpu(spec, torch.float32)
    osim.solve(np.asarray(spec["gravity"]))
    gsim.solve_jit(torch.tensor(spec["gravity"], dtype=torch.float64))
    worst = _compare(osim, gsim, 1e-4, "rollout")
    print(f"rollout drift fp32 vs fp64 oracle after 50 steps: {worst:.3e}")


def _state(gsim):
    """Bit patterns of the particle block, node block and element ids (so
    NaNs compare equal when their bits do)."""
    ps = gsim.mesh.particles[0]
    out = []
    for t in (ps._buf, gsim.mesh.elements.nodes._buf, ps._eid):
        t = t.clone()
        if t.is_floating_point():  # any NaN encoding counts as the same NaN
            t = torch.where(torch.isnan(t), torch.full_like(t, float("nan")), t)
        if t.dtype == torch.float32:
            t = t.view(torch.int32)
        elif t.dtype == torch.float64:
            t = t.view(torch.int64)
        out.append(t)
    return out


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_bitwise_determinism(scheme):
    spec = cases.synthetic_spec(nx=16, ny=12, rows=9, scheme=scheme, nsteps=30, seed=5)
    a = cases.build_gpu(spec)
    b = cases.build_gpu(spec)
    a.solve_jit(torch.tensor(spec["gravity"], dtype=torch.float64))
    b.solve_jit(torch.tensor(spec["gravity"], dtype=torch.float64))
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)


def _per_method_step(gsim, i, gravity):
    """The reference's step fanned out method by method (mesh.py:31-58)."""
    m, s = gsim.mesh, gsim.mpm_scheme
    t = float(np.float32(i) * np.float32(gsim.dt))
    m.apply_on_elements("set_particle_element_ids")
    m.apply_on_particles("update_natural_coords")
    m.apply_on_elements("compute_nodal_mass")
    m.apply_on_elements("compute_nodal_momentum")
    m.apply_on_elements("compute_velocity")
    m.apply_on_elements("apply_boundary_constraints")

    def stress():
        m.apply_on_particles("compute_strain", args=(gsim.dt,))
        m.apply_on_particles("update_volume")
        m.apply_on_particles("compute_stress")
    if gsim.scheme == "usf":
        stress()
    m.apply_on_elements("compute_external_force")
    m.apply_on_elements("compute_body_force", args=(gravity,))
    m.apply_traction_on_particles(t)
    m.apply_on_elements("apply_concentrated_nodal_forces", args=(t,))
    m.apply_on_elements("compute_internal_force")
    m.apply_on_elements("update_nodal_acceleration_velocity", args=(gsim.dt,))
    m.apply_on_particles("update_position_velocity", args=(gsim.dt, gsim.velocity_update))
    if gsim.scheme == "usl":
        stress()


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: "-".join(f"{k}={v[k]}" for k in v))
def test_fused_equals_per_method(variant):
    spec = cases.synthetic_spec(nsteps=6, **variant)
    a = cases.build_gpu(spec)
    b = cases.build_gpu(spec)
    a.solve_jit(torch.tensor(spec["gravity"], dtype=torch.float64))
    for i in range(spec["nsteps"]):
        _per_method_step(b, i, torch.tensor(spec["gravity"], dtype=torch.float64))
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)


def _crowded_spec(n_crowd, seed=11):
    """A 4x4 mesh with n_crowd particles in one cell plus a sparse background
    (exercises the big-cell path of the counting sort)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = 0.25
    crowd = np.stack([rng.uniform(1.05, 1.95, n_crowd), rng.uniform(1.05, 1.95, n_crowd)], 1) * h
    bg = rng.uniform(0.02, 0.98, size=(40, 2))
    pts = np.concatenate([bg[:20], crowd, bg[20:]])
    return dict(element="quad4", nelements=[4, 4], el_len=[h, h], particles=pts.tolist(),
                init_velocity=(0.1 * rng.standard_normal(pts.shape)).tolist(),
                material={"type": "LinearElastic", "density": 1000.0, "youngs_modulus": 1e5,
                          "poisson_ratio": 0.3},
                constraints=[{"nodes": list(range(5)), "dir": 1, "velocity": 0.0}],
                nodal_forces=[], tractions=[], gravity=[0.0, -9.81], scheme="usf", dt=1e-5,
                nsteps=3, velocity_update=False)


@pytest.mark.parametrize("driver", ["run", "methods"])
@pytest.mark.parametrize("n_crowd", [100, 3000, 9000])
def test_crowded_cell_parity_and_determinism(n_crowd, driver):
    """A cell with n_crowd particles (9000: beyond the 8192 that the rank sort
    orders in LDS) through the fused run and the per-method path: two runs
    agree bit for bit (the sort puts every cell in ascending particle index,
    so the summation order is fixed) and both match the oracle."""
    spec = _crowded_spec(n_crowd)
    osim = cases.build_oracle(spec, np.float32)
    o64 = cases.build_oracle(spec, np.float64)
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    osim.solve(np.asarray(spec["gravity"]))
    o64.solve(np.asarray(spec["gravity"]))
    for sim in (a, b):
        if driver == "run":
            sim.solve_jit(g)
        else:
            for i in range(spec["nsteps"]):
                sim._step(i, g)
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)
    if n_crowd <= 3000:
        _compare(osim, a, 1e-5, "crowded")
        return
    # 9000 particles in one cell: its nodal sums have 9000 FP32 terms, and
    # the centroid strain rate (a difference of those nodal velocities)
    # carries their rounding; within the reference's own FP32 precision
    ops, o6, gps = osim.mesh.particles[0], o64.mesh.particles[0], a.mesh.particles[0]
    assert np.array_equal(ops.element_ids, gps.element_ids.cpu().numpy())
    for name in cases.PARTICLE_FIELDS:
        ok, *errs = cases.fp32_within_reference_precision(
            cases.gpu_particle_field(gps, name), cases.oracle_particle_field(ops, name),
            cases.oracle_particle_field(o6, name))
        assert ok, (name, errs)


def test_particles_outside_mesh_follow_the_reference():
    """Particles left of / below the mesh (element 0), beyond the last node
    column (next row), above the mesh (element id >= nx*ny) and at NaN: the
    element ids equal the brute-force restatement of the reference's search
    (element.py:820-840); after 3 steps the set of non-finite particles is the
    oracle's (the NaN particle poisons the nodes of element 0, and through
    them its neighbours, as the reference's scatters do; the particle above
    the mesh is NaN in both) and every other particle matches the oracle."""
    spec = cases.synthetic_spec(nx=8, ny=8, rows=6, seed=2, nsteps=3)
    pts = np.asarray(spec["particles"])
    pts[:5] = [[-0.1, 0.5], [0.5, -0.2], [1.2, 0.3], [np.nan, 0.4], [0.3, 1.5]]
    spec["particles"] = pts.tolist()
    osim = cases.build_oracle(spec, np.float32)
    sim = cases.build_gpu(spec)
    ps, ops = sim.mesh.particles[0], osim.mesh.particles[0]
    brute = osim.mesh.elements.element_ids_bruteforce(pts.astype(np.float32))
    assert brute[:5].tolist() == [0, 0, 24, 0, 66]
    assert np.array_equal(ps.element_ids.cpu().numpy(), brute)
    with np.errstate(all="ignore"):
        osim.solve(np.asarray(spec["gravity"]))
    sim.solve_jit(torch.tensor(spec["gravity"], dtype=torch.float64))
    torch.cuda.synchronize()
    fields = ("loc", "velocity", "stress", "strain", "volume")
    bad_o = np.zeros(len(pts), bool)
    bad_g = np.zeros(len(pts), bool)
    for name in fields:
        bad_o |= ~np.isfinite(cases.oracle_particle_field(ops, name)).all(axis=1)
        bad_g |= ~np.isfinite(cases.gpu_particle_field(ps, name)).all(axis=1)
    assert bad_o[[3, 4]].all() and not bad_o[2] and bad_o.sum() < len(pts) // 2
    assert np.array_equal(bad_g, bad_o), (np.nonzero(bad_g)[0], np.nonzero(bad_o)[0])
    ok = ~bad_o
    for name in fields:
        err = cases.normwise(cases.gpu_particle_field(ps, name)[ok],
                             cases.oracle_particle_field(ops, name)[ok])
        assert err < RTOL_F32, (name, err)


def _moving_spec(scheme, per_particle_E=False, nx=200, ny=40, rows=30, drift=(0.6, 0.3),
                 velocity_scale=0.05):
    """Multi-tile 2D case whose particles cross cells within a few steps
    (exercises re-sorting in the fused path)."""
    spec = cases.synthetic_spec("quad4", nx=nx, ny=ny, rows=rows, seed=21, scheme=scheme,
                                traction=False, nodal_force=True, E=1e3, dt=1e-3,
                                velocity_scale=velocity_scale, nsteps=12,
                                per_particle_E=per_particle_E, drift=drift)
    spec["nodal_forces"][0]["force"] = -0.05
    right = [j * (nx + 1) + nx for j in range(ny + 1)]
    spec["constraints"].append({"nodes": right, "dir": 0, "velocity": 0.0})
    # keep a 3-cell margin between the particles and the mesh edges
    h = spec["el_len"][0]
    pts = np.asarray(spec["particles"]) * np.array([(nx - 6) / nx, 1.0]) + 3 * h
    spec["particles"] = pts.tolist()
    return spec


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("scheme,per_E,drift", [("usf", False, (0.6, 0.3)),
                                                 ("usl", False, (0.6, 0.3)),
                                                 ("usf", True, (0.6, 0.3)),
                                                 ("usf", False, None), ("usl", True, None)])
def test_tile_path_equals_per_phase_path(scheme, per_E, drift, split):
    """The fused step (dmpm_run default: re-sort, P2G cell and node passes,
    G2P) and the per-phase kernels (dmpm_run flag bit 2) give bit-identical
    states, across cell changes: with a drift every particle changes cell in
    every step, without it a few hundred do (incremental re-sorts either
    way)."""
    from diffmpm_amd._plan import plan_for
    spec = _moving_spec(scheme, per_E, drift=drift, velocity_scale=0.05 if drift else 0.1)
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    import diffmpm_amd._lib as L
    pa.set_option(L.DMPM_OPT_USF_SPLIT, 1 if split else 0)
    e0 = a.mesh.particles[0].element_ids.clone()
    for k in range(0, spec["nsteps"], 3):
        pa.run(scheme, False, g, spec["dt"], k, 3, True, per_phase=False)
        pb.run(scheme, False, g, spec["dt"], k, 3, True, per_phase=True)
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), k
    assert not torch.equal(e0, a.mesh.particles[0].element_ids)  # cells did change


@pytest.mark.parametrize("cells_per_block", [64, 256])
def test_block_size_does_not_change_results(cells_per_block):
    from diffmpm_amd._plan import plan_for
    spec = _moving_spec("usf", nx=150, ny=70, rows=60, drift=None, velocity_scale=0.1)
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    pb.set_block(cells_per_block)
    pa.run("usf", False, g, spec["dt"], 0, 6, True)
    pb.run("usf", False, g, spec["dt"], 0, 6, True)
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_tile_path_per_step_parity(scheme, dtype):
    """The fused path, re-synchronised to the oracle before every step of
    the multi-tile moving case, matches it within RTOL_F32 / RTOL_F64."""
    from diffmpm_amd._plan import plan_for
    spec = _moving_spec(scheme)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    rtol = RTOL_F64 if dtype == torch.float64 else RTOL_F32
    osim = cases.build_oracle(spec, npdt)
    gsim = cases.build_gpu(spec, dtype)
    plan = plan_for(gsim.mesh.elements, gsim.mesh.particles[0])
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    for i in range(spec["nsteps"]):
        cases.load_oracle_state_into_gpu(osim, gsim)
        osim.step(i, np.asarray(spec["gravity"]))
        plan.run(scheme, False, g, spec["dt"], i, 1, True)
        _compare(osim, gsim, rtol, f"step {i}")


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_tile_path_particles_below_and_left_of_mesh(scheme):
    """Points below or left of the mesh belong to element 0 (A1 quirk,
    element.py:827-837); the fused path treats them like the per-phase path."""
    from diffmpm_amd._plan import plan_for
    spec = cases.synthetic_spec("quad4", nx=150, ny=20, rows=10, seed=4, scheme=scheme,
                                traction=False, E=1e3, dt=1e-3, velocity_scale=0.05, nsteps=4)
    pts = np.asarray(spec["particles"])
    pts[:40, 1] = -0.002      # below the mesh, spread over x
    pts[40:60, 0] = -0.003    # left of the mesh
    spec["particles"] = pts.tolist()
    spec["constraints"] = []
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    assert (a.mesh.particles[0].element_ids[:60] == 0).all()
    pa.run(scheme, False, g, spec["dt"], 0, 4, True, per_phase=False)
    pb.run(scheme, False, g, spec["dt"], 0, 4, True, per_phase=True)
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_tile_path_many_arrivals_into_one_cell(scheme):
    """A sheet of 200 particles just left of a cell edge, moving right fast
    enough to cross it in one step: more arrivals into one cell than the
    incremental re-sort takes (full re-sort), then incremental steps; the
    fused path equals the per-phase path bit for bit."""
    from diffmpm_amd._plan import plan_for
    h, dt = 0.1, 1e-3
    rng = np.random.Generator(np.random.PCG64(8))
    sheet = np.stack([np.full(200, 0.4 - 0.0005), 0.31 + 0.08 * rng.random(200)], 1)
    bg = cases.synthetic_spec("quad4", nx=8, ny=8, rows=6, seed=4)["particles"]
    pts = np.concatenate([np.asarray(bg) * 0.8 + 0.1, sheet])
    vel = np.zeros_like(pts)
    vel[-200:, 0] = 0.02 * h / dt
    spec = dict(element="quad4", nelements=[10, 10], el_len=[h, h], particles=pts.tolist(),
                init_velocity=vel.tolist(),
                material={"type": "LinearElastic", "density": 1000.0, "youngs_modulus": 1e2,
                          "poisson_ratio": 0.3},
                constraints=[{"nodes": list(range(11)), "dir": 1, "velocity": 0.0}],
                nodal_forces=[], tractions=[], gravity=[0.0, -9.81], scheme=scheme, dt=dt,
                nsteps=4, velocity_update=True)
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    e0 = a.mesh.particles[0].element_ids.clone()
    pa.run(scheme, True, g, dt, 0, 1, True, per_phase=False)
    pb.run(scheme, True, g, dt, 0, 1, True, per_phase=True)
    for k in range(1, 4):
        pa.run(scheme, True, g, dt, k, 1, True, per_phase=False)
        pb.run(scheme, True, g, dt, k, 1, True, per_phase=True)
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), k
    moved = int((e0 != a.mesh.particles[0].element_ids).sum())
    assert moved > 64  # the sheet crossed into the next cell


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_incremental_resort_is_the_stable_counting_sort(scheme):
    """After each fused step of the multi-tile moving case, the sort by cell
    that the step used (exported: slot -> particle, first slot of each cell)
    equals the stable counting sort of the oracle's element ids of the
    step-start positions: cells in order, ascending particle index within a
    cell, particles outside the mesh last."""
    from diffmpm_amd._plan import plan_for
    from diffmpm_amd._lib import call, ptr
    spec = _moving_spec(scheme)
    sim = cases.build_gpu(spec)
    el = cases.build_oracle(spec, np.float32).mesh.elements
    ps, nodes = sim.mesh.particles[0], sim.mesh.elements.nodes
    plan = plan_for(sim.mesh.elements, ps)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    ncells = spec["nelements"][0] * spec["nelements"][1]
    n, dev = len(ps), ps.loc.device
    keys = []
    for k in range(spec["nsteps"]):
        loc = ps.loc.detach().cpu().numpy().astype(np.float32).reshape(n, 2)
        e = np.asarray(el.element_ids(loc))
        key = np.where((e >= 0) & (e < ncells), e, ncells)
        keys.append(key)
        plan.run(scheme, False, g, spec["dt"], k, 1, True)
        vkin = torch.empty((2, nodes._buf.shape[1]), dtype=nodes._buf.dtype, device=dev)
        perm = torch.empty(n, dtype=torch.int32, device=dev)
        cs = torch.empty(ncells + 2, dtype=torch.int32, device=dev)
        call("dmpm_plan_export_step", plan.h, ptr(vkin), ptr(perm), ptr(cs), plan.s)
        torch.cuda.synchronize()
        want_cs = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=ncells + 1))])
        assert np.array_equal(perm.cpu().numpy(), np.argsort(key, kind="stable")), k
        assert np.array_equal(cs.cpu().numpy()[:ncells + 2], want_cs), k
    moved = sum(int((a != b).sum()) for a, b in zip(keys, keys[1:]))
    assert moved > 50  # incremental re-sorts with cell changes


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_tile_path_crowded_cell_equals_per_phase_path(scheme):
    """3000 particles in one cell (rank-sorted by one workgroup in the full
    re-sort): the fused path equals the per-phase path bit for bit."""
    from diffmpm_amd._plan import plan_for
    spec = _crowded_spec(3000)
    spec["scheme"] = scheme
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    pa.run(scheme, False, g, spec["dt"], 0, 3, True, per_phase=False)
    pb.run(scheme, False, g, spec["dt"], 0, 3, True, per_phase=True)
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)


def _run_pair(spec, scheme, nsteps, chunk, vu=False, per_phase_b=False, kind=1):
    """Runs spec on two fresh simulations: a with the one-launch step (kind
    1) or the split step (kind 2), b with one launch per phase (or the
    per-phase kernels); compares after every chunk of steps."""
    from diffmpm_amd import _lib as L
    from diffmpm_amd._plan import plan_for
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    pa.set_option(L.DMPM_OPT_STEP_KERNEL, kind)
    pb.set_option(L.DMPM_OPT_STEP_KERNEL, 0)
    for k in range(0, nsteps, chunk):
        pa.run(scheme, vu, g, spec["dt"], k, chunk, True)
        pb.run(scheme, vu, g, spec["dt"], k, chunk, True, per_phase=per_phase_b)
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), k
    return a


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("chunk", [1, 3, 4])
@pytest.mark.parametrize("scheme,per_E,drift", [("usf", False, (0.6, 0.3)),
                                                 ("usl", False, (0.6, 0.3)),
                                                 ("usf", True, None), ("usl", True, (0.6, 0.3))])
def test_one_launch_step_equals_phase_launches(scheme, per_E, drift, chunk, kind):
    """The one-launch step (tiles of cells with halos, alternating
    particle blocks) and the split step (three launches over node and cell
    tiles) give bit-identical states to one launch per phase, for odd and
    even step counts per run (the state ends in the caller's block either
    way), across cell changes, on a mesh of partial tiles (cases that stay
    finite in the oracle too)."""
    spec = _moving_spec(scheme, per_E, drift=drift, velocity_scale=0.05 if drift else 0.1)
    a = _run_pair(spec, scheme, 12, chunk, kind=kind)
    assert torch.isfinite(a.mesh.particles[0].loc).all()


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_one_launch_step_large_mesh(scheme, dtype, kind):
    """Several full tiles in each direction (96x70 cells), nodal force and
    velocity update: the one-launch and the split step equal the per-phase
    kernels."""
    spec = _moving_spec(scheme, nx=96, ny=70, rows=60, drift=None, velocity_scale=0.1)
    spec["velocity_update"] = True
    from diffmpm_amd import _lib as L
    from diffmpm_amd._plan import plan_for
    a, b = cases.build_gpu(spec, dtype), cases.build_gpu(spec, dtype)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    pa.set_option(L.DMPM_OPT_STEP_KERNEL, kind)
    pa.run(scheme, True, g, spec["dt"], 0, 5, True)
    pb.run(scheme, True, g, spec["dt"], 0, 5, True, per_phase=True)
    for x, y in zip(_state(a), _state(b)):
        assert torch.equal(x, y)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_one_launch_step_crowded_and_outside(scheme, kind):
    """A 3000-particle cell (several staging rounds in one tile) and
    particles outside the mesh (NaN state, handled by the tail workgroups)."""
    spec = _crowded_spec(3000)
    spec["scheme"] = scheme
    pts = np.asarray(spec["particles"])
    pts[:3] = [[1.2, 0.3], [np.nan, 0.4], [0.3, 1.5]]
    spec["particles"] = pts.tolist()
    _run_pair(spec, scheme, 3, 3, kind=kind)


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_step_between_blocks_equals_run(scheme, split):
    """dmpm_step_between alternating two blocks (chained after the first
    step) gives the states of dmpm_run stepping one block in place, bit for
    bit, and leaves its input block unchanged: with a copy of the state rows
    (split=False), and with the state rows read from and written to state
    tensors instead of the blocks (split=True)."""
    from diffmpm_amd import _lib
    from diffmpm_amd._lib import call, ptr
    from diffmpm_amd._plan import Plan, plan_for
    spec = _moving_spec(scheme, drift=(0.6, 0.3))
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    psb = b.mesh.particles[0]
    pb = Plan(b.mesh.elements, psb)
    blocks = [psb._buf.clone(), psb._buf.clone()]
    eid = psb._eid.clone()
    code = _lib.DMPM_USF if scheme == "usf" else _lib.DMPM_USL
    n = len(psb)
    state_rows = [_lib.P_X, _lib.P_X + 1, _lib.P_V, _lib.P_V + 1, _lib.P_STRESS,
                  _lib.P_STRESS + 1, _lib.P_STRESS + 3, _lib.P_VOLUME]
    other_rows = [_lib.P_DENSITY, _lib.P_STRAIN, _lib.P_STRAIN + 1, _lib.P_STRAIN + 3,
                  _lib.P_STRESS + 2, _lib.P_VOLSTRAIN_C]
    sin = torch.stack([psb._buf[r, :n] for r in state_rows]).contiguous()
    for k in range(7):
        src, dst = blocks[k % 2], blocks[1 - k % 2]
        before, sin_before = src.clone(), sin.clone()
        pa.run(scheme, False, g, spec["dt"], k, 1, False)
        sout = torch.full((8, n), float("nan"), dtype=src.dtype, device=src.device)
        call("dmpm_step_between", pb.h, ptr(src), ptr(dst), ptr(eid),
             ptr(b.mesh.elements.nodes._buf), code, 0, pb._gravity(g), float(spec["dt"]), k,
             2 if k else 0, ptr(sin) if split else None, ptr(sout), n, pb.s)
        assert torch.equal(src, before) and torch.equal(sin, sin_before)
        ref = a.mesh.particles[0]._buf
        for r in other_rows + ([] if split else state_rows):
            assert torch.equal(dst[r, :n].view(torch.int32), ref[r, :n].view(torch.int32)), (k, r)
        for q, r in enumerate(state_rows):
            assert torch.equal(sout[q].view(torch.int32), ref[r, :n].view(torch.int32)), (k, q)
        assert torch.equal(eid[:n], a.mesh.particles[0]._eid[:n])
        na, nb = a.mesh.elements.nodes, b.mesh.elements.nodes
        m = na._count
        assert torch.equal(na._buf[:, :m].view(torch.int32), nb._buf[:, :m].view(torch.int32))
        sin = sout


@pytest.mark.parametrize("driver", ["method", "phase_kernels", "phase_launches", "step",
                                    "split"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_nodal_damping_force_per_step_parity(scheme, dtype, driver):
    """A nonzero nodal f_damp enters the total nodal force (node.py:123-125,
    element.py:305-329) and is never reset by the step (the reference calls
    reset_values nowhere), through each driver of the step: the scheme's
    per-method calls, the per-phase kernels, the per-phase launches of the
    tile kernels and the one-launch step.  Re-synchronised to the oracle
    before every step, like test_per_step_parity.  f_damp is set on the
    nodes the particle block covers for the whole run (a massless node under
    a force takes nan_to_num(F / 0) = the largest float as its acceleration,
    in the reference too)."""
    from diffmpm_amd import _lib as L
    from diffmpm_amd._plan import plan_for
    spec = _moving_spec(scheme)
    spec["nsteps"] = 6
    npdt = np.float64 if dtype == torch.float64 else np.float32
    rtol = RTOL_F64 if dtype == torch.float64 else RTOL_F32
    osim = cases.build_oracle(spec, npdt)
    gsim = cases.build_gpu(spec, dtype)
    nx, ny = spec["nelements"]
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    covered = ((ii >= 6) & (ii <= nx - 10) & (jj >= 5) & (jj <= 31)).reshape(-1)
    rng = np.random.Generator(np.random.PCG64(5))
    fd = 0.1 * rng.standard_normal((covered.size, 2)) * covered[:, None]
    on, gn = osim.mesh.elements.nodes, gsim.mesh.elements.nodes
    on.f_damp = fd.reshape(on.f_damp.shape).astype(npdt)
    gn.f_damp = torch.from_numpy(fd.reshape(gn.f_damp.shape))
    expect_fd = gn.f_damp.clone()
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    plan = plan_for(gsim.mesh.elements, gsim.mesh.particles[0])
    plan.set_option(L.DMPM_OPT_STEP_KERNEL, {"step": 1, "split": 2}.get(driver, 0))
    for i in range(spec["nsteps"]):
        cases.load_oracle_state_into_gpu(osim, gsim)
        osim.step(i, np.asarray(spec["gravity"]))
        if driver == "method":
            gsim._step(i, g)
        else:
            plan.run(scheme, False, g, spec["dt"], i, 1, True,
                     per_phase=driver == "phase_kernels")
        _compare(osim, gsim, rtol, f"step {i}")
        assert torch.equal(gn.f_damp, expect_fd), i
    # the damping force changed the particles' motion
    ref = cases.build_oracle(spec, npdt)
    for i in range(spec["nsteps"]):
        ref.step(i, np.asarray(spec["gravity"]))
    assert cases.normwise(osim.mesh.particles[0].velocity,
                          ref.mesh.particles[0].velocity) > 1e-3


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_apply_force_boundary_constraints(dtype):
    """element.py:336-340: the nodes of the first constraint get zero f_int,
    f_ext and f_damp; every other node and field is left as it was."""
    spec = cases.synthetic_spec(nsteps=1)
    gsim = cases.build_gpu(spec, dtype)
    el = gsim.mesh.elements
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    gsim._step(0, g)
    n = el.nodes
    rng = np.random.Generator(np.random.PCG64(2))
    n.f_damp = torch.from_numpy(rng.standard_normal(tuple(n.f_damp.shape)))
    before = n._buf.clone()
    el.apply_force_boundary_constraints()
    from diffmpm_amd import _lib as L
    ids = torch.as_tensor(el.constraints[0][0], device=before.device).long()
    expect = before.clone()
    for r in (L.N_FINT, L.N_FEXT, L.N_FDAMP):
        for d in range(2):
            expect[r + d, ids] = 0
    assert torch.equal(n._buf, expect)
    assert torch.count_nonzero(before[L.N_FINT:L.N_ROWS][:, ids]) > 0
