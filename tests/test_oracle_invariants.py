"""Closed forms, invariants and brute force that pin the oracle where the
reference's own tests leave it open (including the Linear1D path, which the
reference never tests, and element search outside the mesh)."""
import math

import numpy as np
import pytest

import cases
from oracle import mpm as om


def rng(seed=0):
    return np.random.Generator(np.random.PCG64(seed))


@pytest.mark.parametrize("h", [(0.1, 0.1), (0.25, 0.5), (1.0, 1.0)])
def test_element_search_matches_bruteforce_q4(h):
    """Closed form of element.py:820-840 vs the literal nonzero/isin
    construction, including points on node lines, outside the mesh and NaN."""
    el = om.Elements("quad4", (3, 2), h, dtype=np.float32)
    X = el.nodes.loc[:4, 0]
    Y = el.nodes.loc[::4, 1]
    pts = [[x, y] for x in list(X) + [-0.01, 0.05, 3.5 * h[0]] for y in list(Y) + [-1.0, 0.03]]
    r = rng(1).uniform(-0.5, 1.5, size=(200, 2)) * np.array([3 * h[0], 2 * h[1]])
    pts = np.concatenate([np.asarray(pts), r, [[np.nan, 0.05], [0.05, np.nan]]]).astype(np.float32)
    assert np.array_equal(el.element_ids(pts), el.element_ids_bruteforce(pts))


def test_element_search_matches_bruteforce_1d():
    el = om.Elements("linear1d", 5, 0.1, dtype=np.float32)
    X = el.nodes.loc[:, 0]
    pts = np.concatenate([X, [-0.1, 0.7, np.nan], rng(2).uniform(-0.2, 0.7, 100)])
    pts = pts.astype(np.float32).reshape(-1, 1)
    assert np.array_equal(el.element_ids(pts), el.element_ids_bruteforce(pts))


def test_element_search_quirks():
    """Documented A1 behaviour: x >= X_nx wraps into the next row, points
    left of/below the mesh map to element 0, y >= Y_ny gives ids >= nx*ny."""
    el = om.Elements("quad4", (3, 2), (1.0, 1.0), dtype=np.float32)
    ids = el.element_ids(np.array([[3.2, 0.5], [-1, 1.5], [1.5, -2], [1.5, 2.5]], np.float32))
    assert ids.tolist() == [3, 0, 0, 7]


@pytest.mark.parametrize("kind", ["quad4", "linear1d"])
def test_partition_of_unity_and_gradient_sums(kind):
    el = om.Elements(kind, (2, 2) if kind == "quad4" else 2, 0.5, dtype=np.float64)
    nd = el.ndim
    xi = rng(3).uniform(-1, 1, size=(50, nd))
    assert np.allclose(el.shapefn(xi).sum(axis=1), 1.0, atol=1e-15)
    assert np.allclose(el.shapefn_natural_grad(xi).sum(axis=1), 0.0, atol=1e-15)
    coords = om.gather(el.nodes.loc, el.id_to_node_ids(np.zeros(50, int)))
    dN = el.shapefn_grad(xi, coords)
    assert np.allclose(dN.sum(axis=1), 0.0, atol=1e-12)
    # dN/dx = dN/dxi * 2/h on the axis-aligned cell
    assert np.allclose(dN, el.shapefn_natural_grad(xi) * 2 / 0.5, atol=1e-12)


@pytest.mark.parametrize("element", ["quad4", "linear1d"])
def test_p2g_conserves_mass_and_momentum(element):
    spec = cases.synthetic_spec(element, nx=8, ny=6, rows=5, seed=4, nsteps=1)
    sim = cases.build_oracle(spec, np.float64)
    el, ps = sim.mesh.elements, sim.mesh.particles[0]
    ps.set_particle_element_ids(el)
    ps.update_natural_coords(el)
    el.compute_nodal_mass(ps)
    assert math.isclose(el.nodes.mass.sum(), ps.mass.sum(), rel_tol=1e-13)
    el.compute_nodal_momentum(ps)
    want = (ps.mass[:, None] * ps.velocity).sum(axis=0)
    assert np.allclose(el.nodes.momentum.sum(axis=0), want, rtol=1e-12, atol=1e-12)
    # body force sums to total weight
    el.nodes.f_ext[:] = 0
    g = np.asarray(spec["gravity"])
    el.compute_body_force(ps, g)
    assert np.allclose(el.nodes.f_ext.sum(axis=0), ps.mass.sum() * g, rtol=1e-12)
    # internal forces are self-equilibrated for any stress
    ps.stress = rng(5).standard_normal(ps.stress.shape)
    ps.compute_volume(el)
    el.compute_internal_force(ps)
    assert np.allclose(el.nodes.f_int.sum(axis=0), 0.0, atol=1e-9)


def test_linear_velocity_field_gives_exact_strain_rate_q4():
    """A bilinear interpolant reproduces v = A x + b exactly, so the strain
    rate is [A00, A11, 0, A01 + A10, 0, 0] at every particle and centroid."""
    spec = cases.synthetic_spec("quad4", nx=5, ny=4, rows=4, seed=6)
    sim = cases.build_oracle(spec, np.float64)
    el, ps = sim.mesh.elements, sim.mesh.particles[0]
    A = np.array([[0.3, -0.7], [1.1, 0.2]])
    b = np.array([0.05, -0.02])
    el.nodes.velocity = el.nodes.loc @ A.T + b
    ps.set_particle_element_ids(el)
    ps.update_natural_coords(el)
    ps.compute_strain(el, 0.5)
    want = np.array([A[0, 0], A[1, 1], 0, A[0, 1] + A[1, 0], 0, 0])
    assert np.allclose(ps.strain_rate, want[None, :], atol=1e-12)
    assert np.allclose(ps.dvolumetric_strain, 0.5 * (A[0, 0] + A[1, 1]), atol=1e-12)
    # G2P of the same field is exact too
    el.nodes.acceleration[:] = 0
    x0 = ps.loc.copy()
    ps.update_position_velocity(el, 0.1, True)
    assert np.allclose(ps.velocity, x0 @ A.T + b, atol=1e-12)


def test_linear_velocity_field_1d_quirk():
    """1D: L = a for v = a x, and the clamped indices of particle.py:302-305
    give strain_rate = [a, a, 0, 2a, 0, 0]."""
    el = om.Elements("linear1d", 6, 0.2, dtype=np.float64)
    ps = om.Particles(rng(7).uniform(0, 1.19, size=(20, 1)),
                      om.LinearElastic({"density": 1, "youngs_modulus": 1, "poisson_ratio": 0}),
                      dtype=np.float64)
    el.nodes.velocity = 0.4 * el.nodes.loc
    ps.set_particle_element_ids(el)
    ps.update_natural_coords(el)
    ps.compute_strain(el, 1.0)
    assert np.allclose(ps.strain_rate, [[0.4, 0.4, 0, 0.8, 0, 0]], atol=1e-12)
    assert np.allclose(ps.dvolumetric_strain, 0.4, atol=1e-12)


def test_simple_material_per_particle_modulus():
    """SimpleMaterial with a per-particle E: dsigma_p = E_p dstrain_p
    (material.py:130-131 with E_p in place of the scalar E).  Particles
    holding E_p = E or 2E step exactly as the scalar E or 2E material."""
    def run(E, youngs=None):
        el = om.Elements("linear1d", 8, 0.125, [(np.array([0]), om.Constraint(0, 0.0))],
                         dtype=np.float64)
        ps = om.Particles(np.linspace(0.05, 0.95, 16)[:, None],
                          om.SimpleMaterial({"E": E, "density": 1.0}), dtype=np.float64)
        ps.velocity[:, 0] = 0.1 * np.sin(np.pi * ps.loc[:, 0] / 2)
        if youngs is not None:
            ps.youngs_modulus = youngs
        sim = om.MPMExplicit(om.Mesh(el, [ps]), 1e-3, "usf", False, 5)
        sim.solve(np.array([0.0]))
        return ps.stress.copy(), ps.velocity.copy()
    for scale in (1.0, 2.0):
        got = run(100.0, np.full(16, 100.0 * scale))
        want = run(100.0 * scale)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # mixed: each particle's increment uses its own E
    E = np.where(np.arange(16) % 2 == 0, 100.0, 300.0)
    el = om.Elements("linear1d", 8, 0.125, dtype=np.float64)
    ps = om.Particles(np.linspace(0.05, 0.95, 16)[:, None],
                      om.SimpleMaterial({"E": 7.0, "density": 1.0}), dtype=np.float64)
    ps.youngs_modulus = E
    ps.dstrain = rng(3).standard_normal((16, 6))
    ps.compute_stress()
    assert np.allclose(ps.stress, ps.dstrain * E[:, None], rtol=0, atol=0)


@pytest.mark.parametrize("nu", [0.0, 0.25, 0.3, 0.45])
def test_linear_elastic_closed_form(nu):
    E = 2.5e6
    mat = om.LinearElastic({"density": 1, "youngs_modulus": E, "poisson_ratio": nu})
    e = 1e-4
    s = mat.compute_stress(np.array([[e, 0, 0, 0, 0, 0]], np.float64))[0]
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    assert np.allclose(s, [(lam + 2 * mu) * e, lam * e, lam * e, 0, 0, 0], rtol=1e-13)
    s = mat.compute_stress(np.array([[0, 0, 0, e, 0, 0]], np.float64))[0]
    assert np.allclose(s, [0, 0, 0, mu * e, 0, 0], rtol=1e-13)


def test_linear_function_is_np_interp():
    xs, fs = [0.0, 0.5, 1.0, 1.0, 2.0], [0.0, 1.0, 2.0, 3.0, -1.0]
    f = om.Linear(0, xs, fs)
    for t in np.linspace(-1, 3, 81):
        assert math.isclose(f.value(t, np.float64), np.interp(t, xs, fs), abs_tol=1e-15)


@pytest.mark.parametrize("scheme", ["usf", "usl"])
@pytest.mark.parametrize("element", ["quad4", "linear1d"])
def test_fp32_tracks_fp64(scheme, element):
    spec = cases.synthetic_spec(element, scheme=scheme, nsteps=20, seed=8)
    a = cases.build_oracle(spec, np.float32)
    b = cases.build_oracle(spec, np.float64)
    a.solve(np.asarray(spec["gravity"]))
    b.solve(np.asarray(spec["gravity"]))
    for name in ("loc", "velocity", "stress", "strain", "volume"):
        err = cases.normwise(cases.oracle_particle_field(a.mesh.particles[0], name),
                             cases.oracle_particle_field(b.mesh.particles[0], name))
        assert err < 1e-5, (name, err)


def test_nan_to_num_and_empty_node_velocity():
    """element.py:191-202 keeps the old velocity where m == 0 and
    element.py:311 maps F/0 to 0 (F = 0) or +-max (F != 0)."""
    el = om.Elements("quad4", (2, 1), 1.0, dtype=np.float32)
    el.nodes.mass[:] = 0
    el.nodes.momentum[:] = 1
    el.nodes.velocity[:] = 0.25
    el.compute_velocity(None)
    assert np.all(el.nodes.velocity == np.float32(0.25))
    el.nodes.f_ext[:] = 0
    el.nodes.f_ext[0, 0] = 2.0
    el.nodes.f_ext[1, 1] = -2.0
    el.update_nodal_acceleration_velocity(None, 1.0)
    a = el.nodes.acceleration
    fmax = np.finfo(np.float32).max
    assert a[0, 0] == fmax and a[1, 1] == -fmax and a[2, 0] == 0


# ---------------------------------------------------------------------------
# Linear completeness: the shape functions reproduce linear fields exactly, so
# interpolating the node coordinates returns the particle position, and the
# first moments of P2G equal the particle sums.  These pin the natural
# coordinate (sign, centre and scale; for the 2-node element the reference's
# index-2 clamp to node 1, particle.py:180-206) and every transfer that uses
# it, in 1D as in 2D.
# ---------------------------------------------------------------------------
def _setup(element, dtype=np.float64, seed=11):
    spec = cases.synthetic_spec(element, nx=9, ny=7, rows=6, seed=seed, jitter=0.4)
    sim = cases.build_oracle(spec, dtype)
    el, ps = sim.mesh.elements, sim.mesh.particles[0]
    ps.set_particle_element_ids(el)
    ps.update_natural_coords(el)
    return el, ps


def completeness_error(el, ps):
    """max |sum_i X_i N_i(xi_p) - x_p| over particles and axes."""
    N = el.shapefn(ps.reference_loc)
    X = om.gather(el.nodes.loc, el.id_to_node_ids(ps.element_ids))
    return float(np.abs(np.einsum("nk,nka->na", N, X) - ps.loc).max())


@pytest.mark.parametrize("element", ["linear1d", "quad4"])
def test_natural_coords_reproduce_positions(element):
    el, ps = _setup(element)
    assert np.all((ps.element_ids >= 0) & (ps.element_ids < el.total_elements))
    assert np.all(np.abs(ps.reference_loc) <= 1 + 1e-12)
    assert completeness_error(el, ps) < 1e-12


@pytest.mark.parametrize("element", ["linear1d", "quad4"])
@pytest.mark.parametrize("velocity_update", [True, False])
def test_g2p_of_linear_nodal_fields(element, velocity_update):
    """particle.py:208-247 with nodal v = A X + b and a = C X + d: v_p = A x_p
    + b (velocity_update) or v_p0 + dt (C x_p + d), and x_p += dt (A x_p + b)."""
    el, ps = _setup(element, seed=12)
    nd = el.ndim
    r = rng(13)
    A, C = r.standard_normal((nd, nd)), r.standard_normal((nd, nd))
    b, dd = r.standard_normal(nd), r.standard_normal(nd)
    el.nodes.velocity = el.nodes.loc @ A.T + b
    el.nodes.acceleration = el.nodes.loc @ C.T + dd
    x0, v0 = ps.loc.copy(), ps.velocity.copy()
    dt = 0.125
    ps.update_position_velocity(el, dt, velocity_update)
    want_v = x0 @ A.T + b if velocity_update else v0 + dt * (x0 @ C.T + dd)
    assert np.allclose(ps.velocity, want_v, rtol=0, atol=1e-12)
    assert np.allclose(ps.loc, x0 + dt * (x0 @ A.T + b), rtol=0, atol=1e-12)


@pytest.mark.parametrize("element", ["linear1d", "quad4"])
def test_p2g_first_moments(element):
    """sum_i X_i m_i = sum_p m_p x_p (element.py:122-152), likewise for the
    momentum (element.py:154-189) and body force (element.py:236-268)."""
    el, ps = _setup(element, seed=14)
    el.compute_nodal_mass(ps)
    X = el.nodes.loc
    assert np.allclose(X.T @ el.nodes.mass, ps.loc.T @ ps.mass, rtol=1e-12, atol=0)
    el.compute_nodal_momentum(ps)
    mv = ps.mass[:, None] * ps.velocity
    assert np.allclose(X.T @ el.nodes.momentum, ps.loc.T @ mv, rtol=1e-11, atol=1e-12)
    el.nodes.f_ext[:] = 0
    g = np.arange(1, el.ndim + 1) * -2.5
    el.compute_body_force(ps, g)
    want = (ps.loc.T @ ps.mass)[:, None] * g[None, :]
    assert np.allclose(X.T @ el.nodes.f_ext, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("element", ["linear1d", "quad4"])
def test_internal_force_virial(element):
    """sum_i X_i,a f_int,i,b = -sum_p V_p sigma_ab (element.py:548-600 in 1D:
    sigma_xx only; element.py:842-901 in 2D), from sum_i X_i,a dN_i/dx_c =
    delta_ac.  Pins the sign of f_int and the 1D/2D gradient chain."""
    el, ps = _setup(element, seed=15)
    ps.compute_volume(el)
    ps.stress = rng(16).standard_normal(ps.stress.shape)
    el.compute_internal_force(ps)
    M = el.nodes.loc.T @ el.nodes.f_int
    s, V = ps.stress, ps.volume
    if el.ndim == 1:
        want = np.array([[-(V * s[:, 0]).sum()]])
    else:
        want = -np.array([[(V * s[:, 0]).sum(), (V * s[:, 3]).sum()],
                          [(V * s[:, 3]).sum(), (V * s[:, 1]).sum()]])
    assert np.allclose(M, want, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("mistake", ["negate", "scale", "node0"])
def test_planted_natural_coordinate_mistakes_are_caught(monkeypatch, mistake):
    """Plants a wrong 1D natural coordinate in the oracle (a negated or
    rescaled xi, or the far node taken as node 0 instead of the clamped node
    1) and checks that the completeness checks above fail on it."""
    orig = om.Particles.update_natural_coords

    def wrong(self, el):
        orig(self, el)
        if el.kind != "linear1d":
            return
        if mistake == "negate":
            self.reference_loc = -self.reference_loc
        elif mistake == "scale":
            self.reference_loc = 2 * self.reference_loc
        else:
            X0 = om.gather(el.nodes.loc, self.element_ids)
            self.reference_loc = (self.loc - X0) * 2 / el.hx

    monkeypatch.setattr(om.Particles, "update_natural_coords", wrong)
    el, ps = _setup("linear1d")
    assert completeness_error(el, ps) > 1e-3
    el.compute_nodal_mass(ps)
    assert not np.allclose(el.nodes.loc.T @ el.nodes.mass, ps.loc.T @ ps.mass, rtol=1e-9)
