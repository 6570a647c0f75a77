"""Linear completeness through the HIP per-method path (Linear1D and Q4,
FP32 and FP64): the natural coordinates reproduce the positions, G2P of
linear nodal fields is exact, and the first moments of P2G and the virial of
the internal force equal their particle sums.  These are the checks that pin
the oracle's Linear1D natural coordinate (a negated or rescaled xi fails
them); here they pin the kernels the same way, independently of the oracle."""
import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu
DTYPES = [torch.float32, torch.float64]
# FP32: the checks compare sums of O(1)-sized terms, exact up to rounding
TOL = {torch.float32: 2e-5, torch.float64: 1e-11}


def setup(element, dtype, seed=11):
    spec = cases.synthetic_spec(element, nx=9, ny=7, rows=6, seed=seed, jitter=0.4)
    sim = cases.build_gpu(spec, dtype)
    el, ps = sim.mesh.elements, sim.mesh.particles[0]
    el.set_particle_element_ids(ps)
    ps.update_natural_coords(el)
    torch.cuda.synchronize()
    return el, ps


def host(t):
    t = t.detach().double().cpu().numpy()
    return t.reshape(t.shape[0], -1)


def node_ids(element, eid, nx):
    eid = eid.astype(np.int64)
    if element == "linear1d":
        return np.stack([eid, eid + 1], axis=1)
    ll = (eid // nx) * (nx + 1) + eid % nx
    return np.stack([ll, ll + 1, ll + nx + 2, ll + nx + 1], axis=1)


def shapefn(xi):
    if xi.shape[1] == 1:
        x = xi[:, 0]
        return np.stack([0.5 * (1 - x), 0.5 * (1 + x)], axis=1)
    x, y = xi[:, 0], xi[:, 1]
    return 0.25 * np.stack([(1 - x) * (1 - y), (1 + x) * (1 - y), (1 + x) * (1 + y),
                            (1 - x) * (1 + y)], axis=1)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("element", ["linear1d", "quad4"])
def test_natural_coords_reproduce_positions(element, dtype):
    el, ps = setup(element, dtype)
    xi, x = host(ps.reference_loc), host(ps.loc)
    eid = ps.element_ids.cpu().numpy()
    assert np.all((eid >= 0) & (eid < el.total_elements))
    X = host(el.nodes.loc)[node_ids(element, eid, el._nx)]
    got = np.einsum("nk,nka->na", shapefn(xi), X)
    assert np.abs(got - x).max() < TOL[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("velocity_update", [True, False])
@pytest.mark.parametrize("element", ["linear1d", "quad4"])
def test_g2p_of_linear_nodal_fields(element, velocity_update, dtype):
    el, ps = setup(element, dtype, seed=12)
    nd = el.ndim
    r = np.random.Generator(np.random.PCG64(13))
    A, C = r.standard_normal((nd, nd)), r.standard_normal((nd, nd))
    b, d = r.standard_normal(nd), r.standard_normal(nd)
    X = host(el.nodes.loc)
    shape = el.nodes.velocity.shape
    el.nodes.velocity = torch.from_numpy((X @ A.T + b).reshape(shape))
    el.nodes.acceleration = torch.from_numpy((X @ C.T + d).reshape(shape))
    x0, v0 = host(ps.loc), host(ps.velocity)
    dt = 0.125
    ps.update_position_velocity(el, dt, velocity_update)
    torch.cuda.synchronize()
    want_v = x0 @ A.T + b if velocity_update else v0 + dt * (x0 @ C.T + d)
    assert np.abs(host(ps.velocity) - want_v).max() < 10 * TOL[dtype]
    assert np.abs(host(ps.loc) - (x0 + dt * (x0 @ A.T + b))).max() < 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("element", ["linear1d", "quad4"])
def test_p2g_first_moments_and_virial(element, dtype):
    el, ps = setup(element, dtype, seed=14)
    el.compute_nodal_mass(ps)
    el.compute_nodal_momentum(ps)
    torch.cuda.synchronize()
    X, x = host(el.nodes.loc), host(ps.loc)
    m, v = host(ps.mass)[:, 0], host(ps.velocity)
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()  # noqa: E731
    assert rel(X.T @ host(el.nodes.mass)[:, 0], x.T @ m) < TOL[dtype]
    assert rel(X.T @ host(el.nodes.momentum), x.T @ (m[:, None] * v)) < 10 * TOL[dtype]
    # virial of the internal force: sum_i X_i,a f_int,i,b = -sum_p V_p sigma_ab
    r = np.random.Generator(np.random.PCG64(16))
    s = r.standard_normal((len(ps), 6))
    ps.stress = torch.from_numpy(s.reshape(len(ps), 6, 1))
    el.compute_internal_force(ps)
    torch.cuda.synchronize()
    M = X.T @ host(el.nodes.f_int)
    V = host(ps.volume)[:, 0]
    s = host(ps.stress)
    if el.ndim == 1:
        want = np.array([[-(V * s[:, 0]).sum()]])
    else:
        want = -np.array([[(V * s[:, 0]).sum(), (V * s[:, 3]).sum()],
                          [(V * s[:, 3]).sum(), (V * s[:, 1]).sum()]])
    assert rel(M, want) < 10 * TOL[dtype]
