"""Gradients of Linear1D rollouts through the hand-written 1D adjoint
(dmpm_step_adjoint on a Linear1D plan), against central differences of the
FP64 oracle's forward rollout.

* The reference's own gradient case (examples/optim_1d.py:12-61): one
  Linear1D element of length 1, node 0 fixed, one particle at x = 0.5 with
  v = 0.1, SimpleMaterial (density 1), dt = 0.01, 1000 USL steps, loss =
  |v_final(E) - v_final(E_true = 100)|; d loss / d E at E = 105 within 1e-4
  of the oracle's central difference (FP64), and within 2e-3 from the FP32
  adjoint.
* A 256-particle bar (128 elements on [0, 1], two Gauss particles each, v0 =
  0.1 sin(pi x / 2), node 0 fixed, E = 100, rho = 1, dt = 2e-4: a stable
  time step of the C1 bar) for USF and USL, LinearElastic and SimpleMaterial,
  with and without the velocity update, from a seeded stress field: the
  directional derivatives of L = sum(w * final state) along a per-particle
  log E direction and along directions of the initial x, vx, stress and V
  match the oracle's central differences within 1e-4.
* A checkpointed rollout gives the same gradients as a plain one.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- optim_1d bar
def _bar_gpu(E, dtype):
    import diffmpm_amd as d
    el = d.Linear1D(1, 1, 1.0, [(torch.tensor([0]), d.Constraint(0, 0.0))], device="cuda",
                    dtype=dtype)
    ps = d.Particles(torch.tensor([0.5], dtype=dtype).reshape(1, 1, 1),
                     d.SimpleMaterial({"E": E, "density": 1.0}),
                     torch.tensor([0], dtype=torch.int32), device="cuda", dtype=dtype)
    ps.velocity = ps.velocity + 0.1
    mesh = d.Mesh1D({"particles": [ps], "elements": el})
    return d.MPMExplicit(mesh, 0.01, "usl", False, 1000)


def _bar_oracle(E, nsteps=1000):
    from oracle import mpm as om
    el = om.Elements("linear1d", 1, 1.0, [(np.array([0]), om.Constraint(0, 0.0))], [],
                     np.float64)
    ps = om.Particles(np.array([[0.5]]), om.SimpleMaterial({"E": E, "density": 1.0}),
                      dtype=np.float64)
    ps.velocity[:] = 0.1
    sim = om.MPMExplicit(om.Mesh(el, [ps]), 0.01, "usl", False, nsteps)
    sim.solve(np.array([0.0]))
    return ps.velocity[:, 0].copy()


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-4), (torch.float32, 2e-3)])
def test_optim_1d_bar_gradient_matches_oracle_central_differences(dtype, tol):
    from diffmpm_amd.autodiff import Differentiable, rollout
    E0, eps = 105.0, 1e-4
    target = _bar_oracle(100.0)

    def loss_np(E):
        return float(np.linalg.norm(_bar_oracle(E) - target))

    fd = (loss_np(E0 * math.exp(eps)) - loss_np(E0 * math.exp(-eps))) / (2 * eps) / E0
    sim = _bar_gpu(E0, dtype)
    d = Differentiable(sim)
    le = torch.tensor(math.log(E0), dtype=torch.float64, device="cuda", requires_grad=True)
    out = rollout(d, 1000, torch.tensor([0.0], dtype=torch.float64), log_e=le)
    loss = torch.linalg.norm(out[2].double() - torch.from_numpy(target).cuda())
    loss.backward()
    got = float(le.grad) / E0
    assert abs(loss.item() - loss_np(E0)) <= 1e-6 + 1e-4 * loss_np(E0), (loss.item(), loss_np(E0))
    assert math.isclose(got, fd, rel_tol=tol), (got, fd)
    print(f"optim_1d bar {dtype}: dL/dE = {got:.8e}, oracle FD {fd:.8e}")


def test_optim_1d_adam_iterates_follow_the_oracle_gradients():
    """Ten Adam iterations on E (the reference's optimiser: optax.adam at
    learning rate 0.1, written out below with optax's defaults) from E = 105,
    once with the GPU adjoint's dL/dE and once with the FP64 oracle's central
    differences: the iterates agree within 1e-7 relative."""
    from diffmpm_amd.autodiff import Differentiable, rollout
    target_np = _bar_oracle(100.0)
    target = torch.from_numpy(target_np).cuda()
    sim = _bar_gpu(105.0, torch.float64)
    d = Differentiable(sim)
    state0 = d.get_state().clone()
    g0 = torch.tensor([0.0], dtype=torch.float64)

    def grad_gpu(E):
        le = torch.tensor(math.log(E), dtype=torch.float64, device="cuda", requires_grad=True)
        out = rollout(d, 1000, g0, state=state0, log_e=le)
        loss = torch.linalg.norm(out[2] - target)
        loss.backward()
        return float(le.grad) / E, loss.item()

    def grad_fd(E, eps=1e-4):
        f = [float(np.linalg.norm(_bar_oracle(E * math.exp(s * eps)) - target_np))
             for s in (1, -1)]
        return (f[0] - f[1]) / (2 * eps) / E, float(np.linalg.norm(_bar_oracle(E) - target_np))

    def adam(grad, E, iters=10, lr=0.1, b1=0.9, b2=0.999, eps=1e-8):
        m = v = 0.0
        path = []
        for t in range(1, iters + 1):
            g, loss = grad(E)
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            E = E - lr * (m / (1 - b1 ** t)) / (math.sqrt(v / (1 - b2 ** t)) + eps)
            path.append((E, loss))
        return path

    got, want = adam(grad_gpu, 105.0), adam(grad_fd, 105.0)
    for (eg, lg), (ew, lw) in zip(got, want):
        assert math.isclose(eg, ew, rel_tol=1e-7), (got, want)
        assert math.isclose(lg, lw, rel_tol=1e-6, abs_tol=1e-9), (got, want)
    print("optim_1d Adam path (E, loss):", [(round(e, 4), round(l, 6)) for e, l in got])


# ---------------------------------------------------------------- 256-particle bar
def _bar256(scheme, material, vu):
    nx = 128
    h = 1.0 / nx
    g = (1.0 - 1.0 / math.sqrt(3.0)) / 2.0
    base = np.arange(nx, dtype=np.float64)
    pts = np.stack([base + g, base + 1 - g], axis=1).reshape(-1, 1) * h
    vel = 0.1 * np.sin(np.pi * pts / 2.0)
    mat = ({"type": "LinearElastic", "density": 1.0, "youngs_modulus": 100.0,
            "poisson_ratio": 0.0} if material == "LinearElastic"
           else {"type": "SimpleMaterial", "density": 1.0, "E": 100.0})
    rng = np.random.Generator(np.random.PCG64(3))
    youngs = 100.0 * np.exp(0.2 * rng.standard_normal(len(pts)))
    return dict(element="linear1d", nelements=nx, el_len=h, particles=pts.tolist(),
                constraints=[{"nodes": [0], "dir": 0, "velocity": 0.0}], nodal_forces=[],
                gravity=[0.0], init_velocity=vel.tolist(), material=mat, tractions=[],
                scheme=scheme, dt=2e-4, nsteps=1, velocity_update=vu, youngs=youngs.tolist())


# state rows the oracle has in 1D: x, vx, sigma_xx, sigma_yy, sigma_xy, V
ROWS_1D = (0, 2, 4, 5, 6, 7)


def _oracle_final(spec, K, log_e=None, dstate=None):
    import cases
    s = dict(spec)
    if log_e is not None:
        s["youngs"] = np.exp(log_e).tolist()
    osim = cases.build_oracle(s, np.float64)
    ps = osim.mesh.particles[0]
    if dstate is not None:
        ps.loc[:, 0] += dstate[0]
        ps.velocity[:, 0] += dstate[2]
        ps.stress[:, 0] += dstate[4]
        ps.stress[:, 1] += dstate[5]
        ps.stress[:, 3] += dstate[6]
        ps.volume[:] += dstate[7]
    for k in range(K):
        osim.step(k, np.array([0.0]))
    return np.stack([ps.loc[:, 0], ps.velocity[:, 0], ps.stress[:, 0], ps.stress[:, 1],
                     ps.stress[:, 3], ps.volume])


CASES = [(sc, mat, vu) for sc in ("usf", "usl") for mat in ("LinearElastic", "SimpleMaterial")
         for vu in (False, True)]


# perturbation directions of the initial state (row lists) and their step
DIRECTIONS = (("x", (0,), 1e-6), ("vx", (2,), 1e-5), ("stress", (4, 5, 6), 1e-4),
              ("V", (7,), 1e-6))


@pytest.mark.parametrize("scheme,material,vu", CASES)
def test_bar256_directional_derivatives_match_oracle_central_differences(scheme, material, vu):
    import cases
    from diffmpm_amd.autodiff import Differentiable, rollout
    K = 20
    spec = _bar256(scheme, material, vu)
    sim = cases.build_gpu(spec, torch.float64, keep_temporaries=False)
    d = Differentiable(sim)
    n = d.n
    rng = np.random.Generator(np.random.PCG64(11))
    w = rng.standard_normal((6, n))
    le0 = np.log(np.asarray(spec["youngs"]))
    # a stressed start: from sigma = 0 the first USL step has f_int = 0 and
    # a = 0 exactly, where snap(a) (element.py:326-329) is flat and its
    # derivative, like jax.grad's of jnp.where, is 0, while a central
    # difference steps over the 1e-12 wide flat part
    base = np.zeros((8, n))
    base[4:7] = 1e-2 * rng.standard_normal((3, n))
    state0 = (d.get_state() + torch.from_numpy(base).cuda()).requires_grad_(True)
    le = torch.tensor(le0, device="cuda", requires_grad=True)
    out = rollout(d, K, torch.tensor([0.0], dtype=torch.float64), state=state0, log_e=le)
    loss = (torch.from_numpy(w).cuda() * out[list(ROWS_1D)]).sum()
    loss.backward()
    gs, ge = state0.grad.cpu().numpy(), le.grad.cpu().numpy()
    # y and vy are not 1D quantities: the loss does not read them and they
    # pass their (zero) cotangents through
    assert not gs[1].any() and not gs[3].any()

    def fd(eps, log_e=None, dstate=None):
        f = []
        for sgn in (1, -1):
            le_ = le0 if log_e is None else le0 + sgn * eps * log_e
            ds = base if dstate is None else base + sgn * eps * dstate
            f.append(_oracle_final(spec, K, log_e=le_, dstate=ds))
        return float((w * (f[0] - f[1])).sum()) / (2 * eps)

    u = rng.standard_normal(n)
    got, want = float((ge * u).sum()), fd(1e-5, log_e=u)
    assert math.isclose(got, want, rel_tol=1e-4), ("log E", got, want)
    for name, rows, eps in DIRECTIONS:
        us = np.zeros((8, n))
        for r in rows:
            us[r] = rng.standard_normal(n)
        got, want = float((gs * us).sum()), fd(eps, dstate=us)
        assert math.isclose(got, want, rel_tol=1e-4), (name, got, want)


def test_bar256_checkpointed_gradients_equal_plain_ones():
    import cases
    from diffmpm_amd.autodiff import Differentiable, rollout
    spec = _bar256("usf", "SimpleMaterial", False)
    grads = []
    for ck in (0, 4):
        sim = cases.build_gpu(spec, torch.float64, keep_temporaries=False)
        d = Differentiable(sim)
        s0 = d.get_state().clone().requires_grad_(True)
        le = torch.tensor(np.log(np.asarray(spec["youngs"])), device="cuda", requires_grad=True)
        out = rollout(d, 12, torch.tensor([0.0], dtype=torch.float64), state=s0, log_e=le,
                      checkpoint_every=ck)
        (out * torch.arange(1, 9, device="cuda", dtype=out.dtype)[:, None]).sum().backward()
        grads.append((s0.grad.clone(), le.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


# ---------------------------------------------------------------- error paths
def test_differentiable_rejects_particle_tractions_and_external_forces():
    import cases
    from diffmpm_amd.autodiff import Differentiable
    spec = _bar256("usl", "SimpleMaterial", False)
    sim = cases.build_gpu(spec, torch.float64, keep_temporaries=False)
    ps = sim.mesh.particles[0]
    ps.f_ext = torch.ones((len(ps), 1, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="particle external forces"):
        Differentiable(sim)
    spec["tractions"] = [{"pset": [0], "pids": [3], "dir": 0, "traction": 1.0,
                          "function": None}]
    with pytest.raises(ValueError, match="particle tractions"):
        Differentiable(cases.build_gpu(spec, torch.float64, keep_temporaries=False))


def test_1d_adjoint_entry_point_rejects_particle_tractions():
    """dmpm_step_adjoint on a Linear1D plan with particle tractions: EINVAL
    before any work."""
    import cases
    from diffmpm_amd import _lib
    from diffmpm_amd._plan import Plan
    spec = _bar256("usl", "SimpleMaterial", False)
    spec["tractions"] = [{"pset": [0], "pids": [3], "dir": 0, "traction": 1.0,
                          "function": None}]
    sim = cases.build_gpu(spec, torch.float64, keep_temporaries=False)
    ps, el = sim.mesh.particles[0], sim.mesh.elements
    plan = Plan(el, ps)
    pw, nw, eid = ps._buf.clone(), el.nodes._buf.clone(), ps._eid.clone()
    cot = torch.zeros((8, ps._buf.shape[1]), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="particle tractions"):
        _lib.call("dmpm_step_adjoint", plan.h, _lib.ptr(ps._buf), _lib.ptr(pw), _lib.ptr(eid),
                  _lib.ptr(nw), _lib.ptr(cot), None, _lib.DMPM_USL, 0, plan._gravity([0.0]),
                  2e-4, 0, plan.s)


def test_plan_rejects_simple_material_with_a_per_particle_modulus():
    """A SimpleMaterial with a per-particle E is passed as LinearElastic
    coefficients (1, 0, 1); the raw pair is refused with that advice."""
    from ctypes import byref, c_void_p
    from diffmpm_amd import _lib
    d = _lib.PlanDesc()
    d.element, d.dtype = _lib.DMPM_LINEAR1D, _lib.dtype_code(torch.float32)
    d.nx, d.ny, d.hx, d.hy = 4, 1, 0.25, 0.25
    d.nparticles, d.ld, d.nld = 8, 8, 5
    d.material, d.per_particle_E = _lib.DMPM_SIMPLE_MATERIAL, 1
    for k, v in enumerate((1.0, 0.0, 0.0)):
        d.de[k] = v
    h = c_void_p()
    with pytest.raises(ValueError, match=r"de = \(1, 0, 1\)"):
        _lib.call("dmpm_plan_create", byref(d), byref(h))
