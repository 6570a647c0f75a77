"""Adjoint of the fused 2D step (dmpm_step_adjoint via diffmpm_amd.autodiff).

The reference has no gradient test (it relies on jax.grad), so the adjoint is
pinned by central finite differences of the FP64 oracle's forward rollout
(the reference step restated in NumPy, from the same initial state and
per-particle E): per-particle log E on 16 seeded particles (relative step
1e-3, as in the inverse-problem configuration), directional derivatives along
seeded random directions of the whole initial state and of all log E, and
every state component of 4 particles.  A rollout is short and slow enough
that no particle changes cell under the perturbations (checked), so the loss
is smooth there.  The GPU forward itself matches the oracle's within 1e-12.
"""
import math

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu

NSTEPS = 4
FD_H_LOGE = 1e-3
RTOL_FD = 2e-5


def _sim(scheme, vu, dtype=torch.float64, seed=3, material="LinearElastic", nodal_force=True):
    spec = cases.synthetic_spec("quad4", nx=10, ny=8, rows=5, seed=seed, scheme=scheme,
                                velocity_update=vu, traction=False, nodal_force=nodal_force, E=1e6,
                                dt=1e-5, nsteps=NSTEPS, per_particle_E=True, velocity_scale=0.5)
    if material == "SimpleMaterial":
        spec["material"] = {"type": "SimpleMaterial", "E": 1e6, "density": 1000.0}
    return cases.build_gpu(spec, dtype, keep_temporaries=False), spec


def _weights(state0, seed=5):
    """Seeded loss weights, each state row scaled to O(1) contributions."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    w = torch.randn(state0.shape, generator=g, dtype=torch.float64).to(state0.device)
    scale = state0.abs().amax(dim=1, keepdim=True).clamp_min(1.0)
    return (w / scale).to(state0.dtype)


def _setup(scheme, vu, dtype=torch.float64, material="LinearElastic", nodal_force=True):
    from diffmpm_amd.autodiff import Differentiable
    sim, spec = _sim(scheme, vu, dtype, material=material, nodal_force=nodal_force)
    d = Differentiable(sim)
    d.spec = spec  # the oracle's copy of the case
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    state0 = d.get_state()
    # give the stress and volume rows non-trivial values so that every term
    # of the step carries a gradient
    gen = torch.Generator(device="cpu").manual_seed(11)
    noise = torch.randn(state0.shape, generator=gen, dtype=torch.float64).to(state0.device)
    state0[4:7] = (2e3 * noise[4:7]).to(dtype)
    state0[7] = state0[7] * (1 + 0.05 * noise[7]).to(dtype)
    log_e = torch.log(d.ps.youngs_modulus.detach().clone())
    return d, g, state0, log_e


def _oracle_rollout(spec, state0, log_e):
    """The final (8, N) state of NSTEPS FP64 oracle steps from state0 (rows
    x, y, vx, vy, sigma_xx, sigma_yy, sigma_xy, V) with E_p = exp(log_e)."""
    osim = cases.build_oracle(spec, np.float64)
    ps = osim.mesh.particles[0]
    s = state0.detach().double().cpu().numpy()
    ps.loc = np.stack([s[0], s[1]], axis=1)
    ps.velocity = np.stack([s[2], s[3]], axis=1)
    ps.stress[:, 0], ps.stress[:, 1], ps.stress[:, 3] = s[4], s[5], s[6]
    ps.volume = s[7].copy()
    E = np.exp(log_e.detach().double().cpu().numpy())
    ps.youngs_modulus = np.broadcast_to(E, (len(ps),)).copy()
    g = np.asarray(spec["gravity"])
    for i in range(NSTEPS):
        osim.step(i, g)
    return np.stack([ps.loc[:, 0], ps.loc[:, 1], ps.velocity[:, 0], ps.velocity[:, 1],
                     ps.stress[:, 0], ps.stress[:, 1], ps.stress[:, 3], ps.volume])


def _loss(d, g, state0, log_e, w):
    """The loss of the FP64 oracle's rollout (the finite differences' forward)."""
    out = _oracle_rollout(d.spec, state0, log_e)
    return float((w.double().cpu().numpy() * out).sum())


def test_gpu_rollout_matches_the_oracle_rollout():
    for scheme, vu in (("usf", False), ("usl", True)):
        d, g, state0, log_e = _setup(scheme, vu)
        from diffmpm_amd.autodiff import rollout
        with torch.no_grad():
            out = rollout(d, NSTEPS, g, state=state0, log_e=log_e).double().cpu().numpy()
        want = _oracle_rollout(d.spec, state0, log_e)
        for r in range(8):
            assert cases.normwise(out[r], want[r]) < 1e-12, (scheme, r)


def _cells(d, state):
    h = d.el._hx
    return torch.floor(state[0:2].double() / h).long()


def _grad(d, g, state0, log_e, w, checkpoint_every=0):
    from diffmpm_amd.autodiff import rollout
    s = state0.clone().requires_grad_(True)
    le = log_e.clone().requires_grad_(True)
    out = rollout(d, NSTEPS, g, state=s, log_e=le, checkpoint_every=checkpoint_every)
    (w * out).sum().backward()
    return s.grad.detach(), le.grad.detach(), out.detach()


@pytest.mark.parametrize("scheme,vu,material", [
    ("usf", False, "LinearElastic"), ("usl", False, "LinearElastic"),
    ("usf", True, "LinearElastic"), ("usl", True, "LinearElastic"),
    ("usf", False, "SimpleMaterial"), ("usl", True, "SimpleMaterial")])
def test_log_e_gradient_matches_central_differences(scheme, vu, material):
    """Per-particle log E; for SimpleMaterial E_p replaces the material's E."""
    d, g, state0, log_e = _setup(scheme, vu, material=material)
    w = _weights(state0)
    gs, ge, out = _grad(d, g, state0, log_e, w)
    assert torch.equal(_cells(d, state0), _cells(d, out)), "a particle changed cell"
    rng = np.random.Generator(np.random.PCG64(0))
    pids = sorted(rng.choice(d.n, size=16, replace=False).tolist())
    fd = []
    for p in pids:
        lp, lm = log_e.clone(), log_e.clone()
        lp[p] += FD_H_LOGE
        lm[p] -= FD_H_LOGE
        fd.append((_loss(d, g, state0, lp, w) - _loss(d, g, state0, lm, w)) / (2 * FD_H_LOGE))
    fd = torch.tensor(fd, dtype=torch.float64)
    ad = ge[pids].double().cpu()
    scale = fd.abs().max()
    assert scale > 0
    assert torch.allclose(ad, fd, rtol=RTOL_FD, atol=RTOL_FD * scale), (ad, fd)


@pytest.mark.parametrize("scheme,vu", [("usf", False), ("usl", False), ("usf", True),
                                       ("usl", True)])
def test_directional_derivatives_match_central_differences(scheme, vu):
    d, g, state0, log_e = _setup(scheme, vu)
    w = _weights(state0)
    gs, ge, _ = _grad(d, g, state0, log_e, w)
    gen = torch.Generator(device="cpu").manual_seed(7)
    # direction in the state: each row scaled to its magnitude; positions by a
    # small absolute amount so that no particle leaves its cell
    u = torch.randn(state0.shape, generator=gen, dtype=torch.float64).to(state0.device)
    row = state0.abs().amax(dim=1, keepdim=True).clamp_min(1.0)
    u = u * row
    u[0:2] = u[0:2] / row[0:2] * 1e-3 * d.el._hx
    h = 1e-4
    lp = _loss(d, g, state0 + h * u, log_e, w)
    lm = _loss(d, g, state0 - h * u, log_e, w)
    fd = (lp - lm) / (2 * h)
    ad = float((gs.double() * u).sum())
    assert abs(fd) > 1e-6
    assert math.isclose(ad, fd, rel_tol=RTOL_FD, abs_tol=1e-12), (ad, fd)
    ue = torch.randn(log_e.shape, generator=gen, dtype=torch.float64).to(log_e.device)
    lp = _loss(d, g, state0, log_e + FD_H_LOGE * ue, w)
    lm = _loss(d, g, state0, log_e - FD_H_LOGE * ue, w)
    fd = (lp - lm) / (2 * FD_H_LOGE)
    ad = float((ge.double() * ue).sum())
    assert abs(fd) > 1e-6
    assert math.isclose(ad, fd, rel_tol=RTOL_FD, abs_tol=1e-12), (ad, fd)


@pytest.mark.parametrize("scheme,vu", [("usf", False), ("usl", True)])
def test_every_state_component_matches_central_differences(scheme, vu):
    """d loss / d state[r, p] for every row r on 4 particles, one central
    difference each (steps relative to the row's magnitude; positions by a
    small absolute amount)."""
    d, g, state0, log_e = _setup(scheme, vu)
    w = _weights(state0)
    gs, _, _ = _grad(d, g, state0, log_e, w)
    rng = np.random.Generator(np.random.PCG64(1))
    pids = sorted(rng.choice(d.n, size=4, replace=False).tolist())
    row = state0.abs().amax(dim=1)
    for r in range(8):
        h = 1e-6 * d.el._hx if r < 2 else 1e-5 * float(row[r].clamp_min(1.0))
        for p in pids:
            sp, sm = state0.clone(), state0.clone()
            sp[r, p] += h
            sm[r, p] -= h
            fd = (_loss(d, g, sp, log_e, w) - _loss(d, g, sm, log_e, w)) / (2 * h)
            ad = float(gs[r, p])
            assert math.isclose(ad, fd, rel_tol=1e-4, abs_tol=1e-7 * max(1.0, abs(fd))), \
                (r, p, ad, fd)


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_checkpointed_rollout_gives_identical_gradients(scheme):
    d, g, state0, log_e = _setup(scheme, False)
    w = _weights(state0)
    a = _grad(d, g, state0, log_e, w, checkpoint_every=0)
    b = _grad(d, g, state0, log_e, w, checkpoint_every=3)
    c = _grad(d, g, state0, log_e, w, checkpoint_every=0)
    for x, y, z in zip(a, b, c):
        assert torch.equal(x, y)
        assert torch.equal(x, z)


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_fp32_adjoint_matches_fp64(scheme):
    d64, g, s64, le64 = _setup(scheme, False, torch.float64)
    d32, _, s32, le32 = _setup(scheme, False, torch.float32)
    w = _weights(s64)
    gs64, ge64, _ = _grad(d64, g, s64, le64, w)
    gs32, ge32, _ = _grad(d32, g, s32, le32, w.float())
    for a, b in ((gs32, gs64), (ge32, ge64)):
        err = (a.double() - b).norm() / b.norm()
        assert err < 1e-3, float(err)


@pytest.mark.parametrize("scheme,dtype,nodal_force", [
    ("usf", torch.float64, True), ("usl", torch.float32, True), ("usf", torch.float32, False)])
def test_captured_segments_equal_eager_segments(scheme, dtype, nodal_force):
    """rollout(..., graphs=True) gives the eager checkpointed rollout's final
    state and gradients bit for bit, on two backward passes in a row (the
    second replays the captured graphs), with segments of two lengths (3, 3,
    1), per-particle and scalar log E; with a time-dependent nodal force each
    segment has graphs of its own, without one the segments share them."""
    from diffmpm_amd.autodiff import rollout
    d, g, state0, log_e = _setup(scheme, False, dtype=dtype, nodal_force=nodal_force)
    w = _weights(state0)

    def run(graphs, scalar):
        s = state0.clone().requires_grad_(True)
        le = (log_e.mean() if scalar else log_e).clone().requires_grad_(True)
        out = rollout(d, 7, g, state=s, log_e=le, checkpoint_every=3, graphs=graphs)
        (w * out).sum().backward()
        return out.detach(), s.grad.detach(), le.grad.detach()

    for scalar in (False, True):
        a = run(False, scalar)
        b = run(True, scalar)
        c = run(True, scalar)
        for x, y, z in zip(a, b, c):
            assert torch.equal(x, y) and torch.equal(x, z), scalar
        assert torch.isfinite(a[1]).all() and a[2].abs().sum() > 0
    # per segment length: a forward and a backward graph, per segment start
    # when the nodal force depends on time
    assert len(d._graphs.graphs) == (2 * 3 if nodal_force else 2 * 2)


def test_scalar_log_e_gradient_is_the_sum_of_the_per_particle_terms():
    d, g, state0, log_e = _setup("usl", False)
    w = _weights(state0)
    _, ge, _ = _grad(d, g, state0, log_e, w)
    from diffmpm_amd.autodiff import rollout
    le0 = torch.tensor(math.log(1e6), dtype=torch.float64, device=state0.device,
                       requires_grad=True)
    out = rollout(d, NSTEPS, g, state=state0.clone(), log_e=le0)
    (w * out).sum().backward()
    lp = _loss(d, g, state0, le0.detach() + FD_H_LOGE, w)
    lm = _loss(d, g, state0, le0.detach() - FD_H_LOGE, w)
    fd = (lp - lm) / (2 * FD_H_LOGE)
    assert math.isclose(float(le0.grad), fd, rel_tol=RTOL_FD), (float(le0.grad), fd)


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_saved_forward_and_recomputation_give_identical_gradients(scheme):
    """The backward pass from the forward step's saved quantities (default)
    and from a recomputation of the step give bit-identical gradients."""
    d, g, state0, log_e = _setup(scheme, False)
    w = _weights(state0)
    d.save_forward = True
    a = _grad(d, g, state0, log_e, w)
    d.save_forward = False
    b = _grad(d, g, state0, log_e, w)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
