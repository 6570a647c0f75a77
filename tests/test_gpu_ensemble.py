"""Batched ensembles (one batch plan for several independent simulations).

The reference has no batched path; an ensemble must give, member by member,
exactly what each member gives when stepped on its own (same kernels, same
per-member arithmetic and summation order), including after particles change
cells, and its adjoint must give each member's own gradients.
"""
import math

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu

MEMBERS = [  # (nx, ny, rows, seed, E, nu, nodal force)
    (9, 7, 4, 1, 1e3, 0.30, True),
    (16, 12, 9, 2, 2e3, 0.20, False),
    (5, 5, 5, 3, 5e2, 0.40, True),
    (23, 9, 3, 4, 1.5e3, 0.25, False),
    (12, 20, 14, 5, 8e2, 0.35, True),
]


def _spec(nx, ny, rows, seed, E, nu, nf, scheme, per_E=False):
    spec = cases.synthetic_spec("quad4", nx=nx, ny=ny, rows=rows, seed=seed, scheme=scheme,
                                traction=False, nodal_force=nf, E=E, nu=nu, dt=1e-3,
                                nsteps=12, velocity_scale=0.3, drift=(0.4, 0.2),
                                per_particle_E=per_E)
    if nf:
        spec["nodal_forces"][0]["force"] = -0.05
    # keep particles two cells away from the mesh edges
    h = spec["el_len"][0]
    pts = np.asarray(spec["particles"])
    lo = 2 * h
    hi = np.array([nx * h, ny * h]) - 2 * h
    spec["particles"] = np.clip(pts, lo, hi).tolist()
    return spec


def _members(scheme, dtype, per_E=False):
    return [cases.build_gpu(_spec(*m, scheme, per_E), dtype, keep_temporaries=False)
            for m in MEMBERS]


def _grav(spec):
    return torch.tensor(spec["gravity"], dtype=torch.float64)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_ensemble_equals_members_stepped_alone(scheme, dtype):
    from diffmpm_amd._plan import plan_for
    from diffmpm_amd.ensemble import Ensemble
    g = _grav(_spec(*MEMBERS[0], scheme))
    alone = _members(scheme, dtype)
    batch = _members(scheme, dtype)
    ens = Ensemble(batch)
    e0 = ens.eid.clone()
    for k in range(0, 12, 4):
        for sim in alone:
            plan_for(sim.mesh.elements, sim.mesh.particles[0]).run(scheme, False, g, sim.dt, k, 4,
                                                                    False)
        ens.run(4, g, start_step=k)
    ens.scatter()
    assert not torch.equal(e0, ens.eid)  # cells did change
    for a, b in zip(alone, batch):
        pa, pb = a.mesh.particles[0], b.mesh.particles[0]
        n = len(pa)
        assert torch.equal(pa._buf[:, :n].view(torch.int8), pb._buf[:, :n].view(torch.int8))
        assert torch.equal(pa._eid[:n], pb._eid[:n])
        na, nb = a.mesh.elements.nodes, b.mesh.elements.nodes
        m = na._count
        assert torch.equal(na._buf[:, :m].view(torch.int8), nb._buf[:, :m].view(torch.int8))


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_ensemble_log_e_gradients_equal_members_alone(scheme):
    from diffmpm_amd.autodiff import Differentiable, rollout
    from diffmpm_amd.ensemble import Ensemble
    dtype = torch.float64
    g = _grav(_spec(*MEMBERS[0], scheme))
    K = 3
    # each member alone: d/dlogE of a weighted sum of its final state
    alone = []
    for m in MEMBERS:
        sim = cases.build_gpu(_spec(*m, scheme), dtype, keep_temporaries=False)
        d = Differentiable(sim)
        le = torch.tensor(math.log(m[4]), dtype=dtype, device="cuda", requires_grad=True)
        s0 = d.get_state()
        out = rollout(d, K, g, state=s0, log_e=le)
        w = torch.linspace(-1, 1, out.numel(), dtype=dtype, device="cuda").reshape(out.shape)
        (w * out).sum().backward()
        alone.append((float(le.grad), w))
    ens = Ensemble(_members(scheme, dtype), per_particle_E=True)
    d = Differentiable.of_ensemble(ens)
    le = torch.tensor([math.log(m[4]) for m in MEMBERS], dtype=dtype, device="cuda",
                      requires_grad=True)
    idx = ens.sim_index()
    s0 = d.get_state()
    out = rollout(d, K, g, state=s0, log_e=le[idx])
    loss = 0
    for i, (_, w) in enumerate(alone):
        loss = loss + (w * out[:, ens.sim_slice(i)]).sum()
    loss.backward()
    for i, (ga, _) in enumerate(alone):
        assert math.isclose(float(le.grad[i]), ga, rel_tol=1e-10, abs_tol=1e-14), (i, ga)


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_ensemble_checkpointed_rollout_graphs_equal_eager(scheme):
    """An ensemble's checkpointed rollout with captured segment graphs gives
    the eager segments' final state and per-member d/dlogE bit for bit (two
    backward passes in a row replay the graphs), and both equal the
    uncheckpointed rollout; the nodal force's time function gives each
    segment start graphs of its own."""
    from diffmpm_amd.autodiff import Differentiable, rollout
    from diffmpm_amd.ensemble import Ensemble
    dtype = torch.float64
    g = _grav(_spec(*MEMBERS[0], scheme))
    ens = Ensemble(_members(scheme, dtype), per_particle_E=True)
    d = Differentiable.of_ensemble(ens)
    idx = ens.sim_index()
    s0 = d.get_state()
    w = torch.linspace(-1, 1, s0.numel(), dtype=dtype, device="cuda").reshape(s0.shape)

    def run(ck, graphs):
        le = torch.tensor([math.log(m[4]) for m in MEMBERS], dtype=dtype, device="cuda",
                          requires_grad=True)
        out = rollout(d, 7, g, state=s0, log_e=le[idx], checkpoint_every=ck, graphs=graphs)
        (w * out).sum().backward()
        return out.detach(), le.grad.detach()

    plain = run(0, False)
    eager = run(3, False)
    for x, y in zip(plain, eager):
        assert torch.equal(x, y)
    for _ in range(2):
        for x, y in zip(plain, run(3, True)):
            assert torch.equal(x, y)
    assert len(d._graphs.graphs) == 2 * 3


def _mixed_specs(scheme):
    """Members of several tiles per side (one 10 tiles wide: two strips of
    the tile order, the second narrower), a small one, and a 4 x 4 mesh with
    3000 particles in one cell (more than one workgroup stages at once: the
    overflow workgroups) and three particles outside the mesh (the tail
    workgroups)."""
    specs = []
    for nx, ny, rows, seed in ((150, 20, 15, 31), (40, 36, 30, 32), (9, 7, 4, 33)):
        s = cases.synthetic_spec("quad4", nx=nx, ny=ny, rows=rows, seed=seed, scheme=scheme,
                                 traction=False, nodal_force=True, E=1e3, dt=1e-3, nsteps=8,
                                 velocity_scale=0.05, drift=(0.6, 0.3))
        s["nodal_forces"][0]["force"] = -0.05
        h = s["el_len"][0]
        s["particles"] = (np.asarray(s["particles"]) * np.array([(nx - 6) / nx, 1.0])
                          + 3 * h).tolist()
        specs.append(s)
    rng = np.random.Generator(np.random.PCG64(11))
    h = 0.25
    crowd = np.stack([rng.uniform(1.05, 1.95, 3000), rng.uniform(1.05, 1.95, 3000)], 1) * h
    bg = rng.uniform(0.02, 0.98, size=(40, 2))
    pts = np.concatenate([bg[:20], crowd, bg[20:]])
    pts[:3] = [[1.2, 0.3], [np.nan, 0.4], [0.3, 1.5]]
    specs.insert(1, dict(
        element="quad4", nelements=[4, 4], el_len=[h, h], particles=pts.tolist(),
        init_velocity=(0.1 * rng.standard_normal(pts.shape)).tolist(),
        material={"type": "LinearElastic", "density": 1000.0, "youngs_modulus": 1e5,
                  "poisson_ratio": 0.3},
        constraints=[{"nodes": list(range(5)), "dir": 1, "velocity": 0.0}],
        nodal_forces=[], tractions=[], gravity=[0.0, -9.81], scheme=scheme, dt=1e-3,
        nsteps=8, velocity_update=False))
    return specs


def _bits(t):
    """Bit pattern of a block (every NaN encoded alike)."""
    t = t.clone()
    if t.is_floating_point():
        t = torch.where(torch.isnan(t), torch.full_like(t, float("nan")), t)
        t = t.view(torch.int32 if t.dtype == torch.float32 else torch.int64)
    return t


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_batch_one_launch_step_equals_phase_launches(scheme, dtype):
    """A batch plan's one-launch step (every member's tiles in one grid)
    gives bit-identical states to its launches per phase, across cell
    changes, crowded cells and particles outside their mesh."""
    from diffmpm_amd import _lib
    from diffmpm_amd.ensemble import Ensemble
    specs = _mixed_specs(scheme)
    a = Ensemble([cases.build_gpu(s, dtype, keep_temporaries=False) for s in specs])
    b = Ensemble([cases.build_gpu(s, dtype, keep_temporaries=False) for s in specs])
    a.plan.set_option(_lib.DMPM_OPT_STEP_KERNEL, 1)
    b.plan.set_option(_lib.DMPM_OPT_STEP_KERNEL, 0)
    g = torch.tensor(specs[0]["gravity"], dtype=torch.float64)
    e0 = a.eid.clone()
    for k in (0, 4):
        a.run(4, g, start_step=k)
        b.run(4, g, start_step=k)
        for x, y in ((a.pbuf, b.pbuf), (a.nbuf, b.nbuf), (a.eid, b.eid)):
            assert torch.equal(_bits(x), _bits(y)), k
    assert not torch.equal(e0, a.eid)  # cells did change
    n1 = a.counts[0]
    assert torch.isnan(a.pbuf[0, n1 + 1])  # the crowded member's particle at a NaN position


def test_batch_plan_rejects_per_method_calls():
    from diffmpm_amd import _lib
    from diffmpm_amd.ensemble import Ensemble
    ens = Ensemble(_members("usf", torch.float32)[:2])
    with pytest.raises(ValueError, match="batch plans"):
        _lib.call("dmpm_set_particle_element_ids", ens.plan.h, _lib.ptr(ens.pbuf),
                  _lib.ptr(ens.eid), ens.plan.s)


def test_autotune_keeps_state_and_results():
    """Ensemble.autotune leaves the ensemble's state alone, picks one of the
    candidates, and the tuned ensemble steps to the same states as an
    untuned one."""
    from diffmpm_amd.ensemble import Ensemble
    g = _grav(_spec(*MEMBERS[0], "usl"))
    a = Ensemble(_members("usl", torch.float32))
    b = Ensemble(_members("usl", torch.float32))
    before = a.pbuf.clone()
    times = a.autotune(g, steps=2)
    assert set(times) == {"step", 64, 128, 256} and a.cells_per_block in (64, 128, 256)
    assert a.step_kernel == (min(times, key=times.get) == "step")
    assert torch.equal(a.pbuf, before)
    a.run(6, g)
    b.run(6, g)
    assert torch.equal(a.pbuf.view(torch.int32), b.pbuf.view(torch.int32))
    assert torch.equal(a.eid, b.eid)
