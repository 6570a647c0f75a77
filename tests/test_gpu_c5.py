"""The uneven 64-member ensemble at full size (``workloads.ensemble(64,
seed=0)``: 2,390,456 particles, n = 32..256 cells per side, USL, the
configuration bench.py's batched leg times), through the batch plan:

* after 20 steps every member equals the same member stepped alone with its
  own plan, bit for bit (FP32);
* the smallest members (n <= 40) stepped alone match the FP32 oracle after
  every step within 1e-5, and in FP64 the batch matches the FP64 oracle
  within 1e-12 (diffmpm/solver.py:179-213, USL scheme.py:107-133);
* the batch adjoint's per-member d loss / d log E of two members matches
  central differences of the FP64 oracle's forward (relative step 1e-4)
  within 1e-4.
"""
import math

import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu

FIELDS = ("loc", "velocity", "stress", "strain", "volume", "density")


def _oracle_member(m, dtype):
    """The oracle's copy of one member (workloads.block_2d(n, rows, jitter=0,
    E, nu): Gauss points, bottom fixed, side rollers, rho = 1000, USL)."""
    from oracle import mpm as om
    from diffmpm_amd.workloads import gauss_points
    n, rows = m["n"], m["rows"]
    h = 1.0 / n
    pts = gauss_points(n, rows, h, 0.0, 0)
    b, j = np.arange(n + 1), np.arange(n + 1)
    cons = [(b, om.Constraint(0, 0.0)), (b, om.Constraint(1, 0.0)),
            (np.concatenate([j * (n + 1), j * (n + 1) + n]), om.Constraint(0, 0.0))]
    el = om.Elements("quad4", (n, n), (h, h), cons, [], dtype)
    ps = om.Particles(pts, om.LinearElastic({"density": 1000.0, "youngs_modulus": m["E"],
                                             "poisson_ratio": m["nu"]}), dtype=dtype)
    return om.MPMExplicit(om.Mesh(el, [ps]), 1e-5, "usl", False, 1)


def _field_errs(osim, ps):
    ops = osim.mesh.particles[0]
    return {f: cases.normwise(cases.gpu_particle_field(ps, f), cases.oracle_particle_field(ops, f))
            for f in FIELDS}


def test_c5_batch_equals_members_alone_and_small_members_match_the_oracle():
    from diffmpm_amd import workloads
    from diffmpm_amd._plan import plan_for
    from diffmpm_amd.ensemble import Ensemble
    K = 20
    alone, g, specs = workloads.ensemble(64, seed=0)
    batch, _, _ = workloads.ensemble(64, seed=0)
    assert sum(len(s.mesh.particles[0]) for s in batch) == 2390456
    ens = Ensemble(batch)
    plans = [plan_for(s.mesh.elements, s.mesh.particles[0]) for s in alone]
    small = [i for i, m in enumerate(specs) if m["n"] <= 40]
    assert len(small) >= 4
    oracles = {i: _oracle_member(specs[i], np.float32) for i in small}
    gv = np.array([0.0, -9.81])
    worst = 0.0
    for k in range(K):
        if k % 2 == 0:  # the batch in runs of two steps (the one-launch step)
            ens.run(2, g, start_step=k)
        for p, s in zip(plans, alone):
            p.run("usl", False, g, s.dt, k, 1, keep_temporaries=False)
        for i in small:
            oracles[i].step(k, gv)
        torch.cuda.synchronize()
        for i in small:
            errs = _field_errs(oracles[i], alone[i].mesh.particles[0])
            bad = {f: e for f, e in errs.items() if not e <= 1e-5}
            assert not bad, (k, i, specs[i], bad)
            worst = max(worst, max(errs.values()))
    ens.scatter()
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(alone, batch)):
        pa, pb = a.mesh.particles[0], b.mesh.particles[0]
        n = len(pa)
        assert torch.equal(pa._buf[:, :n].view(torch.int32), pb._buf[:, :n].view(torch.int32)), i
        assert torch.equal(pa._eid[:n], pb._eid[:n]), i
        na, nb = a.mesh.elements.nodes, b.mesh.elements.nodes
        m = na._count
        assert torch.equal(na._buf[:, :m].view(torch.int32), nb._buf[:, :m].view(torch.int32)), i
    print(f"C5: {len(small)} small members vs FP32 oracle over {K} steps, worst {worst:.1e}")


def test_c5_fp64_batch_matches_the_fp64_oracle():
    from diffmpm_amd import workloads
    from diffmpm_amd.ensemble import Ensemble
    batch, g, specs = workloads.ensemble(64, seed=0, dtype=torch.float64)
    ens = Ensemble(batch)
    small = [i for i, m in enumerate(specs) if m["n"] <= 40]
    oracles = {i: _oracle_member(specs[i], np.float64) for i in small}
    gv = np.array([0.0, -9.81])
    for k in range(0, 10, 2):
        ens.run(2, g, start_step=k)
        for i in small:
            oracles[i].step(k, gv)
            oracles[i].step(k + 1, gv)
        ens.scatter()
        torch.cuda.synchronize()
        for i in small:
            errs = _field_errs(oracles[i], batch[i].mesh.particles[0])
            bad = {f: e for f, e in errs.items() if not e <= 1e-12}
            assert not bad, (k, i, bad)


def _oracle_state(osim):
    ps = osim.mesh.particles[0]
    return np.stack([ps.loc[:, 0], ps.loc[:, 1], ps.velocity[:, 0], ps.velocity[:, 1],
                     ps.stress[:, 0], ps.stress[:, 1], ps.stress[:, 3], ps.volume])


def test_c5_member_log_e_gradients_match_oracle_central_differences():
    """d L / d log E_i of two members from the batch adjoint (per-member
    scalar E through Ensemble.sim_index) against central differences of the
    FP64 oracle's forward rollout of that member, L = sum of w * final state
    over the member's particles (w seeded)."""
    from diffmpm_amd import workloads
    from diffmpm_amd.autodiff import Differentiable, rollout
    from diffmpm_amd.ensemble import Ensemble
    K, eps = 4, 1e-4
    batch, g, specs = workloads.ensemble(64, seed=0, dtype=torch.float64)
    ens = Ensemble(batch, per_particle_E=True)
    d = Differentiable.of_ensemble(ens)
    le = torch.tensor([math.log(m["E"]) for m in specs], dtype=torch.float64, device="cuda",
                      requires_grad=True)
    out = rollout(d, K, g, state=d.get_state(), log_e=le[ens.sim_index()])
    picks = sorted(range(len(specs)), key=lambda i: specs[i]["n"] * specs[i]["rows"])[:2]
    rng = np.random.Generator(np.random.PCG64(7))
    ws = {i: rng.standard_normal((8, ens.counts[i])) for i in picks}
    loss = sum((torch.from_numpy(ws[i]).cuda() * out[:, ens.sim_slice(i)]).sum() for i in picks)
    loss.backward()
    gv = np.array([0.0, -9.81])
    for i in picks:
        vals = []
        for sgn in (1, -1):
            m = dict(specs[i])
            m["E"] = specs[i]["E"] * math.exp(sgn * eps)
            osim = _oracle_member(m, np.float64)
            for k in range(K):
                osim.step(k, gv)
            vals.append(float((ws[i] * _oracle_state(osim)).sum()))
        fd = (vals[0] - vals[1]) / (2 * eps)
        got = float(le.grad[i])
        assert math.isclose(got, fd, rel_tol=1e-4), (i, specs[i], got, fd)
