"""Rollouts captured into HIP graphs (diffmpm_amd.graphs.StepGraph).

A captured rollout must give, replay after replay, exactly the states of the
same steps run eagerly (dmpm_run with the step indices given by the host):
the replays read the step index from the plan's device step counter, which
a nodal force with a time function (ramp over the first five steps) makes
visible in the results.  Replays are interleaved with other eager work and
synchronisations, and every replay is checked.
"""
import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _spec(scheme, nx=60, ny=40, rows=30, seed=41):
    spec = cases.synthetic_spec("quad4", nx=nx, ny=ny, rows=rows, seed=seed, scheme=scheme,
                                traction=False, nodal_force=True, E=1e3, dt=1e-3, nsteps=12,
                                velocity_scale=0.05, drift=(0.6, 0.3))
    h = spec["el_len"][0]
    spec["particles"] = (np.asarray(spec["particles"]) * np.array([(nx - 6) / nx, 1.0])
                         + 3 * h).tolist()
    return spec


def _bits(t):
    t = t.clone()
    if t.is_floating_point():
        t = torch.where(torch.isnan(t), torch.full_like(t, float("nan")), t)
        t = t.view(torch.int32 if t.dtype == torch.float32 else torch.int64)
    return t


def _state(sim):
    ps = sim.mesh.particles[0]
    return [_bits(ps._buf), _bits(sim.mesh.elements.nodes._buf), ps._eid.clone()]


@pytest.mark.parametrize("nsteps", [3, 4])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_replays_equal_eager_steps(scheme, dtype, nsteps):
    """Three replays of nsteps captured steps (odd: the first step of each
    replay runs per phase) equal 3 * nsteps eager steps, bit for bit."""
    from diffmpm_amd._plan import plan_for
    from diffmpm_amd.graphs import StepGraph
    spec = _spec(scheme)
    a, b = cases.build_gpu(spec, dtype), cases.build_gpu(spec, dtype)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    sg = StepGraph(a, g, nsteps)
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    e0 = a.mesh.particles[0]._eid.clone()
    for r in range(3):
        sg.replay()
        torch.arange(4096, device="cuda", dtype=torch.float32).sum()  # other eager work
        pb.run(scheme, False, g, spec["dt"], r * nsteps, nsteps, b.keep_temporaries)
        torch.cuda.synchronize()
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), r
        assert sg.step == (r + 1) * nsteps
    assert not torch.equal(e0, a.mesh.particles[0]._eid)  # cells did change


def test_clock_matters():
    """Replays that repeated the captured step indices would differ: the
    time function of the nodal force changes over the first steps."""
    from diffmpm_amd._plan import plan_for
    spec = _spec("usf")
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    pa.run("usf", False, g, spec["dt"], 0, 4, True)
    pa.run("usf", False, g, spec["dt"], 4, 4, True)
    pb.run("usf", False, g, spec["dt"], 0, 4, True)
    pb.run("usf", False, g, spec["dt"], 0, 4, True)
    assert not torch.equal(_state(a)[1], _state(b)[1])


def test_ensemble_replays_equal_eager_steps():
    """A captured ensemble rollout (every member's tiles in one launch per
    step) equals the same steps run eagerly on the ensemble."""
    from diffmpm_amd.ensemble import Ensemble
    from diffmpm_amd.graphs import StepGraph
    sizes = [(60, 40, 30, 41), (20, 12, 9, 42), (37, 50, 20, 43)]
    ea = Ensemble([cases.build_gpu(_spec("usl", *m), keep_temporaries=False) for m in sizes])
    eb = Ensemble([cases.build_gpu(_spec("usl", *m), keep_temporaries=False) for m in sizes])
    g = torch.tensor(_spec("usl")["gravity"], dtype=torch.float64)
    sg = StepGraph(ea, g, 4)
    for r in range(3):
        sg.replay()
        torch.cuda.synchronize()
        eb.run(4, g, start_step=4 * r)
        torch.cuda.synchronize()
        for x, y in ((ea.pbuf, eb.pbuf), (ea.nbuf, eb.nbuf), (ea.eid, eb.eid)):
            assert torch.equal(_bits(x), _bits(y)), r
    assert sg.step == 12


def test_clock_flag_needs_the_fused_path():
    from diffmpm_amd import _lib
    from diffmpm_amd._lib import call, ptr
    from diffmpm_amd._plan import plan_for
    spec = cases.synthetic_spec("linear1d", nx=16, seed=3)
    sim = cases.build_gpu(spec)
    ps = sim.mesh.particles[0]
    p = plan_for(sim.mesh.elements, ps)
    with pytest.raises(ValueError, match="DMPM_RUN_CLOCK"):
        call("dmpm_run", p.h, ptr(ps._buf), ptr(ps._eid), ptr(sim.mesh.elements.nodes._buf),
             _lib.DMPM_USF, 0, p._gravity([-9.81]), 1e-5, 0, 2, _lib.DMPM_RUN_CLOCK, p.s)


@pytest.mark.parametrize("start", [0, 3])
def test_captured_rollout_writes_the_output_files(tmp_path, start):
    """A captured rollout with out_steps (graphs.StepGraph(..., out_steps=5)):
    three replays of 10 steps write, through NPZWriter, exactly the files and
    arrays that solve_jit writes for the same 30 steps (snapshots after every
    step i with i % 5 == 0), bit for bit, while the snapshots are copied to
    pinned memory inside the graph and written by the writer thread."""
    from diffmpm_amd.graphs import StepGraph
    from diffmpm_amd.writers import NPZWriter, npz_name
    spec = _spec("usf")
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    dirs = {k: tmp_path / k for k in ("graph", "eager")}
    sims = {}
    for k in dirs:
        sims[k] = cases.build_gpu(spec)
        sims[k].writer_func = NPZWriter().write
        sims[k].out_dir = str(dirs[k])
        sims[k].out_steps = 5
        sims[k].sim_steps = 30
    sg = StepGraph(sims["graph"], g, 10, start_step=start, out_steps=5)
    for _ in range(3):
        sg.replay()
        torch.arange(4096, device="cuda", dtype=torch.float32).sum()  # other eager work
    sg.wait_output()
    sims["eager"].solve_jit(g, start_step=start)
    torch.cuda.synchronize()
    want = sorted(npz_name(i, 30) for i in range(start, start + 30) if i % 5 == 0)
    for k, d in dirs.items():
        assert sorted(p.name for p in d.iterdir()) == want, k
    for name in want:
        with np.load(dirs["graph"] / name) as a, np.load(dirs["eager"] / name) as b:
            assert set(a.files) == {"loc", "velocity", "stress", "strain"}
            for f in a.files:
                assert a[f].shape == b[f].shape and np.array_equal(a[f], b[f]), (name, f)
    # the final states agree too
    for x, y in zip(_state(sims["graph"]), _state(sims["eager"])):
        assert torch.equal(x, y)
