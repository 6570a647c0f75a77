"""dmpm_step_between_noalloc (the step with its workspaces in the library's
static device array) against dmpm_step_between, through the C ABI: the same
output blocks, node blocks and element ids bit for bit, on two calls in a
row (the second on a smaller plan, so the shared array is reused with other
sizes), and a plan too large for the array is rejected."""
import numpy as np
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _spec(nx, ny, rows, seed, scheme):
    return cases.synthetic_spec("quad4", nx=nx, ny=ny, rows=rows, seed=seed, scheme=scheme,
                                traction=False, nodal_force=True, E=1e3, dt=1e-3, nsteps=1,
                                velocity_scale=0.05, drift=(0.6, 0.3))


def _bits(t):
    t = t.clone()
    t = torch.where(torch.isnan(t), torch.full_like(t, float("nan")), t)
    return t.view(torch.int32 if t.dtype == torch.float32 else torch.int64)


def _step(sim, entry, step, scheme):
    from diffmpm_amd import _lib
    from diffmpm_amd._lib import call, ptr
    from diffmpm_amd._plan import plan_for
    ps = sim.mesh.particles[0]
    plan = plan_for(sim.mesh.elements, ps)
    plan.refresh_tables()
    pout = ps._buf.clone()
    code = _lib.DMPM_USF if scheme == "usf" else _lib.DMPM_USL
    call(entry, plan.h, ptr(ps._buf), ptr(pout), ptr(ps._eid),
         ptr(sim.mesh.elements.nodes._buf), code, 0, plan._gravity(sim_gravity(sim)),
         float(sim.dt), int(step), 1, None, None, 0, plan.s)
    return pout


def sim_gravity(sim):
    return torch.tensor([0.0, -9.81], dtype=torch.float64)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_noalloc_step_equals_step_between(scheme, dtype):
    for (nx, ny, rows, seed) in ((90, 60, 45, 51), (20, 12, 9, 52)):  # larger, then smaller
        spec = _spec(nx, ny, rows, seed, scheme)
        a, b = cases.build_gpu(spec, dtype), cases.build_gpu(spec, dtype)
        for k in range(2):  # two steps each: the second from the first's output
            outs = []
            for sim, entry in ((a, "dmpm_step_between"), (b, "dmpm_step_between_noalloc")):
                pout = _step(sim, entry, k, scheme)
                sim.mesh.particles[0]._buf.copy_(pout)
                outs.append(pout)
            torch.cuda.synchronize()
            n = len(a.mesh.particles[0])
            assert torch.equal(_bits(outs[0][:, :n]), _bits(outs[1][:, :n])), (nx, k)
            assert torch.equal(_bits(a.mesh.elements.nodes._buf),
                               _bits(b.mesh.elements.nodes._buf)), (nx, k)
            assert torch.equal(a.mesh.particles[0]._eid, b.mesh.particles[0]._eid), (nx, k)
        assert torch.isfinite(outs[1][0:4, :n]).all()


def test_noalloc_rejects_a_plan_beyond_the_array():
    """A 1500 x 1500 mesh (2.25 M nodes, 54 MB of workspaces in FP32) is
    above the 32 MiB array: DMPM_EINVAL, nothing enqueued."""
    from diffmpm_amd import _lib
    from diffmpm_amd._lib import call, ptr
    from diffmpm_amd._plan import plan_for
    n = 1500
    spec = _spec(n, n, 1, 53, "usf")
    spec["particles"] = spec["particles"][:64]
    spec["init_velocity"] = spec["init_velocity"][:64]
    sim = cases.build_gpu(spec)
    ps = sim.mesh.particles[0]
    plan = plan_for(sim.mesh.elements, ps)
    pout = ps._buf.clone()
    with pytest.raises(ValueError, match="scratch"):
        call("dmpm_step_between_noalloc", plan.h, ptr(ps._buf), ptr(pout), ptr(ps._eid),
             ptr(sim.mesh.elements.nodes._buf), _lib.DMPM_USF, 0,
             plan._gravity([0.0, -9.81]), 1e-5, 0, 0, None, None, 0, plan.s)
