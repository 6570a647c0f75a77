"""Per-kernel-class timing (dmpm_timing_enable / dmpm_timing_read_class), the
source of bench.py's roofline durations.

The one-launch step class is timed by the start and stop events of
k_step_fast's own dispatch; the re-sort class by events around its four
launches.  Checked: every timed launch is counted, the durations are
positive, the step class's time per launch lies within the wall time of a
step, and timing leaves the results unchanged (bit for bit).
"""
import pytest
import torch

import cases

pytestmark = pytest.mark.gpu


def _plan(dtype):
    from diffmpm_amd._plan import plan_for
    spec = cases.synthetic_spec("quad4", nx=64, ny=48, rows=40, seed=7, scheme="usf",
                                traction=False, nodal_force=True, E=1e3, dt=1e-3, nsteps=10,
                                velocity_scale=0.05)
    sim = cases.build_gpu(spec, dtype, keep_temporaries=False)
    ps = sim.mesh.particles[0]
    return sim, plan_for(sim.mesh.elements, ps), torch.tensor(spec["gravity"],
                                                              dtype=torch.float64)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_step_and_sort_classes_count_and_time_their_launches(dtype):
    from diffmpm_amd._plan import timing_enable, timing_read
    sim_a, plan_a, g = _plan(dtype)
    sim_b, plan_b, _ = _plan(dtype)
    K = 12
    plan_a.run("usf", False, g, sim_a.dt, 0, 2, keep_temporaries=False)
    plan_b.run("usf", False, g, sim_b.dt, 0, 2, keep_temporaries=False)
    torch.cuda.synchronize()
    timing_enable(plan_a, "all")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    plan_a.run("usf", False, g, sim_a.dt, 2, K, keep_temporaries=False)
    e1.record()
    torch.cuda.synchronize()
    step_ms, step_n = timing_read(plan_a, "step")
    sort_ms, sort_n = timing_read(plan_a, "sort")
    timing_enable(plan_a, None)
    wall_ms = e0.elapsed_time(e1)
    assert step_n == K and sort_n >= K, (step_n, sort_n)
    assert 0 < step_ms and 0 < sort_ms
    assert step_ms / step_n < wall_ms / K, (step_ms / step_n, wall_ms / K)
    # the same steps without events give the same state
    plan_b.run("usf", False, g, sim_b.dt, 2, K, keep_temporaries=False)
    torch.cuda.synchronize()
    pa, pb = sim_a.mesh.particles[0], sim_b.mesh.particles[0]
    for name in ("loc", "velocity", "stress", "volume"):
        assert torch.equal(getattr(pa, name), getattr(pb, name)), name
