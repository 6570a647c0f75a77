"""The split step (DMPM_OPT_STEP_KERNEL = 2: kinematic P2G, stress and force
P2G, G2P as three launches over node and cell tiles) against the oracle and
against the other step drivers.

Runs of an even number of steps go through the step kernels only (an
odd-length run takes its first step with the per-phase launches), so these
tests run the GPU two steps per call.
"""
import numpy as np
import pytest
import torch

import cases
from test_gpu_parity import RTOL_F32, RTOL_F64, _compare, _moving_spec, _state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_step_kernels_two_step_oracle_parity(scheme, dtype, kind):
    """The one-launch (kind 1) and the split step (kind 2), re-synchronised to
    the oracle before every pair of steps of the multi-tile moving case,
    match two oracle steps within RTOL_F32 / RTOL_F64 (particle and nodal
    fields, element ids)."""
    from diffmpm_amd import _lib as L
    from diffmpm_amd._plan import plan_for
    spec = _moving_spec(scheme)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    rtol = RTOL_F64 if dtype == torch.float64 else RTOL_F32
    osim = cases.build_oracle(spec, npdt)
    gsim = cases.build_gpu(spec, dtype)
    plan = plan_for(gsim.mesh.elements, gsim.mesh.particles[0])
    plan.set_option(L.DMPM_OPT_STEP_KERNEL, kind)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    for i in range(0, spec["nsteps"] - 1, 2):
        cases.load_oracle_state_into_gpu(osim, gsim)
        osim.step(i, np.asarray(spec["gravity"]))
        osim.step(i + 1, np.asarray(spec["gravity"]))
        plan.run(scheme, False, g, spec["dt"], i, 2, True)
        _compare(osim, gsim, rtol, f"steps {i}, {i + 1}")


@pytest.mark.parametrize("vu", [False, True])
@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_split_step_equals_one_launch_step(scheme, vu):
    """The split step and the one-launch step give bit-identical particle and
    node states over 10 steps of the moving multi-tile case (runs of 2 and
    4 steps, velocity update on and off)."""
    from diffmpm_amd import _lib as L
    from diffmpm_amd._plan import plan_for
    spec = _moving_spec(scheme, drift=(0.6, 0.3), velocity_scale=0.05)
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    pa = plan_for(a.mesh.elements, a.mesh.particles[0])
    pb = plan_for(b.mesh.elements, b.mesh.particles[0])
    pa.set_option(L.DMPM_OPT_STEP_KERNEL, 2)
    pb.set_option(L.DMPM_OPT_STEP_KERNEL, 1)
    k = 0
    for n in (2, 4, 2, 2):
        pa.run(scheme, vu, g, spec["dt"], k, n, True)
        pb.run(scheme, vu, g, spec["dt"], k, n, True)
        k += n
        for x, y in zip(_state(a), _state(b)):
            assert torch.equal(x, y), k


def test_step_kernel_option_range():
    """DMPM_OPT_STEP_KERNEL takes 0, 1 or 2; other values are rejected and
    leave the plan's choice as it was."""
    from diffmpm_amd import _lib as L
    from diffmpm_amd._plan import plan_for
    spec = cases.synthetic_spec("quad4", nx=8, ny=6, rows=4, scheme="usf", nsteps=2, seed=1)
    gsim = cases.build_gpu(spec)
    plan = plan_for(gsim.mesh.elements, gsim.mesh.particles[0])
    for v in (0, 1, 2):
        plan.set_option(L.DMPM_OPT_STEP_KERNEL, v)
    for v in (-1, 3):
        with pytest.raises(Exception, match="step kernel"):
            plan.set_option(L.DMPM_OPT_STEP_KERNEL, v)
