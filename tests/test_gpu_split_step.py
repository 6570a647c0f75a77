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


@pytest.mark.parametrize("scheme", ["usf", "usl"])
def test_step_between_saved_equals_step_and_export(scheme):
    """dmpm_step_between_saved (previous node block read in place, fresh node
    block written, kinematic nodal velocity and sort exported by the step)
    gives the particle blocks, state tensors, node rows, kinematic nodal
    velocity and sort of dmpm_step_between on a node block prepared from the
    previous one followed by dmpm_plan_export_step, bit for bit, over five
    chained steps with a nonzero nodal damping force."""
    from diffmpm_amd import _lib
    from diffmpm_amd._lib import call, ptr
    from diffmpm_amd._plan import Plan
    spec = _moving_spec(scheme, drift=(0.6, 0.3))
    a, b = cases.build_gpu(spec), cases.build_gpu(spec)
    g = torch.tensor(spec["gravity"], dtype=torch.float64)
    code = _lib.DMPM_USF if scheme == "usf" else _lib.DMPM_USL
    runs = []
    for sim in (a, b):
        ps = sim.mesh.particles[0]
        nodes = sim.mesh.elements.nodes
        gen = torch.Generator(device="cpu").manual_seed(3)
        fd = 0.01 * torch.randn(nodes.f_damp.shape, generator=gen, dtype=torch.float64)
        nodes.f_damp = fd
        runs.append((Plan(sim.mesh.elements, ps), ps, nodes))
    n = len(runs[0][1])
    rows = [_lib.P_X, _lib.P_X + 1, _lib.P_V, _lib.P_V + 1, _lib.P_STRESS, _lib.P_STRESS + 1,
            _lib.P_STRESS + 3, _lib.P_VOLUME]
    st = [torch.stack([r[1]._buf[q, :n] for q in rows]).contiguous() for r in runs]
    blocks = [[r[1]._buf.clone(), r[1]._buf.clone()] for r in runs]
    prev = [r[2]._buf for r in runs]
    ncs = spec["nelements"][0] * spec["nelements"][1] + 2
    for k in range(5):
        outs = []
        for j, (plan, ps, nodes) in enumerate(runs):
            src, dst = blocks[j][k % 2], blocks[j][1 - k % 2]
            nb = torch.full_like(nodes._buf, float("nan"))
            sout = torch.empty_like(st[j])
            vkin = torch.empty((2, nodes._buf.shape[1]), dtype=nb.dtype, device=nb.device)
            perm = torch.empty(n, dtype=torch.int32, device=nb.device)
            cs = torch.empty(ncs, dtype=torch.int32, device=nb.device)
            flags = 2 if k else 0
            if j == 0:
                nb[_lib.N_V:_lib.N_A] = prev[j][_lib.N_V:_lib.N_A]
                nb[_lib.N_FDAMP:_lib.N_ROWS] = prev[j][_lib.N_FDAMP:_lib.N_ROWS]
                call("dmpm_step_between", plan.h, ptr(src), ptr(dst), ptr(ps._eid), ptr(nb),
                     code, 0, plan._gravity(g), float(spec["dt"]), k, flags, ptr(st[j]),
                     ptr(sout), n, plan.s)
                call("dmpm_plan_export_step", plan.h, ptr(vkin), ptr(perm), ptr(cs), plan.s)
            else:
                call("dmpm_step_between_saved", plan.h, ptr(src), ptr(dst), ptr(ps._eid),
                     ptr(prev[j]), ptr(nb), code, 0, plan._gravity(g), float(spec["dt"]), k,
                     flags, ptr(st[j]), ptr(sout), n, ptr(vkin), ptr(perm), ptr(cs), plan.s)
            outs.append((dst, sout, nb, vkin, perm, cs))
            prev[j] = nb
            st[j] = sout
        (d0, s0, n0, v0, p0, c0), (d1, s1, n1, v1, p1, c1) = outs
        nn = runs[0][2]._count
        assert torch.equal(s0.view(torch.int32), s1.view(torch.int32)), k
        assert torch.equal(d0[:, :n].view(torch.int32), d1[:, :n].view(torch.int32)), k
        assert torch.equal(n0[_lib.N_V:, :nn].view(torch.int32),
                           n1[_lib.N_V:, :nn].view(torch.int32)), k
        assert torch.equal(v0.view(-1)[: 2 * nn].view(torch.int32),
                           v1.view(-1)[: 2 * nn].view(torch.int32)), k
        assert torch.equal(p0, p1) and torch.equal(c0, c1), k
