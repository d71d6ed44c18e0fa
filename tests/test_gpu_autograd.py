"""The adjoint behind torch.autograd (avbd.autograd.rollout).

The Function's backward is the C ABI's avbd_adjoint_backward: on the
differentiable configuration (64 boxes on the 30-degree ramp, 120 steps) the
gradients .backward() leaves in the leaf tensors equal the C ABI's (to the
rounding of the reverse pass's atomic additions), for a loss written in torch. On boxes side by side in the sliding
regime, torch.autograd.gradcheck compares the Function's Jacobian with its own
central differences of GPU rollouts (exact precision: each rollout is the f64
oracle's trajectory bit for bit).
"""
import numpy as np
import pytest

import avbd

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def leaves(scene, device="cuda:0"):
    mu = torch.tensor([b["friction"] for b in scene["bodies"]], dtype=torch.float64, device=device,
                      requires_grad=True)
    v0 = torch.tensor([b.get("vel", [0, 0, 0]) for b in scene["bodies"]], dtype=torch.float64, device=device,
                      requires_grad=True)
    return mu, v0


def test_backward_equals_the_c_abi_gradients():
    from avbd.autograd import rollout
    scene, params, meta = avbd.scenes.ramp(64)
    target = torch.tensor(meta["targets"], dtype=torch.float64, device="cuda:0")
    mu, v0 = leaves(scene)
    be = avbd.HipBackend(precision=avbd.PREC_EXACT)
    x = rollout(scene, params, mu, v0, meta["steps"], 10, backend=be)
    loss = ((x[2:] - target[2:]) ** 2).sum()
    loss.backward()
    # the same through the C ABI directly
    ref = avbd.HipBackend(precision=avbd.PREC_EXACT)
    ref.pack(scene, params)
    ref.adjoint_forward(meta["steps"], 10)
    xr = ref.read_state()["pos"]
    assert np.array_equal(x.detach().cpu().numpy(), xr)
    dL = 2 * (xr - np.array(meta["targets"]))
    dL[:2] = 0
    gmu, gv0 = ref.adjoint_backward(dL)
    # the reverse pass adds neighbours' pose adjoints with atomics, whose
    # order varies from run to run: two backward passes of the same recording
    # agree to rounding, not bit for bit
    assert np.abs(mu.grad.cpu().numpy() - gmu).max() <= 1e-9 * np.abs(gmu).max()
    assert np.abs(v0.grad.cpu().numpy() - gv0).max() <= 1e-9 * np.abs(gv0).max()
    assert torch.isfinite(mu.grad).all() and mu.grad[2:].abs().max() > 0


def test_gradcheck_in_the_sliding_regime():
    from avbd.autograd import rollout
    scene, params, meta = avbd.scenes.ramp_lanes(n=4)
    be = avbd.HipBackend(precision=avbd.PREC_EXACT)
    mu, v0 = leaves(scene)

    def f(m, v):
        # the four boxes' final positions (the statics do not move)
        return rollout(scene, params, m, v, 30, 10, backend=be)[2:]

    assert torch.autograd.gradcheck(f, (mu, v0), eps=1e-6, atol=1e-5, rtol=1e-4)
