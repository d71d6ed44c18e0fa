"""Differentiable rollouts: a torch.autograd.Function over the step's adjoint.

``rollout(scene, params, friction, velocity, steps)`` packs the scene with each
body's friction coefficient and initial linear velocity taken from the two
tensors, steps it ``steps`` times on the GPU keeping a checkpoint every
``ckpt_every`` steps (avbd_adjoint_forward), and returns the final positions
as an (n, 3) float64 tensor. Its backward hands the incoming gradient of
those positions to avbd_adjoint_backward, the hand-written reverse of the
step, and returns the gradients with respect to ``friction`` (n,) and
``velocity`` (n, 3). The forward and the backward are exactly the C ABI's
pair; the wrapper adds only the conversion between tensors and host arrays.

The adjoint differentiates the computed trajectory (a fixed iteration count)
with the contact geometry held piecewise constant; it needs the exact or f64
precision, box contacts and no joints (include/avbd.h). There is no
counterpart in the reference, which has no adjoint.

Importing this module imports torch; the rest of the package does not need it.
"""
import copy

import numpy as np
import torch

from . import _lib as L
from .backend import HipBackend


def _record(backend, scene, params, mu, v0, steps, ckpt_every):
    """Packs the scene with these frictions and initial velocities and runs
    the recording forward; returns the final positions."""
    sc = copy.copy(scene)
    sc["bodies"] = [dict(b, friction=float(mu[i]), vel=[float(x) for x in v0[i]])
                    for i, b in enumerate(scene["bodies"])]
    backend.pack(sc, params)
    backend.adjoint_forward(int(steps), int(ckpt_every))
    backend._avbd_recording = getattr(backend, "_avbd_recording", 0) + 1
    return backend.read_state()["pos"]


class Rollout(torch.autograd.Function):
    """forward(friction, velocity, backend, scene, params, steps, ckpt_every)
    -> final positions (n, 3); backward -> (d friction, d velocity).

    The context's recording is what the backward reverses. When the backend
    has recorded another rollout since this one (a caller that evaluates the
    function again before differentiating, as finite-difference checks do),
    the backward first records this rollout again from its saved inputs: the
    forward is deterministic, so it is the same trajectory, bit for bit."""

    @staticmethod
    def forward(ctx, friction, velocity, backend, scene, params, steps, ckpt_every):
        n = len(scene["bodies"])
        mu = friction.detach().to("cpu", torch.float64).numpy().reshape(n).copy()
        v0 = velocity.detach().to("cpu", torch.float64).numpy().reshape(n, 3).copy()
        pos = _record(backend, scene, params, mu, v0, steps, ckpt_every)
        ctx.backend = backend
        ctx.recording = backend._avbd_recording
        ctx.inputs = (scene, params, mu, v0, steps, ckpt_every)
        ctx.mu_like = (friction.device, friction.dtype)
        ctx.v_like = (velocity.device, velocity.dtype)
        return torch.as_tensor(pos, dtype=torch.float64, device=friction.device)

    @staticmethod
    def backward(ctx, grad_pos):
        if getattr(ctx.backend, "_avbd_recording", None) != ctx.recording:
            _record(ctx.backend, *ctx.inputs)
            ctx.recording = ctx.backend._avbd_recording
        g = np.ascontiguousarray(grad_pos.detach().to("cpu", torch.float64).numpy())
        dmu, dv0 = ctx.backend.adjoint_backward(g)
        gm = torch.as_tensor(dmu, device=ctx.mu_like[0]).to(ctx.mu_like[1]) if ctx.needs_input_grad[0] else None
        gv = torch.as_tensor(dv0, device=ctx.v_like[0]).to(ctx.v_like[1]) if ctx.needs_input_grad[1] else None
        return gm, gv, None, None, None, None, None


def rollout(scene, params, friction, velocity, steps, ckpt_every=10, backend=None):
    """Final positions (n, 3) of `steps` steps of `scene` with per-body
    friction (n,) and initial velocity (n, 3) taken from the tensors;
    differentiable with respect to both. `backend` (a HipBackend of
    precision exact or f64) is created when not given; pass one to reuse its
    device buffers across calls, as an optimisation loop does."""
    if backend is None:
        backend = HipBackend(precision=L.PREC_F64)
    return Rollout.apply(friction, velocity, backend, scene, params, steps, ckpt_every)
