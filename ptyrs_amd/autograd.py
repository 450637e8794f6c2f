"""The reconstruction cost as a torch.autograd.Function whose backward is the engine's
hand-written adjoint (HIP, on the engine's device): the forward model object window → probe
multiply → multislice propagation → exit transform → Poisson (or amplitude) cost, its gradient
with respect to every object slice and the probe. Torch tensors are the interface; the engine
holds the scan and computes on the GPU (there is no CPU path).

    loss = ptyrs_amd.autograd.cost(engine, objects, probe)   # objects [nslice, nobj, nyo, nxo]
    loss.backward()

ptyrs/src/engine/ptycho.rs:492-561 (gradients), ptyrs/src/engine/multislice.rs:266-442 (their
multislice adjoint).
"""

import numpy as np
import torch


class Cost(torch.autograd.Function):
    """value = L(objects, probe) over every frame of `engine`'s scan; backward returns
    grad_output · 2·∂L/∂z* for each complex input, torch's convention for a real loss."""

    @staticmethod
    def forward(ctx, engine, objects, probe, cost="poisson"):
        objs = objects.detach().to("cpu", torch.complex128).numpy()
        engine.set_object(objs if engine.nslice > 1 else objs.reshape(objs.shape[-3:]))
        engine.set_probe(probe.detach().to("cpu", torch.complex128).numpy())
        value, go, gp = engine.grad(cost)
        ctx.save_for_backward(torch.from_numpy(go.reshape(objects.shape)).to(objects),
                              torch.from_numpy(gp).to(probe))
        return torch.tensor(value, dtype=torch.float64, device=objects.device)

    @staticmethod
    def backward(ctx, grad_output):
        go, gp = ctx.saved_tensors
        g = grad_output.to(go.real.dtype)
        return None, g * go, g * gp, None


def cost(engine, objects, probe, cost="poisson"):
    """The scan's cost at `objects` (`[nslice, nobj, nyo, nxo]` or `[nobj, nyo, nxo]`) and
    `probe` (`[nprobe, ny, nx]`), complex tensors, differentiable in both."""
    return Cost.apply(engine, objects, probe, cost)
