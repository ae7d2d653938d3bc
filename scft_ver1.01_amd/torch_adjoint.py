"""Optional torch.autograd.Function over the C-ABI loop adjoint.

    loss = SCFTLoopLoss.apply(h, solver, w0A, w0B, n_iter, target)
    loss.backward()          # h.grad = dloss/dh from scft_adjoint_loop

The forward runs the unrolled SCFT loop through the HIP library and keeps the
discrete-adjoint gradient; backward scales it by the incoming gradient.  torch
is only the autograd plumbing here: tensors are moved to host float64 for the
C ABI, and the gradient comes back on h's device and dtype.
"""
import numpy as np
import torch


class SCFTLoopLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h, solver, w0A, w0B, n_iter, target):
        def host(t):
            return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy().ravel())

        loss, grad, _ = solver.adjoint_loop(host(h), host(w0A), host(w0B), int(n_iter), host(target), grad=True)
        ctx.save_for_backward(torch.from_numpy(grad.reshape(h.shape)).to(h.device, h.dtype))
        return torch.tensor(float(loss.sum()), dtype=h.dtype, device=h.device)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return gout * g, None, None, None, None, None
