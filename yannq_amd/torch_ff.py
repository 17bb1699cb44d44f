"""torch.autograd wrapper of the FeedForward machine's HIP kernels.

``FFMachine(ctx_device, n, channels, kernel, hidden, dtype)`` holds the
network of Machines/FeedForward.hpp used in the C4 configuration (or, with
``layers``, any stack of Conv1D / FullyConnected / activation / Mean layers)
and exposes
``logpsi(sigma)`` whose backward pass is the hand-written per-sample reverse
sweep (``yq_ff_vjp_dev``), so ``torch.autograd.gradcheck`` and finite
differences can be run against it.  The forward stores the layer
activations for the backward while ``batch * npar * itemsize`` stays within
``act_budget`` bytes and recomputes them inside the backward kernel beyond.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Context


class _LogPsi(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, sigma, machine):
        m = machine
        m.ctx.ff_set_params(params.detach().to("cpu", torch.float64).numpy())
        ns = sigma.shape[0]
        out = torch.empty(ns, dtype=m.dtype, device=sigma.device)
        store = ns * m.npar * out.element_size() <= m.act_budget
        act = torch.empty((ns, m.act_width), dtype=m.dtype, device=sigma.device) if store else None
        torch.cuda.synchronize(sigma.device)
        m.ctx.ff_forward_dev(ns, sigma.data_ptr(), out.data_ptr(),
                             act.data_ptr() if act is not None else None)
        ctx.save_for_backward(params, sigma)
        ctx.act = act
        ctx.machine = m
        m.last_stored = store
        return out

    @staticmethod
    def backward(ctx, g):
        params, sigma = ctx.saved_tensors
        m = ctx.machine
        m.ctx.ff_set_params(params.detach().to("cpu", torch.float64).numpy())
        r = g.detach().to(sigma.device, m.dtype).contiguous()
        grad = torch.empty(m.npar, dtype=m.dtype, device=sigma.device)
        torch.cuda.synchronize(sigma.device)
        m.ctx.ff_vjp_dev(sigma.shape[0], sigma.data_ptr(),
                         ctx.act.data_ptr() if ctx.act is not None else None,
                         r.data_ptr(), grad.data_ptr())
        return grad.to(params.device, params.dtype), None, None


class FFMachine:
    """Conv1D(1->channels, kernel) -> LnCosh -> FC(channels*n -> hidden) ->
    Tanh -> Mean on the GPU, or with ``layers`` (tuples (kind, a, b, c, d,
    use_bias, param) of yq_ff_layer, in addLayer order) any layer stack;
    parameters live in ``self.params``."""

    def __init__(self, n, channels=None, kernel=None, hidden=None, dtype=torch.float32,
                 device=0, act_budget=1 << 30, seed=None, init_mode=3, ctx=None, layers=None):
        self.ctx = ctx if ctx is not None else Context(device)
        self.dtype = dtype
        npdt = np.float64 if dtype == torch.float64 else np.float32
        if layers is not None:
            self.ctx.ff_create_stack(n, layers, npdt)
        else:
            self.ctx.ff_create(n, channels, kernel, hidden, npdt)
        self.n, self.npar = n, self.ctx.ff_npar
        self.act_width = self.ctx.ff_act_len()
        self.act_budget = act_budget
        self.last_stored = None
        if seed is not None:
            self.ctx.ff_init_random(seed, init_mode)
        self.params = torch.tensor(self.ctx.ff_get_params(), dtype=dtype,
                                   device=f"cuda:{device}", requires_grad=True)

    def logpsi(self, sigma, params=None):
        """log psi of a batch of int8 configurations on the GPU."""
        sigma = sigma.to(torch.int8).contiguous()
        return _LogPsi.apply(self.params if params is None else params, sigma, self)

    def jacobian(self, sigma):
        """Per-sample d log psi / d params (ns x npar) without autograd."""
        sigma = sigma.to(torch.int8).contiguous()
        self.ctx.ff_set_params(self.params.detach().to("cpu", torch.float64).numpy())
        J = torch.empty((sigma.shape[0], self.npar), dtype=self.dtype, device=sigma.device)
        torch.cuda.synchronize(sigma.device)
        self.ctx.ff_jacobian_dev(sigma.shape[0], sigma.data_ptr(), J.data_ptr())
        return J
