"""The oracle's general FeedForward stack (oracle/oracle_ffs.cpp) pinned
against an independent torch (CPU, float64) build of the same network --
Conv1D as a circularly padded torch conv1d, FullyConnected as a matrix
product, torch's own activations and means -- for the forward value and,
through torch autograd, the Jacobian rows of FeedForward::backward
(Machines/FeedForward.hpp:186-205); and against central finite differences
of the oracle's own forward.

Sigmoid is the exception: the reference's activations.hpp:211-230 discards
both inverse() expressions, so its value is exp(z) + 1 (pinned here against
that closed form) and its ApplyJacobian returns (1 - a) a without the
incoming gradient, which is not a derivative of anything: that backward is
the reference's literal formula and parity rests on reading it.
"""
import math

import numpy as np
import pytest
import torch

import oracle as orc

from ffstacks import C1, FC, ACT, MEAN, STACKS  # noqa: E402


def torch_forward(n, layers, p, sigma):
    """The network in torch, from the layer definitions (layers/Conv1D.hpp
    forward: output r of channel oc reads sites (r*stride + ki - K/2) mod
    inSize of every input channel; FullyConnected y = W^T x + b with W
    inputDim x outputDim column-major; Mean over the axis of an a x b
    column-major view)."""
    x = torch.as_tensor(sigma, dtype=torch.float64)
    off = 0
    for (kind, a, b, c, d, bias, prm) in layers:
        if kind == C1:
            ic, oc, K, st = a, b, c, d
            ins = x.numel() // ic
            w = p[off:off + ic * K * oc].reshape(oc, ic, K)
            off += ic * K * oc
            bb = None
            if bias:
                bb = p[off:off + oc]
                off += oc
            idx = (torch.arange(ins + K - 1) - K // 2) % ins
            xp = x.reshape(ic, ins)[:, idx]
            y = torch.nn.functional.conv1d(xp[None], w, bb, stride=st)[0]
            x = y[:, :ins // st].reshape(-1)
        elif kind == FC:
            W = p[off:off + a * b].reshape(b, a)
            off += a * b
            x = W @ x
            if bias:
                x = p[off:off + b] + x
                off += b
        elif kind == ACT:
            if a == orc.A_LNCOSH:
                x = torch.log(torch.cosh(x))
            elif a == orc.A_TANH:
                x = prm * torch.tanh(x / prm)
            elif a == orc.A_WEAKTANH:
                x = x - torch.tanh(x) / 2
            elif a == orc.A_SIGMOID:
                x = torch.exp(x) + 1
            elif a == orc.A_RELU:
                x = torch.relu(x)
            elif a == orc.A_LEAKYRELU:
                x = torch.nn.functional.leaky_relu(x, prm)
            elif a == orc.A_HARDTANH:
                x = torch.nn.functional.hardtanh(x)
            elif a == orc.A_SOFTSHRINK:
                x = torch.nn.functional.softshrink(x, 1.0)
            elif a == orc.A_LEAKYHARDTANH:
                x = torch.nn.functional.hardtanh(x) + prm * (x - torch.clamp(x, -1, 1))
            elif a == orc.A_SOFTSIGN:
                x = torch.nn.functional.softsign(x)
            elif a == orc.A_COS:
                x = prm * torch.cos(x * math.pi / prm)
        else:
            m = x.reshape(b, a).T  # a x b, column-major view
            x = m.mean(0) if c == 0 else m.mean(1)
    assert off == p.numel()
    return x[0]


def params_for(desc, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, orc.ff_npar(desc))


@pytest.mark.parametrize("name", sorted(STACKS))
def test_stack_forward_and_jacobian_vs_torch(name):
    n, layers = STACKS[name]
    desc = orc.ff_stack(n, layers)
    p = params_for(desc, 7)
    rng = np.random.default_rng(11)
    sigma = rng.choice(np.array([-1, 1], dtype=np.int8), size=(6, n))
    J, out = orc.ff_jacobian(desc, p, sigma)
    assert np.allclose(orc.ff_forward(desc, p, sigma), out, rtol=0, atol=0)
    for s in range(sigma.shape[0]):
        pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
        y = torch_forward(n, layers, pt, sigma[s])
        (g,) = torch.autograd.grad(y, pt)
        assert abs(y.item() - out[s]) <= 1e-12 * max(1.0, abs(out[s]))
        assert np.allclose(J[s], g.numpy(), rtol=1e-10, atol=1e-12), name


@pytest.mark.parametrize("name", ["stride_bias", "activations_mean", "wide_kernel_odd_stride"])
def test_stack_jacobian_vs_finite_differences(name):
    n, layers = STACKS[name]
    desc = orc.ff_stack(n, layers)
    p = params_for(desc, 3)
    sigma = np.random.default_rng(5).choice(np.array([-1, 1], dtype=np.int8), size=(2, n))
    J, _ = orc.ff_jacobian(desc, p, sigma)
    h = 1e-6
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        fd = (orc.ff_forward(desc, p + e, sigma) - orc.ff_forward(desc, p - e, sigma)) / (2 * h)
        assert np.allclose(J[:, k], fd, rtol=1e-5, atol=1e-7), (name, k)


def test_c4_stack_equals_c4_restatement():
    n, layers = STACKS["c4"]
    stack = orc.ff_stack(n, layers)
    c4 = orc.FFDesc(n, 3, 3, 6, 0, None)
    p = params_for(stack, 9)
    sigma = np.random.default_rng(1).choice(np.array([-1, 1], dtype=np.int8), size=(5, n))
    J1, o1 = orc.ff_jacobian(stack, p, sigma)
    J2, o2 = orc.ff_jacobian(c4, p, sigma)
    assert np.allclose(o1, o2, rtol=1e-14, atol=0) and np.allclose(J1, J2, rtol=1e-13, atol=1e-15)


def test_sigmoid_value_is_the_reference_closed_form():
    n = 6
    layers = [(FC, 6, 1, 0, 0, 1, 0), (ACT, orc.A_SIGMOID, 0, 0, 0, 0, 0)]
    desc = orc.ff_stack(n, layers)
    p = params_for(desc, 2)
    sigma = np.random.default_rng(4).choice(np.array([-1, 1], dtype=np.int8), size=(4, n))
    z = sigma.astype(np.float64) @ p[:6] + p[6]
    assert np.allclose(orc.ff_forward(desc, p, sigma), np.exp(z) + 1.0, rtol=1e-14)


def test_stack_shape_errors():
    with pytest.raises(ValueError):
        orc.ff_stack(10, [(FC, 9, 1, 0, 0, 0, 0)])          # inputDim != width
    with pytest.raises(ValueError):
        orc.ff_stack(10, [(FC, 10, 2, 0, 0, 0, 0)])         # two outputs
    with pytest.raises(ValueError):
        orc.ff_stack(10, [(C1, 3, 1, 3, 1, 0, 0), (MEAN, 3, 1, 0, 0, 0, 0)])  # 10 % 3
