"""Pins of the FeedForward oracle (or_ff_*): central finite differences in FP64
(eps = 1e-4, the C4 configuration; Tests/TestFeedForward.cpp:42-79 and
Tests/LayerHelper.hpp:11-60 use the same technique) and torch.autograd on an
independent float64 restatement of the same network."""
import numpy as np
import pytest

C4 = dict(n=40, channels=8, kernel=5, hidden=64)


def _desc(orc, n, channels, kernel, hidden):
    return orc.FFDesc(n, channels, kernel, hidden)


@pytest.mark.parametrize("shape", [C4, dict(n=12, channels=3, kernel=3, hidden=7),
                                   dict(n=9, channels=2, kernel=7, hidden=5)])
def test_ff_jacobian_finite_differences(orc, shape):
    d = _desc(orc, **shape)
    npar = orc.ff_npar(d)
    assert npar == shape["kernel"] * shape["channels"] + \
        shape["channels"] * shape["n"] * shape["hidden"]
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, npar)
    sig = rng.choice([-1, 1], (3, shape["n"])).astype(np.int8)
    J, out = orc.ff_jacobian(d, p, sig)
    np.testing.assert_allclose(out, orc.ff_forward(d, p, sig), rtol=0, atol=0)
    eps = 1e-4
    cols = np.concatenate([np.arange(min(npar, 60)), rng.choice(npar, 60, replace=False)])
    for q in cols:
        pp, pm = p.copy(), p.copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (orc.ff_forward(d, pp, sig) - orc.ff_forward(d, pm, sig)) / (2 * eps)
        np.testing.assert_allclose(J[:, q], fd, rtol=1e-6, atol=1e-9)


def test_ff_matches_torch_autograd(orc):
    torch = pytest.importorskip("torch")
    d = _desc(orc, **C4)
    n, C, K, H = C4["n"], C4["channels"], C4["kernel"], C4["hidden"]
    rng = np.random.default_rng(2)
    p = rng.normal(0, 0.2, orc.ff_npar(d))
    sig = rng.choice([-1, 1], (4, n)).astype(np.int8)
    J, out = orc.ff_jacobian(d, p, sig)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    ker = pt[:K * C].reshape(C, K)              # ker[oc, ki] = p[ki + oc*K]
    W = pt[K * C:].reshape(H, C * n)            # W[o, i] = p[K*C + i + o*C*n]
    x = torch.tensor(sig, dtype=torch.float64)
    idx = torch.tensor([[(r + ki - K // 2) % n for ki in range(K)] for r in range(n)])
    z = torch.einsum("ck,brk->bcr", ker, x[:, idx]).reshape(len(sig), C * n)
    a1 = torch.log(torch.cosh(z))
    a2 = torch.tanh(a1 @ W.T)
    f = a2.mean(dim=1)
    np.testing.assert_allclose(f.detach().numpy(), out, rtol=1e-13, atol=1e-15)
    for s in range(len(sig)):
        g, = torch.autograd.grad(f[s], pt, retain_graph=True)
        np.testing.assert_allclose(g.numpy(), J[s], rtol=1e-10, atol=1e-14)
