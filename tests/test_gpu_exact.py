"""Exact sums on the GPU (SRMatExact / ExactSampler, GroundState/SRMatExact.hpp:
97-149, Samplers/ExactSampler.hpp:49-77): the weighted basis batch against the
oracle's weighted SR restatement (or_sr_build / or_c_sr_build with |psi|^2
weights over the same basis), the exact energy against <psi|H|psi> from the
ED column functors, F against finite differences of the GPU's own exact
energy, and the exact sampler's histogram against the weights (chi^2)."""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu


def _sector(n, nup):
    """Basis/BasisJz.hpp order: the integers with nup set bits, increasing."""
    v = np.arange(1 << n, dtype=np.uint64)
    pc = np.array([bin(int(x)).count("1") for x in v])
    return v[pc == nup] if nup >= 0 else v


def _spins(n, vals):
    bits = (vals[:, None] >> np.arange(n, dtype=np.uint64)[None, :]) & 1
    return (1 - 2 * bits).astype(np.int8)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("use_bias", [True, False])
def test_real_exact_sums(ctx, orc, use_bias):
    n, m = 10, 20
    rng = np.random.default_rng(21)
    rbm = orc.RBM.random(n, m, 0.3, rng, use_bias=use_bias)
    ctx.rbm_create(n, m, use_bias=use_bias)
    ctx.set_params(rbm.params())
    ctx.ham_set(0, n, J=-1.0, h=-0.7)
    B = ctx.exact_batch(-1)
    assert B == 1 << n
    sig = orc.basis_sigma(n)
    th = orc.theta(rbm, sig)
    lp = np.array([orc.log_coeff(rbm, sig[s], th[s]) for s in range(B)])
    p = np.exp(2 * (lp - lp.max()))
    p /= p.sum()
    np.testing.assert_allclose(ctx.exact_probs(), p, rtol=1e-12, atol=1e-300)
    # energy and variance: <psi|H|psi> with H from the ED column functors
    H = orc.ham_matrix(orc.ham(0, n, J=-1.0, h=-0.7))
    psi = np.sqrt(p) * np.sign(np.exp(lp - lp.max()))
    k = H @ psi
    e, mean, var = ctx.eloc()
    assert abs(mean - psi @ k) < 1e-12 * abs(psi @ k)
    assert abs(var - (k @ k - (psi @ k) ** 2)) < 1e-10 * max(1.0, k @ k)
    # O mean, F, S against the oracle's weighted restatement
    el = orc.eloc(orc.ham(0, n, J=-1.0, h=-0.7), rbm, sig, th)
    _, om, F, S = orc.sr_build(rbm, sig, th, el, weights=p)
    ctx.sr_build(0)
    assert _rel(ctx.sr_get(1), om) < 1e-12
    assert _rel(ctx.sr_get(0), F) < 1e-11
    assert _rel(ctx.sr_get(3), S) < 1e-11
    x, res, _ = ctx.sr_solve(1e-3, tol=1e-12, max_iter=1)
    assert _rel(x, orc.llt_solve(S, 1e-3, F)) < 1e-9
    # the dense CG on the weighted rows solves the same system
    ctx.sr_build(1)
    x2, res2, it = ctx.sr_solve(1e-3, tol=1e-11, max_iter=20000)
    assert _rel(x2, x) < 1e-8
    with pytest.raises(Exception):
        ctx.sr_build(2)   # the factored solver has no weights


def test_gradient_is_derivative_of_exact_energy(ctx, orc):
    """dE/dtheta_k = 2 F_k, with E the GPU's own exact energy at theta +- eps e_k."""
    n, m = 8, 8
    rng = np.random.default_rng(4)
    rbm = orc.RBM.random(n, m, 0.4, rng)
    ctx.rbm_create(n, m)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    p0 = rbm.params()
    ctx.set_params(p0)
    ctx.exact_batch(-1)
    ctx.sr_build(0)
    F = ctx.sr_get(0)
    eps = 1e-5
    for k in rng.choice(len(p0), 8, replace=False):
        E = []
        for sgn in (1, -1):
            q = p0.copy()
            q[k] += sgn * eps
            ctx.set_params(q)
            ctx.exact_batch(-1)
            E.append(ctx.eloc()[1])
        d = (E[0] - E[1]) / (2 * eps)
        assert abs(d - 2 * F[k]) < 1e-6 * max(1.0, abs(d)), (k, d, 2 * F[k])


def test_complex_sector_sums(ctx, orc):
    """A complex RBM on the XXZ chain in the Sz = 0 sector (BasisJz(n, n/2))."""
    n, m = 10, 20
    rng = np.random.default_rng(8)
    rbm = orc.CRBM.random(n, m, 0.2, rng)
    ctx.rbm_create(n, m, dtype=np.complex128)
    ctx.set_params(rbm.params())
    ctx.ham_set(1, n, J=1.0, delta=1.0, sign_rule=1)
    B = ctx.exact_batch(n // 2)
    vals = _sector(n, n // 2)
    assert B == len(vals) == 252
    sig = _spins(n, vals)
    th = orc.c_theta(rbm, sig)
    lp = np.array([orc.c_log_coeff(rbm, sig[s], th[s]) for s in range(B)])
    p = np.exp(2 * (lp.real - lp.real.max()))
    p /= p.sum()
    np.testing.assert_allclose(ctx.exact_probs(), p, rtol=1e-11)
    H = orc.ham(1, n, J=1.0, delta=1.0, sign_rule=1)
    el = orc.c_eloc(H, rbm, sig, th)
    e, mean, var = ctx.eloc()
    Er = np.sum(p * el.real)
    assert abs(mean - Er) < 1e-12 * abs(Er)
    assert abs(var - (np.sum(p * np.abs(el) ** 2) - Er ** 2)) < 1e-10
    # and the exact energy equals <psi|H|psi> in the sector
    Hm = orc.ham_matrix(H)[vals][:, vals]
    psi = np.exp(lp - lp.real.max())
    psi /= np.linalg.norm(psi)
    assert abs(mean - np.real(np.vdot(psi, Hm @ psi))) < 1e-12 * abs(mean)
    _, om, F, S = orc.c_sr_build(rbm, sig, th, el, weights=p)
    ctx.sr_build(0)
    assert _rel(ctx.sr_get(1), om) < 1e-12
    assert _rel(ctx.sr_get(0), F) < 1e-11
    assert _rel(ctx.sr_get(3), S) < 1e-11


def test_exact_sampler_distribution(ctx, orc):
    n, m = 6, 6
    rng = np.random.default_rng(30)
    rbm = orc.RBM.random(n, m, 0.6, rng)
    ctx.rbm_create(n, m)
    ctx.set_params(rbm.params())
    ctx.exact_batch(-1)
    p = ctx.exact_probs()
    N = 200000
    idx = ctx.exact_sample(N, seed=123)
    s, _ = ctx.samples()
    np.testing.assert_array_equal(s, orc.basis_sigma(n)[idx])
    cnt = np.bincount(idx, minlength=1 << n)
    chi2 = np.sum((cnt - N * p) ** 2 / (N * p))
    assert stats.chi2.sf(chi2, (1 << n) - 1) > 1e-4, chi2
    # the same seed draws the same configurations; another seed does not
    np.testing.assert_array_equal(ctx.exact_sample(N, seed=123), idx)
    assert not np.array_equal(ctx.exact_sample(N, seed=124), idx)


def test_exact_sr_reaches_ground_state(ctx, orc):
    """RunRBMExact's loop (RunRBMExact.hpp): exact S + lambda, Cholesky, SGD."""
    n = 8
    ctx.rbm_create(n, n)
    ctx.init_random(5, 0.01)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    e0 = -2.0 / np.sin(np.pi / (2 * n))
    for it in range(200):
        ctx.exact_batch(-1)
        _, E, _ = ctx.eloc()
        ctx.sr_build(0)
        x, _, _ = ctx.sr_solve(1e-3, tol=1e-12, max_iter=1)
        ctx.update_params(-0.05 * x)
    assert abs(E - e0) / abs(e0) < 2e-3, (E, e0)


def test_large_basis_split_syrk(ctx, orc):
    """A basis larger than 4 x 16384 configurations forms S as partial SYRKs
    over sample chunks plus a remainder (BasisJz(19, 8): 75582 states)."""
    n, m, nup = 19, 19, 8
    rng = np.random.default_rng(40)
    rbm = orc.RBM.random(n, m, 0.15, rng)
    ctx.rbm_create(n, m)
    ctx.set_params(rbm.params())
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    B = ctx.exact_batch(nup)
    vals = _sector(n, nup)
    assert B == len(vals) == 75582
    sig = _spins(n, vals)
    th = orc.theta(rbm, sig)
    p = ctx.exact_probs()
    _, om, _, S = orc.sr_build(rbm, sig, th, np.zeros(B), weights=p)
    ctx.sr_build(0)
    assert _rel(ctx.sr_get(1), om) < 1e-12
    assert _rel(ctx.sr_get(3), S) < 1e-11
