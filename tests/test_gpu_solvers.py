"""Solver options on the GPU: Jacobi-preconditioned CG and warm starts solve
the same system as the Cholesky path (the oracle's or_llt_solve) to the
requested tolerance, in fewer iterations; the FP64 factored products on
MFMA f64 agree with the dense operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _batch(ctx, orc, n, m, ns, seed=9, scale=0.1):
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(seed)
    rbm = orc.RBM.random(n, m, scale, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    th = orc.theta(rbm, sig)
    el = orc.eloc(orc.ham(0, n), rbm, sig, th)
    O, om, F, S = orc.sr_build(rbm, sig, th, el)
    return rbm, sig, O, om, F, S


def _np_pcg(A, b, pre, tol, x0=None):
    """Textbook (P)CG in numpy on an explicit matrix; returns x, iterations."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A @ x
    z = r * pre if pre is not None else r.copy()
    p, rho, bn = z.copy(), r @ z, b @ b
    for it in range(1, 100000):
        Ap = A @ p
        al = rho / (p @ Ap)
        x += al * p
        r -= al * Ap
        if r @ r < tol * tol * bn:
            return x, it
        z = r * pre if pre is not None else r
        rn = r @ z
        p = z + rn / rho * p
        rho = rn
    return x, it


@pytest.mark.parametrize("prec", [0, 1])
def test_jacobi_pcg_matches_cholesky(ctx, orc, prec):
    """Jacobi PCG solves the same system as Cholesky.  diag(S) = Var(O_k) is
    pinned by the iteration count of a numpy PCG with the oracle S's diagonal
    (a wrong diagonal changes the Krylov sequence)."""
    n, m, ns = 24, 48, 700
    rbm, sig, O, om, F, S = _batch(ctx, orc, n, m, ns)
    ctx.samples_set(sig, precision=prec)
    ctx.sr_build(2, precision=prec)
    lam = 1e-3
    tol = 1e-10 if prec == 0 else 1e-8
    ctx.sr_options(precond=0)
    x0, res0, it0 = ctx.sr_solve(lam, tol=tol, max_iter=20000)
    ctx.sr_options(precond=1)
    x1, res1, it1 = ctx.sr_solve(lam, tol=tol, max_iter=20000)
    ctx.sr_options(precond=0)
    assert res0 < tol and res1 < tol
    if prec == 0:
        xr = orc.llt_solve(S, lam, F)
        assert _rel(x1, xr) < 1e-7 and _rel(x0, xr) < 1e-7
        A = S + lam * np.eye(len(F))
        _, itp = _np_pcg(A, F, 1.0 / np.diag(A), tol)
        _, itc = _np_pcg(A, F, None, tol)
        assert abs(it1 - itp) <= 0.05 * itp, (it1, itp)
        assert abs(it0 - itc) <= 0.05 * itc, (it0, itc)
    else:
        # the mixed batch rounds theta to float: compare the two CG answers,
        # each within 1e-8 relative residual (so apart by up to cond(A) 1e-8)
        assert _rel(x1, x0) < 1e-5


def test_warm_start_reduces_iterations(ctx, orc):
    n, m, ns = 24, 48, 700
    rbm, sig, O, om, F, S = _batch(ctx, orc, n, m, ns)
    ctx.samples_set(sig)
    ctx.sr_build(2)
    lam = 1e-3
    F2 = F * (1 + 1e-3 * np.random.default_rng(1).normal(size=F.shape))
    ctx.sr_options(precond=0, warm_start=0)
    _, res_c, it_cold = ctx.sr_solve(lam, tol=1e-10, max_iter=20000, rhs=F2)
    ctx.sr_options(precond=0, warm_start=1)
    ctx.sr_solve(lam, tol=1e-10, max_iter=20000)           # leaves x(F) behind
    # a nearby right-hand side: start from the previous solution
    x2, res2, it2 = ctx.sr_solve(lam, tol=1e-10, max_iter=20000, rhs=F2)
    ctx.sr_options(0, 0)
    assert res_c < 1e-10 and res2 < 1e-10 and it2 < it_cold, (it2, it_cold)
    assert _rel(x2, orc.llt_solve(S, lam, F2)) < 1e-7


def test_factored_fp64_apply_large(ctx, orc):
    """MFMA f64 factored products at the C3 machine shape vs a dense numpy O."""
    n, m, ns = 100, 400, 1000
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(10)
    rbm = orc.RBM.random(n, m, 0.05, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(2)
    T = np.tanh(orc.theta_replay(rbm, sig))
    Ow = (sig[:, :, None] * T[:, None, :]).reshape(ns, -1)
    O = np.concatenate([Ow, sig.astype(float), T], axis=1)
    np.testing.assert_allclose(ctx.sr_get(1), O.mean(axis=0), rtol=0, atol=1e-13)
    Oc = O - O.mean(axis=0)
    v = rng.normal(size=O.shape[1])
    assert _rel(ctx.sr_apply(1e-3, v), Oc.T @ (Oc @ v) / ns + 1e-3 * v) < 1e-12


def test_sr_step_options(ctx):
    """C3-shaped epoch bodies with the mixed solver, Jacobi PCG and warm start."""
    ctx.rbm_create(100, 400)
    ctx.init_random(7, 0.01)
    ctx.ham_set(0, 100, J=-1.0, h=-1.0)
    ctx.sampler_config(0, 2048, precision=1, seed=7)
    its = []
    for i in range(4):
        o = ctx.sr_step(n_sweeps=1, n_discard=20 if i == 0 else 0, solver=2, precision=1,
                        lam=1e-3, cg_tol=1e-8, lr=0.01, precond=1, warm_start=1)
        assert o["cg_err"] < 1e-8 and np.isfinite(o["energy"])
        its.append(o["cg_iters"])
    assert min(its) > 0
