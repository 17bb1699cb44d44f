"""Solver options on the GPU: Jacobi-preconditioned CG and warm starts solve
the same system as the Cholesky path (the oracle's or_llt_solve) to the
requested tolerance (the Jacobi iteration count is pinned against a numpy
PCG; on these SR matrices it is higher than plain CG's); the FP64 factored
products on MFMA f64 agree with the dense operator; the graph-captured mixed
CG matches the eagerly launched one bit for bit."""
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


def test_graph_chunks_match_eager(ctx, orc):
    """The mixed CG replayed as captured 8-iteration graphs (stopping on the
    device) and the same kernels issued one by one give bit-identical
    solutions and iteration counts, over repeated solves with the FP64
    refinement running between the graph replays, and honour an iteration
    budget that is not a multiple of the chunk."""
    n, m, ns = 40, 80, 3000
    _batch(ctx, orc, n, m, ns, scale=0.05)
    rng = np.random.default_rng(12)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=1)
    ctx.sr_build(2, precision=1)
    out = {False: [], True: []}
    for mode in (False, True, False, True):
        ctx.graph_mode(mode)
        x, res, it = ctx.sr_solve(1e-3, tol=1e-8, max_iter=5000)
        assert res < 1e-8
        out[mode].append((x, it))
    for mode in (False, True):
        ctx.graph_mode(mode)
        x, res, it = ctx.sr_solve(1e-3, tol=1e-8, max_iter=13, allow_unconverged=True)
        assert it == 13 and res > 1e-8
        out[mode].append((x, it))
    ctx.graph_mode(True)
    for (xe, ie), (xg, ig) in zip(out[False], out[True]):
        assert ie == ig
        np.testing.assert_array_equal(xe, xg)
    # repeated solves of the same system are identical too
    np.testing.assert_array_equal(out[True][0][0], out[True][1][0])


def test_mixed_solve_budgets_chunks_and_zero_rhs(ctx, orc):
    """The host's queueing of captured chunks (one ahead, or none when the
    published <r, r> forecasts convergence inside the queued chunk) never
    changes what the solve computes: chunks of 8 and 16 iterations and eager
    launches give bit-identical solutions and iteration counts, converged and
    for budgets of 1 (no chunk at all), 2, 9 and 17 iterations; a zero
    right-hand side returns x = 0 after no iteration with a zero residual."""
    n, m, ns = 40, 80, 3000
    _batch(ctx, orc, n, m, ns, scale=0.05)
    rng = np.random.default_rng(21)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=1)
    ctx.sr_build(2, precision=1)
    dim = n * m + n + m
    runs = {}
    for mode in (0, 8, 16):
        ctx.graph_mode(mode if mode != 8 else True)
        x, res, it = ctx.sr_solve(1e-3, tol=1e-8, max_iter=5000)
        assert res < 1e-8 and it > 17
        runs[(mode, None)] = (x, it)
        for budget in (1, 2, 9, 17):
            x, res, it = ctx.sr_solve(1e-3, tol=1e-8, max_iter=budget, allow_unconverged=True)
            assert it == budget and res > 1e-8
            runs[(mode, budget)] = (x, it)
        x, res, it = ctx.sr_solve(1e-3, tol=1e-8, max_iter=5000, rhs=np.zeros(dim))
        assert it == 0 and res == 0.0
        np.testing.assert_array_equal(x, np.zeros(dim))
    ctx.graph_mode(True)
    for budget in (None, 1, 2, 9, 17):
        xe, ie = runs[(0, budget)]
        for mode in (8, 16):
            xg, ig = runs[(mode, budget)]
            assert ig == ie, (mode, budget)
            np.testing.assert_array_equal(xg, xe)


@pytest.mark.parametrize("n,m,ns", [(100, 400, 2048), (40, 80, 3000)])
def test_tile_shapes_and_autotune(ctx, orc, n, m, ns):
    """Every tile shape of the mixed products (forward column group, backward
    chunk) computes the same operator to float accuracy, and the autotuned
    solve converges to the same answer."""
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(14)
    rbm = orc.RBM.random(n, m, 0.05, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=1)
    th = orc.theta_replay(rbm, sig).astype(np.float32).astype(np.float64)
    T = np.tanh(th)
    Ow = (sig[:, :, None] * T[:, None, :]).reshape(ns, -1)
    Od = np.concatenate([Ow, sig.astype(float), T], axis=1)
    Oc = Od - Od.mean(axis=0)
    v = np.random.default_rng(4).normal(size=Od.shape[1])
    ref = Oc.T @ (Oc @ v) / ns + 1e-3 * v
    kp = (n + 1 + 31) // 32 * 32
    for ct in (2, 4, 8):
        if 2 * 3 * ct * 16 * (kp + 8) > 64 * 1024:
            with pytest.raises(Exception):
                ctx.tiles(ct, 256)
                ctx.sr_build(2, precision=1)
            continue
        for chunk in (128, 256, 512):
            ctx.tiles(ct, chunk)
            ctx.sr_build(2, precision=1)
            assert _rel(ctx.sr_apply(1e-3, v), ref) < 2e-6, (ct, chunk)
    ctx.tiles(0, 0)
    ctx.sr_build(2, precision=1)
    x0, r0, _ = ctx.sr_solve(1e-3, tol=1e-8, max_iter=20000)
    ctx.autotune(True)
    ctx.sr_build(2, precision=1)
    x1, r1, _ = ctx.sr_solve(1e-3, tol=1e-8, max_iter=20000)
    ctx.autotune(False)
    assert r0 < 1e-8 and r1 < 1e-8 and _rel(x1, x0) < 1e-5


def test_mixed_solve_budget_equal_to_converged_count_and_nan_rhs(ctx, orc):
    """A budget equal to the iteration count of an unbounded solve converges
    to the same solution bit for bit (the last inner run ends its budget on
    the iteration that reaches the threshold, which is then tested on the
    device), for eager launches and chunks of 8 and 16; a right-hand side with
    a NaN or an infinity is reported as YQ_ERR_NAN by the mixed and the FP64
    CG instead of a silent zero solution."""
    import yannq_amd as yq
    n, m, ns = 40, 80, 3000
    _batch(ctx, orc, n, m, ns, scale=0.05)
    rng = np.random.default_rng(33)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    dim = n * m + n + m
    ctx.samples_set(sig, precision=1)
    ctx.sr_build(2, precision=1)
    for mode in (0, 8, 16):
        ctx.graph_mode(mode if mode != 8 else True)
        x, res, it = ctx.sr_solve(1e-3, tol=1e-8, max_iter=5000)
        assert res < 1e-8
        xb, resb, itb = ctx.sr_solve(1e-3, tol=1e-8, max_iter=it)
        assert itb == it and resb == res
        np.testing.assert_array_equal(xb, x)
        for bad in (np.nan, np.inf):
            b = rng.normal(size=dim)
            b[dim // 3] = bad
            with pytest.raises(yq.YqError) as ei:
                ctx.sr_solve(1e-3, tol=1e-8, max_iter=5000, rhs=b)
            assert ei.value.code == yq.ERR_NAN, str(ei.value)
    ctx.graph_mode(True)
    ctx.samples_set(sig, precision=0)
    ctx.sr_build(2, precision=0)
    b = rng.normal(size=dim)
    b[7] = np.nan
    with pytest.raises(yq.YqError) as ei:
        ctx.sr_solve(1e-3, tol=1e-8, max_iter=5000, rhs=b)
    assert ei.value.code == yq.ERR_NAN, str(ei.value)


def test_fisher_average(ctx, orc):
    """yq_sr_fisher_average is ObsAvg<Matrix> of RunRBM.hpp:13-36 on the
    explicit S: after builds on two batches with beta2 = 0.7 the matrix that
    S reads and the Cholesky solves is (0.7 (0.3 S1) + 0.3 S2) / (1 - 0.7^2);
    a reset starts over."""
    n, m, ns = 10, 20, 800
    rbm = orc.RBM.random(n, m, 0.1, np.random.default_rng(3))
    ctx.rbm_create(n, m)
    ctx.set_params(rbm.params())
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    rng = np.random.default_rng(5)
    Ss, Fs = [], []
    b = 0.7
    for q in range(2):
        sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
        th = orc.theta(rbm, sig)
        el = orc.eloc(orc.ham(0, n), rbm, sig, th)
        _, _, F, S = orc.sr_build(rbm, sig, th, el)
        Ss.append(S)
        Fs.append(F)
        ctx.samples_set(sig)
        ctx.sr_build(0)
        ctx.sr_fisher_average(b, reset=(q == 0))
    want = (b * ((1 - b) * Ss[0]) + (1 - b) * Ss[1]) / (1 - b ** 2)
    assert _rel(ctx.sr_get(3), want) < 1e-12
    x, res, _ = ctx.sr_solve(1e-3, rhs=Fs[1])
    assert _rel(x, orc.llt_solve(want, 1e-3, Fs[1])) < 1e-9 and res < 1e-12
    ctx.sr_fisher_average(b, reset=True)   # a fresh average of the current S
    assert _rel(ctx.sr_get(3), want) < 1e-12
    with pytest.raises(Exception):
        ctx.sr_fisher_average(1.0)


@pytest.mark.parametrize("n,m,ns", [(20, 40, 2000), (7, 13, 37), (12, 30, 5001)])
def test_mixed_explicit_s_and_refined_solve(ctx, orc, n, m, ns):
    """The split-bf16 S of the mixed explicit build (six hi/mid/lo products on
    bf16 MFMA, fp32 accumulation) agrees with the oracle's FP64 S to float
    accuracy, and the mixed explicit solve (FP32 Cholesky inside FP64
    refinement with the FP64 operator) reaches the FP64 Cholesky solution."""
    rbm, sig, O, om, F, S = _batch(ctx, orc, n, m, ns, scale=0.1)
    ctx.samples_set(sig, precision=0)   # FP64 samples and E_loc, mixed S
    ctx.sr_build(0, precision=1)
    Sm = ctx.sr_get(3)
    assert np.abs(Sm - S).max() <= 3e-6 * np.abs(S).max()
    assert _rel(ctx.sr_get(0), F) < 1e-10
    lam = 1e-3
    x, res, rounds = ctx.sr_solve(lam, tol=1e-12)
    assert res < 1e-12 and 1 <= rounds <= 30
    assert _rel(x, orc.llt_solve(S, lam, F)) < 1e-9
