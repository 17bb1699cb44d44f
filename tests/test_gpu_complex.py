"""Complex RBM (RBM<std::complex<double>>) on the GPU against the complex
oracle (oracle/oracle_complex.cpp).

Sampler: spins, theta, counters and accept counts bit-exact against the
replay (the decisions use the replayed Re logcosh).  E_loc: within 1e-13
of the kernel replay (rp_c_eloc; the phase uses the platform tanh / atan2,
so no bit-exact claim) and 1e-10 of the libm restatement in FP64; in float
within 1e-6 of the replay and, like the replay, within 1e-5 norm-wise of the
restatement.
O_c, omean, F, S and the solutions: within 1e-10 relative.  The SR loop
reaches the exact ground-state energy of a small XXZ chain."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _setup(ctx, orc, n, m, seed=21, scale=0.05):
    ctx.rbm_create(n, m, dtype=np.complex128)
    rng = np.random.default_rng(seed)
    rbm = orc.CRBM.random(n, m, scale, rng)
    ctx.set_params(rbm.params())
    np.testing.assert_array_equal(ctx.get_params(), rbm.params())
    return rbm


@pytest.mark.parametrize("n,m,chains,kind,prec,sweeps,discard", [
    (20, 40, 16, 0, 0, 4, 10),      # local moves, fp64
    (20, 40, 16, 0, 1, 4, 10),      # local moves, mixed
    (16, 32, 40, 1, 0, 4, 5),       # exchange moves, Sz = 0
    (16, 32, 40, 1, 1, 4, 5),
    (64, 128, 64, 1, 0, 2, 1),      # C2 shape (N=64, alpha=2), exchange, fp64
    (64, 128, 64, 1, 1, 2, 1),      # C2 shape, mixed
    (37, 70, 9, 0, 0, 3, 2),        # ragged sizes
])
def test_complex_sampler_bit_exact(ctx, orc, n, m, chains, kind, prec, sweeps, discard):
    rbm = _setup(ctx, orc, n, m, scale=0.3)
    seed = 0xC2 + n
    ctx.sampler_config(kind, chains, precision=prec, seed=seed)
    ctx.randomize(n // 2 if kind else -1)
    s0, _, _ = ctx.sampler_state()
    ctx.sample(sweeps, discard)
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.c_sample_replay(rbm, s0, seed=seed, n_sweeps=sweeps,
                                                   n_discard=discard, kind=kind, precision=prec)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    assert (acc > 0).all()
    if kind == 1:
        assert (gs.astype(int).sum(axis=1) == 0).all()
    # theta cache invariant against the restatement
    np.testing.assert_allclose(gt, orc.c_theta(rbm, gs), atol=2e-6 if prec else 1e-12)


def test_complex_init_random(ctx):
    ctx.rbm_create(10, 20, dtype=np.complex128)
    ctx.init_random(5, 0.1)
    p = ctx.get_params()
    assert p.dtype == np.complex128 and not ctx.has_nan()
    assert 0.07 < p.real.std() < 0.13 and 0.07 < p.imag.std() < 0.13
    assert abs(np.corrcoef(p.real, p.imag)[0, 1]) < 0.2


HAMS = [
    (0, dict(J=-1.0, h=-0.8)),
    (1, dict(J=1.0, delta=1.0)),
    (1, dict(J=1.0, delta=0.5, sign_rule=1)),
    (2, dict(J=1.0, J2=0.5, sign_rule=1)),
]


@pytest.mark.parametrize("kind,kw", HAMS)
@pytest.mark.parametrize("prec", [0, 1])
def test_complex_eloc(ctx, orc, kind, kw, prec):
    n, m, ns = 24, 48, 300
    rbm = _setup(ctx, orc, n, m, scale=0.2)
    ctx.ham_set(kind, n, **kw)
    rng = np.random.default_rng(kind + 10 * prec)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=prec)
    e, mean, var = ctx.eloc()
    H = orc.ham(kind, n, **kw)
    th = orc.c_theta(rbm, sig)
    want = orc.c_eloc(H, rbm, sig, th)
    # the kernels' arithmetic on the CPU (same theta rounding); the device
    # differs only through the platform tanh / atan2 of the phase
    th_k = th.astype(np.complex64).astype(np.complex128) if prec else th
    rep = orc.c_eloc_replay(H, rbm, sig, th_k, precision=prec)
    if prec == 0:
        assert _rel(e, want) < 1e-10
        assert _rel(e, rep) < 1e-13
    else:
        assert _rel(e, rep) < 1e-6
        assert np.max(np.abs(e - want) / np.maximum(np.abs(want), 1.0)) < 1e-4
        assert _rel(e, want) < 1e-5
    assert abs(mean - e.real.mean()) < 1e-12 * max(1.0, abs(mean))
    assert abs(var - ((e.real ** 2).mean() - e.real.mean() ** 2)) < 1e-9 * max(1.0, var)


@pytest.mark.parametrize("solver", [0, 1])
def test_complex_sr_build_and_solve(ctx, orc, solver):
    n, m, ns = 16, 24, 500
    rbm = _setup(ctx, orc, n, m, scale=0.1)
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    rng = np.random.default_rng(31)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(solver)
    th = orc.c_theta(rbm, sig)
    el = orc.c_eloc(orc.ham(1, n, J=1.0, delta=1.0), rbm, sig, th)
    O, om, F, S = orc.c_sr_build(rbm, sig, th, el)
    assert _rel(ctx.sr_get(1), om) < 1e-12
    assert _rel(ctx.sr_get(0), F) < 1e-10
    assert _rel(ctx.sr_get(2), O) < 1e-12
    if solver == 0:
        assert _rel(ctx.sr_get(3), S) < 1e-12
    lam = 1e-3
    A = S + lam * np.eye(len(F))
    x_ref = np.linalg.solve(A, F)
    v = rng.normal(size=len(F)) + 1j * rng.normal(size=len(F))
    assert _rel(ctx.sr_apply(lam, v), A @ v) < 1e-12
    x, res, it = ctx.sr_solve(lam, tol=1e-11, max_iter=20000)
    assert res < 1e-10
    assert _rel(x, x_ref) < 1e-8
    if solver == 1:
        assert it > 0
        # warm start and Jacobi options keep the answer
        ctx.sr_options(precond=1, warm_start=1)
        x2, res2, _ = ctx.sr_solve(lam, tol=1e-11, max_iter=20000)
        ctx.sr_options(0, 0)
        assert res2 < 1e-10 and _rel(x2, x_ref) < 1e-8


@pytest.mark.parametrize("n,m,ns", [(16, 24, 500), (64, 128, 2048)])
def test_complex_factored_build_and_solve(ctx, orc, n, m, ns):
    """The factored build of a complex machine (O never stored; every product
    four real FP64 MFMA factored products on the planes of T): omean, F and the
    operator against the oracle, and the CG solution against a direct solve
    (small shape) or the dense CG (C2 shape)."""
    rbm = _setup(ctx, orc, n, m, scale=0.1 if n < 64 else 0.02)
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    rng = np.random.default_rng(n + 1)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(2)
    lam = 1e-3
    v = rng.normal(size=ctx.sr_dim) + 1j * rng.normal(size=ctx.sr_dim)
    with pytest.raises(Exception):
        ctx.sr_get(2)  # O is not materialised
    if n < 64:
        th = orc.c_theta(rbm, sig)
        el = orc.c_eloc(orc.ham(1, n, J=1.0, delta=1.0), rbm, sig, th)
        O, om, F, S = orc.c_sr_build(rbm, sig, th, el)
        assert _rel(ctx.sr_get(1), om) < 1e-12
        assert _rel(ctx.sr_get(0), F) < 1e-10
        A = S + lam * np.eye(len(F))
        assert _rel(ctx.sr_apply(lam, v), A @ v) < 1e-12
        x_ref = np.linalg.solve(A, F)
    else:
        F = ctx.sr_get(0)
        Av = ctx.sr_apply(lam, v)
        ctx.sr_build(1)   # the dense build of the same batch
        assert _rel(ctx.sr_get(0), F) < 1e-12
        assert _rel(ctx.sr_apply(lam, v), Av) < 1e-12
        x_ref, r_ref, _ = ctx.sr_solve(lam, tol=1e-11, max_iter=20000)
        ctx.sr_build(2)
    x, res, it = ctx.sr_solve(lam, tol=1e-11, max_iter=20000)
    assert res < 1e-10 and it > 0
    assert _rel(x, x_ref) < 1e-8
    ctx.sr_options(precond=1)
    with pytest.raises(Exception):
        ctx.sr_solve(lam, tol=1e-11, max_iter=20000)
    ctx.sr_options(0, 0)


@pytest.mark.parametrize("n,m,ns", [(12, 20, 700), (32, 64, 2048)])
def test_complex_explicit_fp64_and_mixed_s(ctx, orc, n, m, ns):
    """The explicit S of a complex machine comes from the hand-written SYRK of
    the centred conj(O) rows read as 2 dim reals (FP64 MFMA): equal to the
    oracle's S to 1e-12.  With YQ_MIXED the same product runs in split-bf16
    (six hi/mid/lo products, fp32 accumulation): S to float accuracy, and the
    FP64-refined Cholesky solve reaches the FP64 solution."""
    rbm = _setup(ctx, orc, n, m, scale=0.1)
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    rng = np.random.default_rng(n)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(0)
    S64 = ctx.sr_get(3)
    th = orc.c_theta(rbm, sig)
    el = orc.c_eloc(orc.ham(1, n, J=1.0, delta=1.0), rbm, sig, th)
    _, _, F, S = orc.c_sr_build(rbm, sig, th, el)
    assert _rel(S64, S) < 1e-12
    lam = 1e-3
    x_ref = np.linalg.solve(S + lam * np.eye(len(F)), F)
    ctx.sr_build(0, precision=1)
    Sm = ctx.sr_get(3)
    assert np.abs(Sm - S).max() <= 3e-6 * np.abs(S).max()
    assert _rel(ctx.sr_get(0), F) < 1e-10
    x, res, rounds = ctx.sr_solve(lam, tol=1e-12)
    assert res < 1e-12 and 1 <= rounds <= 30
    assert _rel(x, x_ref) < 1e-9


def test_complex_mixed_batch_matches_fp64(ctx, orc):
    """C2 shape (N=64, alpha=2, complex): the same configurations evaluated
    with float E_loc and with FP64 give the same energy and force to float
    accuracy (SR itself runs in FP64 for complex machines)."""
    n, m, ns = 64, 128, 2048
    _setup(ctx, orc, n, m, scale=0.02)
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    sig = orc.randomize(3, 0, np.arange(ns), n, nup=n // 2)
    out = {}
    for prec in (0, 1):
        ctx.samples_set(sig, precision=prec)
        e, mean, var = ctx.eloc()
        ctx.sr_build(1, precision=prec)
        out[prec] = (e, mean, ctx.sr_get(0))
    assert abs(out[1][1] - out[0][1]) < 1e-5 * abs(out[0][1])
    assert _rel(out[1][0], out[0][0]) < 1e-5
    assert _rel(out[1][2], out[0][2]) < 1e-4


def test_complex_factored_float_batch_matches_dense(ctx, orc):
    """A batch sampled in float (theta stored as float): the factored build
    reads tanh of the widened theta and agrees with the dense build of the
    same batch."""
    n, m, ns = 16, 24, 500
    _setup(ctx, orc, n, m, scale=0.1)
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    rng = np.random.default_rng(5)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=1)
    ctx.eloc()
    v = rng.normal(size=ctx.sr_dim) + 1j * rng.normal(size=ctx.sr_dim)
    out = {}
    for solver in (1, 2):
        ctx.sr_build(solver)
        out[solver] = (ctx.sr_get(0), ctx.sr_get(1), ctx.sr_apply(1e-3, v))
    for a, b in zip(out[2], out[1]):
        assert _rel(a, b) < 1e-12


@pytest.mark.parametrize("solver,prec", [(0, 0), (2, 1)])
def test_complex_xxz_reaches_ground_state(ctx, orc, solver, prec):
    """SR with exchange moves on the Sz = 0 sector of the N=10 Heisenberg
    chain (Marshall sign rule) reaches the exact ground-state energy: the
    explicit solver on FP64 chains, and the factored CG on float chains."""
    n = 10
    H = orc.ham(1, n, J=1.0, delta=1.0, sign_rule=1)
    e0 = np.linalg.eigvalsh(orc.ham_matrix(H).toarray())[0]
    ctx.rbm_create(n, 2 * n, dtype=np.complex128)
    ctx.init_random(17, 0.01)
    ctx.ham_set(1, n, J=1.0, delta=1.0, sign_rule=1)
    ctx.sampler_config(1, 256, precision=prec, seed=17)
    ctx.randomize(n // 2)
    es = []
    for i in range(150):
        o = ctx.sr_step(n_sweeps=4, n_discard=20 if i == 0 else 2, solver=solver,
                        precision=prec, lam=1e-3, lr=0.02, cg_tol=1e-6)
        assert o["cg_err"] < 1e-6 or solver == 0
        es.append(o["energy"])
    e = np.mean(es[-20:])
    assert abs(e - e0) < 0.01 * abs(e0), (e, e0)


def test_complex_log_ratio_vector(ctx, orc):
    """Complex logRatio(vector) on the GPU against the oracle's restatement
    (RBMState.hpp:108-126), modulo 2 pi i, C2 shape."""
    n, m, ns, k = 64, 128, 200, 9
    rbm = _setup(ctx, orc, n, m, scale=0.2)
    rng = np.random.default_rng(42)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    sites = np.full((ns, k), -1, np.int32)
    for s in range(ns):
        cnt = 1 + s % k
        sites[s, :cnt] = rng.choice(n, cnt, replace=False)
    got = ctx.log_ratio(sig, sites)
    th = orc.c_theta(rbm, sig)
    for s in range(ns):
        want = orc.c_log_ratio_vec(rbm, sig[s], th[s], sites[s][sites[s] >= 0])
        d = got[s] - want
        d = complex(d.real, (d.imag + np.pi) % (2 * np.pi) - np.pi)
        assert abs(d) <= 1e-11 * max(1.0, abs(want)), (s, got[s], want)
