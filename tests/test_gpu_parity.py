"""GPU parity of the HIP path against the CPU oracle (oracle/).

Sampler: spins, theta, proposal counters and accept counts bit-exact against
the replay of the same Philox streams.  E_loc: bit-exact against the kernel
replay and within 1e-10 relative of the libm restatement.  O, mean, F, S and
the SR solution: within 1e-10 relative (norm-wise) of the restatement.  The
mixed path: its error against FP64 is bounded and stated in each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(ctx, orc, n, m, seed=11, scale=0.01, use_bias=True):
    ctx.rbm_create(n, m, use_bias)
    rng = np.random.default_rng(seed)
    rbm = orc.RBM.random(n, m, scale, rng, use_bias)
    ctx.set_params(rbm.params())
    return rbm


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("n,m,chains,kind,prec,sweeps,discard", [
    (20, 20, 16, 0, 0, 8, 20),      # C1 shape, fp64
    (20, 20, 16, 0, 1, 8, 20),      # C1 shape, mixed
    (16, 32, 40, 1, 0, 4, 5),       # exchange moves, Sz = 0
    (16, 32, 40, 1, 1, 4, 5),
    (100, 400, 64, 0, 0, 2, 1),     # C3 shape (N=100, alpha=4), fp64
    (100, 400, 64, 0, 1, 2, 1),     # C3 shape, mixed
    (64, 128, 32, 1, 0, 2, 1),      # C2 shape (real), exchange
    (37, 70, 9, 0, 0, 3, 2),        # ragged sizes
])
def test_sampler_bit_exact(ctx, orc, n, m, chains, kind, prec, sweeps, discard):
    rbm = _setup(ctx, orc, n, m, scale=0.3)
    seed = 0xC0FFEE + n
    ctx.sampler_config(kind, chains, precision=prec, seed=seed)
    ctx.randomize(n // 2 if kind else -1)
    s0, ctr0, acc0 = ctx.sampler_state()
    assert (ctr0 == 0).all() and (acc0 == 0).all()
    ns = ctx.sample(sweeps, discard)
    assert ns == chains * sweeps
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.sample_replay(rbm, s0, seed=seed, n_sweeps=sweeps,
                                                 n_discard=discard, kind=kind, precision=prec)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    assert (acc > 0).all()
    if kind == 1:
        assert (gs.astype(int).sum(axis=1) == 0).all()
    # a second call continues the same streams
    ctx.sample(1, 0)
    fin2, ctr2, acc2 = ctx.sampler_state()
    rfin2, rctr2, racc2, _, _ = orc.sample_replay(rbm, rfin, seed=seed, n_sweeps=1, n_discard=0,
                                                  kind=kind, precision=prec, ctr=rctr, acc=racc)
    np.testing.assert_array_equal(fin2, rfin2)
    np.testing.assert_array_equal(acc2, racc2)


def test_randomize_matches_oracle(ctx, orc):
    _setup(ctx, orc, 24, 24)
    ctx.sampler_config(0, 33, seed=99)   # config randomizes once (epoch 0)
    s, _, _ = ctx.sampler_state()
    np.testing.assert_array_equal(s, orc.randomize(99, 0, np.arange(33), 24))
    ctx.randomize(12)                    # epoch 1, Sz = 0 sector
    s, _, _ = ctx.sampler_state()
    np.testing.assert_array_equal(s, orc.randomize(99, 1, np.arange(33), 24, nup=12))


HAMS = [
    (0, dict(J=-1.0, h=-1.0)),
    (1, dict(J=1.0, delta=1.0, sign_rule=1)),
    (2, dict(J=1.0, J2=0.5, sign_rule=1)),
]


@pytest.mark.parametrize("kind,kw", HAMS)
@pytest.mark.parametrize("prec", [0, 1])
def test_eloc_parity(ctx, orc, kind, kw, prec):
    n, m, ns = 20, 40, 300
    rbm = _setup(ctx, orc, n, m, scale=0.2)
    ctx.ham_set(kind, n, **kw)
    H = orc.ham(kind, n, **kw)
    rng = np.random.default_rng(5)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=prec)
    e, mean, var = ctx.eloc()
    th = orc.theta_replay(rbm, sig)
    if prec == 1:
        th = th.astype(np.float32).astype(np.float64)
    rep = orc.eloc_replay(H, rbm, sig, th, precision=prec)
    np.testing.assert_array_equal(_bits(e), _bits(rep))
    ref = orc.eloc(H, rbm, sig, orc.theta(rbm, sig))
    tol = 1e-10 if prec == 0 else 2e-5
    assert np.abs(e - ref).max() <= tol * max(1.0, np.abs(ref).max())
    assert abs(mean - ref.mean()) <= tol * abs(ref.mean()) + 1e-12
    assert abs(var - (np.mean(ref ** 2) - ref.mean() ** 2)) <= 1e-6 * max(1, var)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("use_bias", [True, False])
def test_sr_build_explicit_parity(ctx, orc, use_bias):
    n, m, ns = 20, 20, 1008   # C1 sizes (dim = 440)
    rbm = _setup(ctx, orc, n, m, scale=0.1, use_bias=use_bias)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    rng = np.random.default_rng(6)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(0)
    th = orc.theta(rbm, sig)
    el = orc.eloc(orc.ham(0, n), rbm, sig, th)
    O, om, F, S = orc.sr_build(rbm, sig, th, el)
    assert _rel(ctx.sr_get(2), O) < 1e-12
    assert _rel(ctx.sr_get(1), om) < 1e-12
    assert _rel(ctx.sr_get(0), F) < 1e-10
    assert _rel(ctx.sr_get(3), S) < 1e-12
    lam = 1e-3
    x, res, it = ctx.sr_solve(lam)
    xr = orc.llt_solve(S, lam, F)
    assert _rel(x, xr) < 1e-10 and res < 1e-12 and it == 0
    # apply == (S + shift) v
    v = rng.normal(size=rbm.dim)
    assert _rel(ctx.sr_apply(lam, v), S @ v + lam * v) < 1e-12


@pytest.mark.parametrize("n,m,ns,use_bias", [
    (7, 13, 37, True),      # one partial tile, ragged samples
    (9, 30, 517, False),    # three tile rows, no bias
    (12, 24, 4001, True),   # ragged everything
    (8, 8, 20000, True),    # one tile, many samples: split into sample chunks
])
def test_syrk_shapes(ctx, orc, n, m, ns, use_bias):
    """The hand-written FP64 MFMA S = O_c^T O_c / Ns (formed from the spins,
    tanh(theta) and the column mean) against the oracle's S at ragged sizes:
    dims that are not multiples of the 128 tile, sample counts that are not
    multiples of the 16-sample stage, without bias, and a tall batch whose
    samples are split into chunks summed in order."""
    rbm = _setup(ctx, orc, n, m, scale=0.2, use_bias=use_bias)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    rng = np.random.default_rng(n * m + ns)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(0)
    th = orc.theta(rbm, sig)
    el = orc.eloc(orc.ham(0, n), rbm, sig, th)
    _, _, F, S = orc.sr_build(rbm, sig, th, el)
    Sg = ctx.sr_get(3)
    assert _rel(Sg, S) < 1e-12
    np.testing.assert_array_equal(Sg, Sg.T)
    lam = 1e-3
    x, res, _ = ctx.sr_solve(lam)
    assert _rel(x, orc.llt_solve(S, lam, F)) < 1e-9 and res < 1e-12


@pytest.mark.parametrize("solver", [1, 2])
def test_cg_solvers_match_cholesky(ctx, orc, solver):
    n, m, ns = 20, 40, 600
    rbm = _setup(ctx, orc, n, m, scale=0.1)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    rng = np.random.default_rng(7)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    th = orc.theta(rbm, sig)
    el = orc.eloc(orc.ham(0, n), rbm, sig, th)
    O, om, F, S = orc.sr_build(rbm, sig, th, el)
    ctx.sr_build(solver)
    assert _rel(ctx.sr_get(1), om) < 1e-12
    assert _rel(ctx.sr_get(0), F) < 1e-10
    lam = 1e-2
    v = rng.normal(size=rbm.dim)
    assert _rel(ctx.sr_apply(lam, v), S @ v + lam * v) < 1e-12
    x, res, it = ctx.sr_solve(lam, tol=1e-12, max_iter=5000)
    xr = orc.llt_solve(S, lam, F)
    assert res < 1e-11 and it > 0
    assert _rel(x, xr) < 1e-9
    xo, ito, erro, rc = orc.cg_solve(O, lam, F, 1e-12, 5000)
    assert abs(it - ito) <= max(3, 0.05 * ito)


def test_mixed_operator_and_refinement(ctx, orc):
    """Split-bf16 MFMA products vs FP64: operator error and refined solution."""
    n, m, ns = 100, 400, 2048     # C3 machine shape, fewer samples
    rbm = _setup(ctx, orc, n, m, scale=0.01)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    rng = np.random.default_rng(8)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=1)
    ctx.sr_build(2, precision=1)
    F = ctx.sr_get(0)
    # FP64 reference on the same samples (theta rounded to float like the mixed batch)
    ctx2_theta = orc.theta_replay(rbm, sig).astype(np.float32).astype(np.float64)
    T = np.tanh(ctx2_theta)
    lam = 1e-3
    # dense FP64 operator on the factored O
    Ow = (sig[:, :, None] * T[:, None, :]).reshape(ns, -1)
    O = np.concatenate([Ow, sig.astype(float), T], axis=1)
    Oc = O - O.mean(axis=0)
    v = rng.normal(size=O.shape[1])
    ref = Oc.T @ (Oc @ v) / ns + lam * v
    got = ctx.sr_apply(lam, v)
    err = _rel(got, ref)
    assert err < 2e-6, err
    x, res, it = ctx.sr_solve(lam, tol=1e-8, max_iter=20000)
    assert res < 1e-8
    Ax = Oc.T @ (Oc @ x) / ns + lam * x
    r = np.linalg.norm(Ax - F) / np.linalg.norm(F)
    assert r < 2e-8, r


def test_error_paths(ctx, orc, yq):
    n, m = 12, 12
    rbm = _setup(ctx, orc, n, m)
    with pytest.raises(yq.YqError) as e:
        ctx.sr_build(0)
    assert e.value.code == yq.ERR_STATE
    bad = rbm.params()
    bad[3] = np.nan
    with pytest.raises(yq.YqError) as e:
        ctx.set_params(bad)
    assert e.value.code == yq.ERR_NAN
    ctx.ham_set(0, n)
    sig = np.random.default_rng(3).choice([-1, 1], size=(40, n)).astype(np.int8)
    ctx.samples_set(sig)
    ctx.sr_build(0)
    with pytest.raises(yq.YqError) as e:   # ns < dim: S is singular without a shift
        ctx.sr_solve(0.0)
    assert e.value.code == yq.ERR_NOT_SPD
    ctx.sr_build(1)
    with pytest.raises(yq.YqError) as e:
        ctx.sr_solve(1e-3, tol=1e-14, max_iter=2)
    assert e.value.code == yq.ERR_NOT_CONVERGED
    with pytest.raises(yq.YqError) as e:
        ctx.samples_set(np.zeros((2, n), np.int8))
    assert e.value.code == yq.ERR_ARG
    with pytest.raises(yq.YqError) as e:   # the dense (materialised O) CG is FP64 only
        ctx.sr_build(1, precision=1)
    assert e.value.code == yq.ERR_UNSUPPORTED


def test_sr_steps_reach_ground_state(ctx, orc):
    """C1: TFI N=20, alpha=1, 16 chains, SR with lambda=1e-3, lr=0.01: the
    variational energy approaches the exact -2/sin(pi/40)."""
    n = 20
    ctx.rbm_create(n, n)
    ctx.init_random(1234, 0.01)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    ctx.sampler_config(0, 16, seed=4321)
    es = []
    for it in range(300):
        o = ctx.sr_step(n_sweeps=63, n_discard=200 if it == 0 else 5, solver=0, lam=1e-3,
                        lr=0.01)
        es.append(o["energy"])
    e0 = -2 / np.sin(np.pi / 40)
    tail = np.mean(es[-20:])
    assert tail < es[0]
    assert abs(tail - e0) / abs(e0) < 0.005, (tail, e0)


def test_log_ratio_vector(ctx, orc):
    """logRatio(k), (k, l) and (vector) on the GPU (yq_rbm_log_ratio) against
    the oracle's restatements (RBMState.hpp:48-63, 78-93, 108-126), for rows
    of 1 to n listed sites with skipped (-1) entries; C3 shape."""
    n, m, ns, k = 100, 400, 300, 12
    rbm = _setup(ctx, orc, n, m, scale=0.05)
    rng = np.random.default_rng(41)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    sites = np.full((ns, k), -1, np.int32)
    for s in range(ns):
        cnt = 1 + s % k
        sites[s, :cnt] = rng.choice(n, cnt, replace=False)
    got = ctx.log_ratio(sig, sites)
    th = orc.theta(rbm, sig)
    for s in range(ns):
        v = sites[s][sites[s] >= 0]
        want = orc.log_ratio_vec(rbm, sig[s], th[s], v)
        assert abs(got[s] - want) <= 1e-12 * max(1.0, abs(want)), (s, got[s], want)
        if len(v) == 1:
            assert abs(got[s] - orc.log_ratio1(rbm, sig[s], th[s], int(v[0]))) < 1e-12
        if len(v) == 2:
            assert abs(got[s] - orc.log_ratio2(rbm, sig[s], th[s], int(v[0]), int(v[1]))) < 1e-12
    with pytest.raises(Exception):
        ctx.log_ratio(sig[:1], np.array([[n]], np.int32))
