"""The benchmark configurations at their stated sizes on the GPU.

SR at scale (TFI N=100, alpha=4, 40,500 parameters, 16,384 samples):
 * the explicit FP64 path -- S = O_c^T O_c / Ns on the hand-written FP64
   MFMA SYRK (13 GB) and its Cholesky -- against the FP64 factored operator,
   which never forms S: (S + lambda) v for random v to 1e-11, and the
   Cholesky solution's residual measured with the factored operator;
 * the mixed factored operator against the FP64 factored operator on the
   same samples (float accuracy, 2e-6) and one mixed SR step whose FP64
   refined residual reaches 1e-8.
Complex XXZ (N=64, alpha=2, exchange moves, Sz = 0): 4096 chains x 4
sweeps sampled bit-exact against the replay in FP64 and in float theta, and
the complex SR build of the 16,384 samples (column mean, F and S v) against
the oracle's O_c, F and S = O_c^H O_c / Ns."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, M, NS = 100, 400, 16384
DIM = N * M + N + M


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _c3_samples(yq, ctx):
    ctx.rbm_create(N, M)
    ctx.init_random(20241203, 0.01)
    ctx.ham_set(yq.TFI, N, J=-1.0, h=-1.0)
    ctx.sampler_config(yq.LOCAL, NS, precision=yq.FP64, seed=20241203)
    ctx.sample(1, 20)
    sig, _ = ctx.samples()
    return sig, ctx.get_params()


def test_c3_explicit_s_against_factored_operator(yq, ctx):
    sig, p = _c3_samples(yq, ctx)
    rng = np.random.default_rng(3)
    lam = 1e-3
    with yq.Context(0) as f:
        f.rbm_create(N, M)
        f.set_params(p)
        f.ham_set(yq.TFI, N, J=-1.0, h=-1.0)
        f.samples_set(sig, precision=yq.FP64)
        f.sr_build(yq.SOLVE_CG_FACTORED, precision=yq.FP64)
        F_f = f.sr_get(yq.SR_F)
        vs = [rng.normal(size=DIM) for _ in range(2)]
        ref = [f.sr_apply(lam, v) for v in vs]
        ctx.samples_set(sig, precision=yq.FP64)
        ctx.sr_build(yq.SOLVE_EXPLICIT, precision=yq.FP64)
        assert _rel(ctx.sr_get(yq.SR_F), F_f) < 1e-12
        for v, r in zip(vs, ref):
            assert _rel(ctx.sr_apply(lam, v), r) < 1e-11
        x, res, it = ctx.sr_solve(lam)
        assert it == 0 and res < 1e-12
        # residual of the Cholesky solution under the operator that has no S
        assert _rel(f.sr_apply(lam, x), F_f) < 1e-9


def test_c3_mixed_operator_and_step(yq, ctx):
    sig, p = _c3_samples(yq, ctx)
    rng = np.random.default_rng(4)
    v = rng.normal(size=DIM)
    out = {}
    for prec in (yq.FP64, yq.MIXED):
        ctx.samples_set(sig, precision=prec)
        ctx.sr_build(yq.SOLVE_CG_FACTORED, precision=prec)
        out[prec] = ctx.sr_apply(1e-3, v)
    assert _rel(out[yq.MIXED], out[yq.FP64]) < 2e-6
    ctx.sampler_config(yq.LOCAL, NS, precision=yq.MIXED, seed=7)
    o = ctx.sr_step(n_sweeps=1, n_discard=5, solver=yq.SOLVE_CG_FACTORED, precision=yq.MIXED,
                    lam=1e-3, cg_tol=1e-8, cg_max_iter=50000, lr=0.01)
    assert o["cg_err"] < 1e-8 and o["cg_iters"] > 0 and o["n_samples"] == NS
    assert np.isfinite(o["energy"])


@pytest.mark.parametrize("prec", [0, 1])
def test_c2_complex_sampler_full_size(ctx, orc, prec):
    n, m, chains, sweeps = 64, 128, 4096, 4
    ctx.rbm_create(n, m, dtype=np.complex128)
    rng = np.random.default_rng(64)
    rbm = orc.CRBM.random(n, m, 0.01, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    seed = 0xC2C2
    ctx.sampler_config(1, chains, precision=prec, seed=seed)
    ctx.randomize(n // 2)
    s0, _, _ = ctx.sampler_state()
    assert (s0.sum(axis=1) == 0).all()
    ctx.sample(sweeps, 0)
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.c_sample_replay(rbm, s0, seed=seed, n_sweeps=sweeps,
                                                   n_discard=0, kind=1, precision=prec)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(np.ascontiguousarray(gt).view(np.uint64),
                                  np.ascontiguousarray(rt).view(np.uint64))
    assert (gs.sum(axis=1) == 0).all()
    if prec == 0:
        # the complex SR build of the 16,384 samples
        ctx.sr_build(0)
        H = orc.ham(1, n, J=1.0, delta=1.0)
        th = orc.c_theta(rbm, gs)
        el = orc.c_eloc(H, rbm, gs, th)
        O, om, F, _ = orc.c_sr_build(rbm, gs, th, el, with_S=False)
        assert _rel(ctx.sr_get(1), om) < 1e-12
        assert _rel(ctx.sr_get(0), F) < 1e-10
        v = rng.normal(size=len(F)) + 1j * rng.normal(size=len(F))
        Sv = O.conj().T @ (O @ v) / O.shape[0]
        assert _rel(ctx.sr_apply(1e-3, v), Sv + 1e-3 * v) < 1e-11


def test_c3_mixed_explicit(yq, ctx):
    """Mixed explicit solve at full size: split-bf16 S ((S + lambda) v within
    3e-5 of the FP64 factored operator: the pair products are exact to 2^-24,
    the fp32 accumulation over 16,384 samples leaves about 1.6e-5) and FP32
    Cholesky refined in FP64 to a 1e-10 residual of the exact operator."""
    sig, p = _c3_samples(yq, ctx)
    lam = 1e-3
    rng = np.random.default_rng(6)
    v = rng.normal(size=DIM)
    with yq.Context(0) as f:
        f.rbm_create(N, M)
        f.set_params(p)
        f.ham_set(yq.TFI, N, J=-1.0, h=-1.0)
        f.samples_set(sig, precision=yq.FP64)
        f.sr_build(yq.SOLVE_CG_FACTORED, precision=yq.FP64)
        ref = f.sr_apply(lam, v)
        ctx.samples_set(sig, precision=yq.FP64)   # FP64 E_loc and F, split-bf16 S
        ctx.sr_build(yq.SOLVE_EXPLICIT, precision=yq.MIXED)
        assert _rel(ctx.sr_apply(lam, v), ref) < 3e-5
        x, res, rounds = ctx.sr_solve(lam, tol=1e-10)
        assert res < 1e-10 and rounds >= 1
        assert _rel(f.sr_apply(lam, x), f.sr_get(yq.SR_F)) < 1e-9
