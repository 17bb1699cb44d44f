"""Parallel tempering on the GPU (SamplerMT with nTmps > 1, SamplerMT.hpp:44-60,
120-153): spins, theta, counters and accept counts of every chain state
bit-exact against the oracle's tempering replay (rp_sample_pt, itself pinned
against a libm restatement with or_log_coeff swaps and a chi^2 test), for
FP64 and mixed, local and exchange sweeps, 2 to 16 temperatures; continued
streams; an SR step on tempered samples.  Complex machines the same way
against rp_c_sample_pt (pinned by a libm restatement, by the real replay at
Im = 0 and by a chi^2 test), including the XXZ shape N = 64, alpha = 2 with
16 temperatures of Examples/RBMXXZ.cpp:45,81."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("n,m,T,groups,kind,prec,sweeps,discard", [
    (12, 24, 4, 8, 0, 0, 3, 2),
    (12, 24, 4, 8, 0, 1, 3, 2),
    (16, 32, 2, 5, 1, 0, 2, 3),
    (16, 32, 16, 3, 1, 1, 2, 1),
    (20, 40, 16, 4, 0, 0, 2, 2),
    (37, 70, 6, 3, 0, 1, 2, 1),
])
def test_tempering_bit_exact(ctx, orc, n, m, T, groups, kind, prec, sweeps, discard):
    rng = np.random.default_rng(n + T)
    rbm = orc.RBM.random(n, m, 0.3, rng)
    ctx.rbm_create(n, m)
    ctx.set_params(rbm.params())
    seed = 0xBEEF + n
    ctx.sampler_config(kind, groups, precision=prec, seed=seed, n_temps=T)
    ctx.randomize(n // 2 if kind else -1)
    s0, ctr0, acc0 = ctx.sampler_state()
    assert s0.shape == (T * groups, n)
    ns = ctx.sample(sweeps, discard)
    assert ns == groups * sweeps
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.sample_replay_pt(rbm, s0, n_temps=T, seed=seed,
                                                    n_sweeps=sweeps, n_discard=discard,
                                                    kind=kind, precision=prec)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    # swaps happened: some final state differs from a replay without swap moves
    # is not available; instead the hotter slots accept more often
    a = acc.reshape(groups, T).mean(axis=0)
    assert a[-1] >= a[0]
    if kind == 1:
        assert (fin.astype(int).sum(axis=1) == 0).all()
    # a second call continues every stream
    ctx.sample(1, 0)
    fin2, ctr2, acc2 = ctx.sampler_state()
    rfin2, rctr2, racc2, _, _ = orc.sample_replay_pt(rbm, rfin, n_temps=T, seed=seed,
                                                     n_sweeps=1, n_discard=0, kind=kind,
                                                     precision=prec, ctr=rctr, acc=racc)
    np.testing.assert_array_equal(fin2, rfin2)
    np.testing.assert_array_equal(ctr2, rctr2)
    np.testing.assert_array_equal(acc2, racc2)


def test_tempering_sr_step(ctx):
    """An SR epoch body on samples of a tempered sampler (16 temperatures as in
    Examples/RBMTFI.cpp:25)."""
    n = 20
    ctx.rbm_create(n, n)
    ctx.init_random(3, 0.01)
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    ctx.sampler_config(0, 64, precision=0, seed=5, n_temps=16)
    for i in range(3):
        o = ctx.sr_step(n_sweeps=16, n_discard=20 if i == 0 else 0, solver=0, precision=0,
                        lam=1e-3, cg_tol=1e-8, lr=0.01)
        assert np.isfinite(o["energy"]) and o["n_samples"] == 64 * 16


def test_tempering_rejects_bad_configs(ctx, yq):
    ctx.rbm_create(8, 8)
    for T in (3, 18, 0):
        with pytest.raises(yq.YqError):
            ctx.sampler_config(0, 4, n_temps=T)


@pytest.mark.parametrize("n,m,T,groups,kind,prec,sweeps,discard", [
    (12, 24, 4, 8, 0, 0, 3, 2),
    (12, 24, 4, 8, 1, 1, 3, 2),
    (16, 32, 2, 5, 1, 0, 2, 3),
    (20, 40, 16, 3, 0, 1, 2, 1),
    (37, 70, 6, 3, 1, 0, 2, 1),
    (64, 128, 16, 4, 1, 0, 2, 2),
    (64, 128, 16, 4, 1, 1, 2, 2),
])
def test_complex_tempering_bit_exact(ctx, orc, n, m, T, groups, kind, prec, sweeps, discard):
    rng = np.random.default_rng(100 + n + T)
    rbm = orc.CRBM.random(n, m, 0.3 if n < 64 else 0.05, rng)
    ctx.rbm_create(n, m, dtype=np.complex128)
    ctx.set_params(rbm.params())
    seed = 0xC0FFEE + n
    ctx.sampler_config(kind, groups, precision=prec, seed=seed, n_temps=T)
    ctx.randomize(n // 2 if kind else -1)
    s0, _, _ = ctx.sampler_state()
    assert s0.shape == (T * groups, n)
    assert ctx.sample(sweeps, discard) == groups * sweeps
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.c_sample_replay_pt(rbm, s0, n_temps=T, seed=seed,
                                                      n_sweeps=sweeps, n_discard=discard,
                                                      kind=kind, precision=prec)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    if kind == 1:   # exchange moves keep every chain's magnetisation
        np.testing.assert_array_equal(fin.astype(int).sum(axis=1), s0.astype(int).sum(axis=1))
    ctx.sample(1, 0)
    fin2, ctr2, acc2 = ctx.sampler_state()
    rfin2, rctr2, racc2, _, _ = orc.c_sample_replay_pt(rbm, rfin, n_temps=T, seed=seed,
                                                       n_sweeps=1, n_discard=0, kind=kind,
                                                       precision=prec, ctr=rctr, acc=racc)
    np.testing.assert_array_equal(fin2, rfin2)
    np.testing.assert_array_equal(ctr2, rctr2)
    np.testing.assert_array_equal(acc2, racc2)


def test_complex_tempering_sr_step(ctx):
    """An SR epoch body of the complex XXZ machine on tempered exchange-move
    samples (16 temperatures x 1 chain per group as in Examples/RBMXXZ.cpp:81,
    here 32 groups)."""
    n = 16
    ctx.rbm_create(n, 2 * n, dtype=np.complex128)
    ctx.init_random(4, 0.01)
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    ctx.sampler_config(1, 32, precision=0, seed=6, n_temps=16)
    ctx.randomize(n // 2)
    for i in range(3):
        o = ctx.sr_step(n_sweeps=8, n_discard=10 if i == 0 else 0, solver=0, precision=0,
                        lam=1e-3, cg_tol=1e-8, lr=0.01)
        assert np.isfinite(o["energy"]) and o["n_samples"] == 32 * 8
