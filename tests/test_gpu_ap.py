"""AmplitudePhase machine on the GPU (Machines/AmplitudePhase.hpp,
States/AmplitudePhaseState.hpp, GroundState/NGDExact.hpp): psi = sqrt(RBM)
exp(i pi FF) against the oracle's restatement (oracle/oracle_ap.cpp).
  * the sampler at beta = 1/2 is bit-exact against the kernel replay;
  * complex E_loc, the rows [O_RBM / 2 | pi dFF], F and the block-diagonal S
    of a sampled batch agree with the oracle to 1e-10 relative (FP64);
  * the exact basis batch reproduces NGDExact's energy, gradient (= 2 F) and
    corrMatAmp / corrMatPhase, and a few steps of the reference test's exact
    SR loop (Tests/TestAmplitudePhase.cpp:50-76) track the oracle's loop."""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

FF_SHAPE = dict(channels=2, kernel=3, hidden=4)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _setup(ctx, orc, n, m, rng, use_bias=True, amp=0.15, phase=0.4):
    rbm = orc.RBM.random(n, m, amp, rng, use_bias=use_bias)
    desc = orc.FFDesc(n, FF_SHAPE["channels"], FF_SHAPE["kernel"], FF_SHAPE["hidden"])
    ffp = rng.normal(0, phase, orc.ff_npar(desc))
    ctx.rbm_create(n, m, use_bias=use_bias)
    ctx.set_params(rbm.params())
    ctx.ff_create(n, dtype=np.float64, **FF_SHAPE)
    ctx.ff_set_params(ffp)
    ctx.ap_enable(True)
    return rbm, desc, ffp


HAMS = {
    "tfi": dict(kind=0, J=-1.0, h=-0.7),
    "xxz": dict(kind=1, J=1.0, delta=0.6, sign_rule=1),
    "j1j2": dict(kind=2, J=1.0, J2=0.44),
}


def _ham(ctx, orc, name, n):
    kw = dict(HAMS[name])
    kind = kw.pop("kind")
    ctx.ham_set(kind, n, **kw)
    return orc.ham(kind, n, **kw)


def test_ap_dims_params_log_psi(ctx, orc):
    rng = np.random.default_rng(41)
    n, m = 12, 24
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng)
    dA, dP = ctx.ap_dims()
    assert dA == rbm.dim and dP == orc.ff_npar(desc)
    p = ctx.ap_get_params()
    np.testing.assert_array_equal(p, np.concatenate([rbm.params(), ffp]))
    sig = orc.basis_sigma(n)[::37]
    got = ctx.ap_log_psi(sig)
    want = orc.ap_log_psi(rbm, desc, ffp, sig)
    np.testing.assert_allclose(got.real, want.real, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(got.imag, want.imag, rtol=0, atol=1e-12)
    # updateParams moves both halves (AmplitudePhase.hpp:153-158)
    d = rng.normal(0, 1e-2, dA + dP)
    ctx.ap_update_params(d)
    np.testing.assert_allclose(ctx.ap_get_params(), p + d, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ctx.ff_get_params(), ffp + d[dA:], rtol=0, atol=1e-15)


@pytest.mark.parametrize("kind,prec", [(0, 0), (0, 1), (1, 0)])
def test_ap_sampler_is_half_beta_replay(ctx, orc, kind, prec):
    rng = np.random.default_rng(42)
    n, m, chains, sweeps, discard = 16, 32, 48, 4, 3
    rbm, _, _ = _setup(ctx, orc, n, m, rng, amp=0.3)
    seed = 0xA11CE + kind
    ctx.sampler_config(kind, chains, precision=prec, seed=seed)
    ctx.randomize(n // 2 if kind else -1)
    s0, _, _ = ctx.sampler_state()
    ctx.sample(sweeps, discard)
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.sample_replay(rbm, s0, seed=seed, n_sweeps=sweeps,
                                                 n_discard=discard, kind=kind, precision=prec,
                                                 beta=0.5)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(np.ascontiguousarray(gt).view(np.uint64),
                                  np.ascontiguousarray(rt).view(np.uint64))
    # and not the beta = 1 chain
    r1 = orc.sample_replay(rbm, s0, seed=seed, n_sweeps=sweeps, n_discard=discard, kind=kind,
                           precision=prec)
    assert not np.array_equal(r1[2], racc)


@pytest.mark.parametrize("name", ["tfi", "xxz", "j1j2"])
def test_ap_eloc_and_sr_of_samples(ctx, orc, name):
    rng = np.random.default_rng(43)
    n, m = 14, 28
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng, amp=0.2, phase=0.5)
    H = _ham(ctx, orc, name, n)
    ctx.sampler_config(1 if name != "tfi" else 0, 64, seed=7)
    ctx.randomize(n // 2 if name != "tfi" else -1)
    ns = ctx.sample(8, 4)
    sig, _ = ctx.samples()
    th = orc.theta(rbm, sig)
    el, mean, var = ctx.eloc()
    want = orc.ap_eloc(H, rbm, desc, ffp, sig, th)
    assert _rel(el, want) < 1e-10
    assert abs(mean - want.real.mean()) < 1e-10 * max(1.0, abs(mean))
    assert abs(var - ((want.real ** 2).mean() - want.real.mean() ** 2)) < 1e-9 * max(1.0, var)
    # SR: rows [O_RBM / 2 | pi dFF], F = [cov(O_amp, Re E) | cov(O_phase, Im E)], S blocks
    rows = orc.ap_logderiv(rbm, desc, ffp, sig, th)
    dA = rbm.dim
    om = rows.mean(axis=0)
    Oc = rows - om
    F = np.concatenate([Oc[:, :dA].T @ want.real, Oc[:, dA:].T @ want.imag]) / ns
    S = np.zeros((len(om), len(om)))
    S[:dA, :dA] = Oc[:, :dA].T @ Oc[:, :dA] / ns
    S[dA:, dA:] = Oc[:, dA:].T @ Oc[:, dA:] / ns
    ctx.sr_build(0)
    assert _rel(ctx.sr_get(1), om) < 1e-11
    assert _rel(ctx.sr_get(2), Oc) < 1e-11
    assert _rel(ctx.sr_get(0), F) < 1e-10
    assert _rel(ctx.sr_get(3), S) < 1e-10
    shift = 1e-3
    x, res, _ = ctx.sr_solve(shift, tol=1e-12, max_iter=1)
    want_x = np.linalg.solve(S + shift * np.eye(len(om)), F)
    assert _rel(x, want_x) < 1e-8
    v = rng.normal(size=len(om))
    assert _rel(ctx.sr_apply(shift, v), S @ v + shift * v) < 1e-11
    # the dense CG solves the same block system
    ctx.sr_build(1)
    x2, _, it = ctx.sr_solve(shift, tol=1e-11, max_iter=20000)
    assert _rel(x2, want_x) < 1e-8 and it > 0
    assert _rel(ctx.sr_apply(shift, v), S @ v + shift * v) < 1e-11


def test_ap_mixed_batch(ctx, orc, yq):
    """A float-theta batch (mixed sampling): E_loc and the SR rows in float
    arithmetic for the amplitude ratios, against the FP64 oracle on the same
    configurations (the stated mixed-precision error)."""
    rng = np.random.default_rng(49)
    n, m = 20, 40
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng, amp=0.2, phase=0.5)
    H = _ham(ctx, orc, "j1j2", n)
    ctx.sampler_config(1, 256, precision=yq.MIXED, seed=13)
    ctx.randomize(n // 2)
    ns = ctx.sample(4, 4)
    sig, _ = ctx.samples()
    th = orc.theta(rbm, sig)
    el, mean, _ = ctx.eloc()
    want = orc.ap_eloc(H, rbm, desc, ffp, sig, th)
    assert _rel(el, want) < 2e-5
    assert abs(mean - want.real.mean()) < 2e-5 * abs(mean)
    ctx.sr_build(0)
    rows = orc.ap_logderiv(rbm, desc, ffp, sig, th)
    assert _rel(ctx.sr_get(2), rows - rows.mean(axis=0)) < 1e-5
    dA = rbm.dim
    Oc = rows - rows.mean(axis=0)
    F = np.concatenate([Oc[:, :dA].T @ want.real, Oc[:, dA:].T @ want.imag]) / ns
    assert _rel(ctx.sr_get(0), F) < 1e-4


@pytest.mark.parametrize("name,nup,use_bias", [("tfi", -1, True), ("j1j2", 5, False),
                                               ("xxz", 5, True)])
def test_ap_exact_matches_ngd_exact(ctx, orc, name, nup, use_bias):
    rng = np.random.default_rng(44)
    n, m = 10, 20
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng, use_bias=use_bias, amp=0.25, phase=0.5)
    H = _ham(ctx, orc, name, n)
    B = ctx.exact_batch(nup)
    basis = np.arange(1 << n, dtype=np.uint32) if nup < 0 else orc.basis_jz(n, nup)
    assert B == len(basis)
    r = orc.ap_ngd_exact(H, rbm, desc, ffp, basis)
    # weights |psi|^2 = RBM coeff / sum
    sig = orc.basis_to_sigma(basis, n)
    lp = orc.ap_log_psi(rbm, desc, ffp, sig).real
    p = np.exp(2 * (lp - lp.max()))
    np.testing.assert_allclose(ctx.exact_probs(), p / p.sum(), rtol=1e-12, atol=1e-300)
    _, mean, _ = ctx.eloc()
    assert abs(mean - r["E"]) < 1e-11 * max(1.0, abs(r["E"]))
    ctx.sr_build(0)
    dA = rbm.dim
    np.testing.assert_allclose(ctx.sr_get(1), np.concatenate([r["oloc_amp"], r["oloc_phase"]]),
                               rtol=0, atol=1e-12)
    # NGDExact's gradient is twice F (the reference's phase part drops the
    # centring term, which vanishes: sum_s p_s Im E_loc(s) = Im <psi|H|psi> = 0)
    assert _rel(2 * ctx.sr_get(0), r["grad"]) < 1e-10
    S = ctx.sr_get(3)
    assert _rel(S[:dA, :dA], r["S_amp"]) < 1e-10
    assert _rel(S[dA:, dA:], r["S_phase"]) < 1e-10
    assert not S[:dA, dA:].any() and not S[dA:, :dA].any()


def test_ap_exact_sr_loop_tracks_oracle(ctx, orc):
    """Tests/TestAmplitudePhase.cpp:50-76 for a few steps: SRMatExact with
    corrMat().real() + lambda I, LLT, SGD(0.02); J1-J2 in the Sz = 0 sector."""
    rng = np.random.default_rng(45)
    n, m = 10, 20
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng, amp=0.05, phase=0.3)
    H = _ham(ctx, orc, "j1j2", n)
    basis = orc.basis_jz(n, n // 2)
    lam, lr = 1e-3, 0.02
    dA = rbm.dim
    p = np.concatenate([rbm.params(), ffp])
    energies = []
    for _ in range(4):
        ctx.exact_batch(n // 2)
        _, e_gpu, _ = ctx.eloc()
        ctx.sr_build(0)
        x, _, _ = ctx.sr_solve(lam, tol=1e-12, max_iter=1)
        ctx.ap_update_params(-lr * x)
        r = orc.ap_ngd_exact(H, rbm, desc, p[dA:], basis)
        assert abs(e_gpu - r["E"]) < 1e-10 * abs(r["E"])
        energies.append(r["E"])
        xa = np.linalg.solve(r["S_amp"] + lam * np.eye(dA), r["grad"][:dA] / 2)
        xp = np.linalg.solve(r["S_phase"] + lam * np.eye(len(p) - dA), r["grad"][dA:] / 2)
        p = p - lr * np.concatenate([xa, xp])
        rbm.set_params(p[:dA])
        assert _rel(ctx.ap_get_params(), p) < 1e-9
    assert energies[-1] < energies[0]


def test_ap_sr_step_equals_its_parts(ctx, orc, yq):
    rng = np.random.default_rng(46)
    n, m = 16, 32
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng, amp=0.1, phase=0.3)
    _ham(ctx, orc, "xxz", n)
    ctx.sampler_config(1, 128, seed=11)
    ctx.randomize(n // 2)
    s0, _, _ = ctx.sampler_state()
    p0 = ctx.ap_get_params()
    out = ctx.sr_step(n_sweeps=4, n_discard=2, solver=yq.SOLVE_EXPLICIT, lam=1e-2, lr=0.05)
    p1 = ctx.ap_get_params()
    assert np.isfinite(out["energy"]) and out["n_samples"] == 512
    # the same step by hand on a second context from the same state
    with yq.Context(0) as c2:
        c2.rbm_create(n, m)
        c2.set_params(rbm.params())
        c2.ff_create(n, dtype=np.float64, **FF_SHAPE)
        c2.ff_set_params(ffp)
        c2.ap_enable(True)
        c2.ham_set(1, n, J=1.0, delta=0.6, sign_rule=1)
        c2.sampler_config(1, 128, seed=11)
        c2.set_sigma(s0)
        c2.sample(4, 2)
        _, e, _ = c2.eloc()
        c2.sr_build(0)
        x, _, _ = c2.sr_solve(1e-2, tol=1e-12, max_iter=1)
        assert abs(out["energy"] - e) <= 1e-14 * abs(e)
        np.testing.assert_allclose(p1, p0 - 0.05 * x, rtol=0, atol=1e-14)
    assert np.abs(p1[rbm.dim:] - p0[rbm.dim:]).max() > 0
    assert np.abs(p1[:rbm.dim] - p0[:rbm.dim]).max() > 0


def test_ap_exact_sampler_draws_rbm_weights(ctx, orc):
    rng = np.random.default_rng(47)
    n, m = 8, 8
    rbm, desc, ffp = _setup(ctx, orc, n, m, rng, amp=0.5)
    ctx.ham_set(0, n)
    idx = ctx.exact_sample(200000, seed=5)
    sig = orc.basis_sigma(n)
    th = orc.theta(rbm, sig)
    lc = np.array([orc.log_coeff(rbm, sig[s], th[s]) for s in range(len(sig))])
    p = np.exp(lc - lc.max())
    p /= p.sum()
    cnt = np.bincount(idx, minlength=len(p))
    chi2 = ((cnt - len(idx) * p) ** 2 / (len(idx) * p)).sum()
    assert stats.chi2.sf(chi2, len(p) - 1) > 1e-4, chi2


def test_ap_rejections(ctx, orc, yq):
    rng = np.random.default_rng(48)
    n, m = 8, 16
    _setup(ctx, orc, n, m, rng)
    with pytest.raises(yq.YqError):
        ctx.sampler_config(0, 8, n_temps=2)       # tempering
    ctx.ham_set(0, n)
    ctx.sampler_config(0, 8)
    ctx.sample(2, 0)
    with pytest.raises(yq.YqError):
        ctx.sr_build(2)                           # factored solver: RBM only
    with pytest.raises(yq.YqError):
        ctx.sr_build(1, yq.MIXED)
    ctx.ff_create(n + 1, dtype=np.float64, **FF_SHAPE)   # disables AP
    assert not ctx.ap
    with pytest.raises(yq.YqError):
        ctx.ap_enable(True)                       # sizes differ
    ctx.ff_create(n, dtype=np.float32, **FF_SHAPE)
    with pytest.raises(yq.YqError):
        ctx.ap_enable(True)                       # float phase network
    ctx.rbm_create(n, m, dtype=np.complex128)
    ctx.ff_create(n, dtype=np.float64, **FF_SHAPE)
    with pytest.raises(yq.YqError):
        ctx.ap_enable(True)                       # complex amplitude
