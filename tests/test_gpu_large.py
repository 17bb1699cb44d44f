"""Machines past one wave's registers and past the old site limit.

Hidden layers wider than 512 units run one chain per block of G = 64 NW
threads (yq_chain_lanes(m)); their spins, theta, counters and accept counts
must be bit-exact against the oracle's replay with the same reduction width
(rp_sample / rp_c_sample with lanes = G: lane-strided partial sums, then the
tree with strides G/2 .. 1).  E_loc of wide machines reads theta from global
memory instead of staging it; it must keep the bits of the kernel replay.
Chains of up to 4096 sites sample, form E_loc and the FP64 factored SR
operator (whose backward product splits the rows into groups of 9 tiles) to
the oracle's values.  Past n = 159 the mixed factored operator runs its
chunked forward and row-grouped backward products, to float accuracy against
a dense FP64 operator, and its refined solve agrees with the FP64 one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n,m,chains,kind,prec", [
    (24, 600, 8, 0, 0),      # G = 128, 6 slots per thread
    (24, 1000, 8, 1, 1),     # G = 128, 8 slots, exchange moves, float
    (30, 2100, 4, 0, 1),     # G = 512, 6 slots
    (20, 8192, 2, 0, 0),     # G = 1024, 8 slots: the widest machine
])
def test_wide_sampler_bit_exact(ctx, orc, yq, n, m, chains, kind, prec):
    lanes = yq.lib().yq_chain_lanes(m)
    assert lanes == {600: 128, 1000: 128, 2100: 512, 8192: 1024}[m]
    print(f"wide sampler n={n} m={m} lanes={lanes} kind={kind} prec={prec}", flush=True)
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(n + m)
    rbm = orc.RBM.random(n, m, 0.3 / np.sqrt(m / 64), rng)
    ctx.set_params(rbm.params())
    seed = 0xABC + m
    ctx.sampler_config(kind, chains, precision=prec, seed=seed)
    ctx.randomize(n // 2 if kind else -1)
    s0, _, _ = ctx.sampler_state()
    sweeps, discard = 2, 1
    ctx.sample(sweeps, discard)
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.sample_replay(rbm, s0, seed=seed, n_sweeps=sweeps,
                                                 n_discard=discard, kind=kind, precision=prec,
                                                 lanes=lanes)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    assert (acc > 0).all()
    if kind == 1:
        assert (gs.astype(int).sum(axis=1) == 0).all()


@pytest.mark.parametrize("kind,prec", [(0, 0), (1, 1)])
def test_wide_complex_sampler_bit_exact(ctx, orc, yq, kind, prec):
    n, m, chains = 20, 700, 6
    lanes = yq.lib().yq_chain_lanes(m)
    assert lanes == 128
    ctx.rbm_create(n, m, dtype=np.complex128)
    rng = np.random.default_rng(77)
    rbm = orc.CRBM.random(n, m, 0.1, rng)
    ctx.set_params(rbm.params())
    seed = 0xC3 + kind
    ctx.sampler_config(kind, chains, precision=prec, seed=seed)
    ctx.randomize(n // 2 if kind else -1)
    s0, _, _ = ctx.sampler_state()
    ctx.sample(2, 1)
    gs, gt = ctx.samples()
    fin, ctr, acc = ctx.sampler_state()
    rfin, rctr, racc, rs, rt = orc.c_sample_replay(rbm, s0, seed=seed, n_sweeps=2, n_discard=1,
                                                   kind=kind, precision=prec, lanes=lanes)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(fin, rfin)
    np.testing.assert_array_equal(ctr, rctr)
    np.testing.assert_array_equal(acc, racc)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    assert (acc > 0).all()


@pytest.mark.parametrize("kind,kw", [(0, dict(J=-1.0, h=-1.0)),
                                     (2, dict(J=1.0, J2=0.5, sign_rule=1))])
@pytest.mark.parametrize("prec", [0, 1])
def test_wide_eloc_unstaged(ctx, orc, kind, kw, prec):
    """m = 2000: theta is read from global memory (the staged tile would need
    more than 48 KB of LDS); bit-exact against the kernel replay."""
    n, m, ns = 40, 2000, 96
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(3 + kind)
    rbm = orc.RBM.random(n, m, 0.05, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(kind, n, **kw)
    H = orc.ham(kind, n, **kw)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig, precision=prec)
    e, mean, var = ctx.eloc()
    th = orc.theta_replay(rbm, sig)
    if prec == 1:
        th = th.astype(np.float32).astype(np.float64)
    rep = orc.eloc_replay(H, rbm, sig, th, precision=prec)
    np.testing.assert_array_equal(_bits(e), _bits(rep))
    ref = orc.eloc(H, rbm, sig, orc.theta(rbm, sig))
    tol = 1e-10 if prec == 0 else 1e-4
    assert np.abs(e - ref).max() <= tol * max(1.0, np.abs(ref).max())


def test_wide_complex_eloc_unstaged(ctx, orc):
    n, m, ns = 24, 1200, 64
    ctx.rbm_create(n, m, dtype=np.complex128)
    rng = np.random.default_rng(8)
    rbm = orc.CRBM.random(n, m, 0.05, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(1, n, J=1.0, delta=1.0)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    e, _, _ = ctx.eloc()
    H = orc.ham(1, n, J=1.0, delta=1.0)
    th = orc.c_theta(rbm, sig)
    assert _rel(e, orc.c_eloc(H, rbm, sig, th)) < 1e-10
    assert _rel(e, orc.c_eloc_replay(H, rbm, sig, th, precision=0)) < 1e-13


def test_long_chain_sample_eloc_and_factored_operator(ctx, orc, yq):
    """n = 300 > 143: the FP64 backward product runs three row groups."""
    n, m, chains = 300, 64, 16
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(300)
    rbm = orc.RBM.random(n, m, 0.05, rng)
    ctx.set_params(rbm.params())
    ctx.ham_set(0, n, J=-1.0, h=-1.0)
    seed = 0x300
    ctx.sampler_config(0, chains, precision=0, seed=seed)
    s0, _, _ = ctx.sampler_state()
    ctx.sample(2, 1)
    gs, gt = ctx.samples()
    _, _, _, rs, rt = orc.sample_replay(rbm, s0, seed=seed, n_sweeps=2, n_discard=1, kind=0,
                                        precision=0, lanes=64)
    np.testing.assert_array_equal(gs, rs)
    np.testing.assert_array_equal(_bits(gt), _bits(rt))
    ns = 256
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    ctx.samples_set(sig)
    e, _, _ = ctx.eloc()
    H = orc.ham(0, n, J=-1.0, h=-1.0)
    ref = orc.eloc(H, rbm, sig, orc.theta(rbm, sig))
    assert np.abs(e - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
    ctx.sr_build(2)
    T = np.tanh(orc.theta_replay(rbm, sig))
    Ow = (sig[:, :, None] * T[:, None, :]).reshape(ns, -1)
    O = np.concatenate([Ow, sig.astype(float), T], axis=1)
    Oc = O - O.mean(axis=0)
    v = rng.normal(size=O.shape[1])
    assert _rel(ctx.sr_apply(1e-3, v), Oc.T @ (Oc @ v) / ns + 1e-3 * v) < 1e-12


def test_site_limits(ctx, yq):
    ctx.rbm_create(4096, 8)
    with pytest.raises(yq.YqError):
        ctx.rbm_create(4097, 8)
    with pytest.raises(yq.YqError):
        ctx.rbm_create(16, 8193)


@pytest.mark.parametrize("n,m,ns", [(200, 48, 2048), (333, 40, 1536)])
def test_mixed_factored_long_chain(ctx, orc, n, m, ns):
    """n >= 160: the chunked forward product (K in chunks of 4 k-steps) and the
    row-grouped backward product compute the operator to float accuracy for
    both column-group widths and every backward chunk; the refined solve
    reaches 1e-8 and agrees with the FP64 factored solve of the same batch; captured graphs
    and eager launches give the same bits."""
    print(f"mixed long chain n={n} m={m} ns={ns}", flush=True)
    ctx.rbm_create(n, m)
    rng = np.random.default_rng(n)
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
    for ct in (2, 4):
        for chunk in (128, 256, 512):
            ctx.tiles(ct, chunk)
            ctx.sr_build(2, precision=1)
            assert _rel(ctx.sr_apply(1e-3, v), ref) < 2e-6, (ct, chunk)
    ctx.tiles(8, 256)
    with pytest.raises(Exception):
        ctx.sr_build(2, precision=1)
    ctx.tiles(0, 0)
    ctx.sr_build(2, precision=1)
    out = {}
    for mode in (False, True):
        ctx.graph_mode(mode)
        out[mode] = ctx.sr_solve(1e-3, tol=1e-8, max_iter=20000)
    ctx.graph_mode(True)
    (xe, re_, ie), (xg, rg, ig) = out[False], out[True]
    assert re_ < 1e-8 and ie == ig
    np.testing.assert_array_equal(xe, xg)
    # the FP64 factored solve of the same (float theta) batch; both answers
    # are within 1e-8 relative residual, so apart by up to cond(A) 1e-8
    ctx.sr_build(2)
    x64, r64, _ = ctx.sr_solve(1e-3, tol=1e-10, max_iter=20000)
    assert r64 < 1e-10
    assert _rel(xg, x64) < 1e-5
