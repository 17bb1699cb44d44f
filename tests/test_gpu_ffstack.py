"""General FeedForward stacks on the GPU (yq_ff_create_stack, ffstack.hip)
against the oracle's restatement (oracle/oracle_ffs.cpp, itself pinned
against torch and finite differences by test_oracle_ffs.py):
  * log psi and the Jacobian rows of FeedForward::backward
    (Machines/FeedForward.hpp:152-205) for Conv1D with channels, stride and
    bias, FullyConnected with bias, all twelve activations and Mean over
    both axes: FP64 to 1e-12, FP32 to 2e-5 relative;
  * the VJP with stored and recomputed activations against J^T r;
  * the C4 stack routed to the C4 kernels gives the C4 network's results;
  * an AmplitudePhase machine with the reference example's phase network
    (Examples/AmplitudePhase.cpp:42-55): complex E_loc, rows, F and S of a
    sampled batch against the oracle's AmplitudePhase restatement."""
import numpy as np
import pytest

from ffstacks import AP_EXAMPLE_40, SIGMOID_STACK, STACKS

pytestmark = pytest.mark.gpu

CASES = dict(STACKS, sigmoid=SIGMOID_STACK)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _batch(rng, ns, n):
    return rng.choice(np.array([-1, 1], dtype=np.int8), size=(ns, n))


def _make(ctx, orc, n, layers, dtype, seed, scale=0.4):
    print(f"stack n={n} layers={len(layers)} dtype={np.dtype(dtype).name}", flush=True)
    desc = orc.ff_stack(n, layers)
    ctx.ff_create_stack(n, layers, dtype)
    assert ctx.ff_npar == orc.ff_npar(desc)
    p = np.random.default_rng(seed).normal(0.0, scale, ctx.ff_npar)
    ctx.ff_set_params(p)
    np.testing.assert_array_equal(ctx.ff_get_params(), p)
    return desc, p


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
@pytest.mark.parametrize("name", sorted(CASES))
def test_stack_forward_jacobian(ctx, orc, name, dtype, tol):
    n, layers = CASES[name]
    desc, p = _make(ctx, orc, n, layers, dtype, 21)
    sig = _batch(np.random.default_rng(22), 37, n)   # a ragged last group
    Jw, ow = orc.ff_jacobian(desc, p, sig)
    out = ctx.ff_forward(sig)
    J, out2 = ctx.ff_jacobian(sig)
    assert np.abs(out - ow).max() <= tol * max(1.0, np.abs(ow).max())
    assert np.abs(out2 - ow).max() <= tol * max(1.0, np.abs(ow).max())
    assert np.abs(J - Jw).max() <= tol * max(1.0, np.abs(Jw).max()), name


def test_stack_amplitude_phase_example_full_size(ctx, orc):
    """The example's own network at N = 40 over 4,096 configurations (rows
    checked on a subset against the oracle)."""
    n, layers = AP_EXAMPLE_40
    desc, p = _make(ctx, orc, n, layers, np.float64, 23, scale=0.3)
    sig = _batch(np.random.default_rng(24), 4096, n)
    J, out = ctx.ff_jacobian(sig)
    idx = np.arange(0, 4096, 97)
    Jw, ow = orc.ff_jacobian(desc, p, sig[idx])
    assert np.abs(out[idx] - ow).max() <= 1e-12 * max(1.0, np.abs(ow).max())
    assert np.abs(J[idx] - Jw).max() <= 1e-12 * max(1.0, np.abs(Jw).max())
    np.testing.assert_allclose(ctx.ff_forward(sig), out, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", ["amplitude_phase_example", "stride_bias", "activations_mean"])
def test_stack_vjp_store_and_recompute(ctx, orc, name, dtype):
    torch = pytest.importorskip("torch")
    from yannq_amd.torch_ff import FFMachine
    n, layers = CASES[name]
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    print(f"vjp {name} {np.dtype(dtype).name}", flush=True)
    m = FFMachine(n, ctx=ctx, dtype=tdt, layers=layers, seed=5, init_mode=1)
    assert m.act_width == ctx.ff_act_len() == n + sum(
        _widths(n, layers))
    rng = np.random.default_rng(25)
    sig = _batch(rng, 1000, n)
    st = torch.tensor(sig, device="cuda:0")
    J = m.jacobian(st).double().cpu().numpy()
    desc = orc.ff_stack(n, layers)
    Jw, _ = orc.ff_jacobian(desc, ctx.ff_get_params(), sig[:50])
    tol = 1e-11 if dtype == np.float64 else 3e-5
    assert np.abs(J[:50] - Jw).max() <= tol * max(1.0, np.abs(Jw).max())
    r = rng.normal(size=len(sig))
    want = J.T @ r
    for budget in (1 << 40, 0):       # stored activations, then recomputed
        m.act_budget = budget
        f = m.logpsi(st)
        assert m.last_stored == (budget > 0)
        g, = torch.autograd.grad(f, m.params, torch.tensor(r, dtype=tdt, device="cuda:0"))
        got = g.double().cpu().numpy()
        assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max())


def _widths(n, layers):
    w, out = n, []
    for (kind, a, b, c, d, _, _) in layers:
        if kind == 1:
            w = (w // d // a) * b
        elif kind == 2:
            w = b
        elif kind == 4:
            w = b if c == 0 else a
        out.append(w)
    return out


def test_stack_torch_gradcheck(ctx):
    torch = pytest.importorskip("torch")
    from yannq_amd.torch_ff import FFMachine
    n, layers = CASES["stride_bias"]
    m = FFMachine(n, ctx=ctx, dtype=torch.float64, layers=layers, seed=8, init_mode=0)
    sig = torch.tensor(np.random.default_rng(8).choice([-1, 1], (5, n)), dtype=torch.int8,
                       device="cuda:0")
    p = m.params.detach().clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda q: m.logpsi(sig, q), (p,), eps=1e-6, atol=1e-8)


def test_c4_stack_uses_the_c4_network(ctx, orc):
    n, layers = CASES["c4"]
    _, p = _make(ctx, orc, n, layers, np.float64, 26)
    # the C4 kernels store 2 C n + H activation values per configuration
    assert ctx.ff_act_len() == 2 * 3 * n + 6
    sig = _batch(np.random.default_rng(27), 64, n)
    J1, o1 = ctx.ff_jacobian(sig)
    ctx.ff_create(n, 3, 3, 6, np.float64)
    ctx.ff_set_params(p)
    J2, o2 = ctx.ff_jacobian(sig)
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(J1, J2)


def test_stack_init_random_fans(ctx, orc):
    n, layers = CASES["stride_bias"]
    ctx.ff_create_stack(n, layers, np.float64)
    ctx.ff_init_random(99, 1)   # Xavier: sigma = sqrt(2 / (fanIn + fanOut))
    p = ctx.ff_get_params()
    # Conv1D(1, 4, 3, bias): 12 + 4 parameters, fanIn 3, fanOut 12
    s0 = np.sqrt(2.0 / (3 + 12))
    assert 0.2 * s0 < p[:16].std() < 3 * s0
    ctx.ff_init_random(99, 1)
    np.testing.assert_array_equal(ctx.ff_get_params(), p)


def test_stack_rejections(ctx, yq):
    C1, FC, ACT = 1, 2, 3
    with pytest.raises(yq.YqError):
        ctx.ff_create_stack(10, [(FC, 9, 1, 0, 0, 0, 0)])
    with pytest.raises(yq.YqError):
        ctx.ff_create_stack(10, [(FC, 10, 2, 0, 0, 0, 0)])
    with pytest.raises(yq.YqError):
        ctx.ff_create_stack(10, [(C1, 3, 1, 3, 1, 0, 0), (FC, 3, 1, 0, 0, 0, 0)])
    with pytest.raises(yq.YqError):
        ctx.ff_create_stack(10, [(FC, 10, 1, 0, 0, 0, 0), (ACT, 12, 0, 0, 0, 0, 0)])
    with pytest.raises(yq.YqError):
        ctx.ff_create_stack(10, [(FC, 10, 1, 0, 0, 0, 0), (ACT, 2, 0, 0, 0, 0, 0.0)])


@pytest.mark.parametrize("name", ["tfi", "j1j2"])
def test_ap_with_example_network(ctx, orc, name):
    """AmplitudePhase with the example's phase network (N = 12): E_loc of
    sampled configurations (the phase network evaluated on every neighbour),
    rows [O_RBM / 2 | pi dFF], F and both S blocks against the oracle."""
    rng = np.random.default_rng(28)
    n, layers = STACKS["amplitude_phase_example"]
    m = 24
    rbm = orc.RBM.random(n, m, 0.2, rng, use_bias=True)
    desc = orc.ff_stack(n, layers)
    ffp = rng.normal(0, 0.4, orc.ff_npar(desc))
    ctx.rbm_create(n, m, use_bias=True)
    ctx.set_params(rbm.params())
    ctx.ff_create_stack(n, layers, np.float64)
    ctx.ff_set_params(ffp)
    ctx.ap_enable(True)
    if name == "tfi":
        ctx.ham_set(0, n, J=-1.0, h=-0.7)
        H = orc.ham(0, n, J=-1.0, h=-0.7)
        ctx.sampler_config(0, 64, seed=3)
        ctx.randomize(-1)
    else:
        ctx.ham_set(2, n, J=1.0, J2=0.44)
        H = orc.ham(2, n, J=1.0, J2=0.44)
        ctx.sampler_config(1, 64, seed=3)
        ctx.randomize(n // 2)
    ns = ctx.sample(6, 4)
    sig, _ = ctx.samples()
    th = orc.theta(rbm, sig)
    el, _, _ = ctx.eloc()
    want = orc.ap_eloc(H, rbm, desc, ffp, sig, th)
    assert _rel(el, want) < 1e-10
    rows = orc.ap_logderiv(rbm, desc, ffp, sig, th)
    dA = rbm.dim
    om = rows.mean(axis=0)
    Oc = rows - om
    F = np.concatenate([Oc[:, :dA].T @ want.real, Oc[:, dA:].T @ want.imag]) / ns
    ctx.sr_build(0)
    assert _rel(ctx.sr_get(1), om) < 1e-11
    assert _rel(ctx.sr_get(0), F) < 1e-10
    S = ctx.sr_get(3)
    assert _rel(S[dA:, dA:], Oc[:, dA:].T @ Oc[:, dA:] / ns) < 1e-10
    assert _rel(S[:dA, :dA], Oc[:, :dA].T @ Oc[:, :dA] / ns) < 1e-10
