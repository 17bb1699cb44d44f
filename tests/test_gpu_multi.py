"""Many machines in one context (yq_multi_*): one persistent sampler launch
over every (machine, chain) pair and one E_loc launch for all machines.

Each machine must behave exactly like a single-machine context with the same
parameters and seed: spins, theta and counters bit-exact against the
sampler replay per machine, local energies against the E_loc replay, and
whole epoch bodies bit-identical to independent contexts.  The C5
configuration (32 TFI machines of uneven size) runs end to end."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16), (24, 48), (40, 160), (64, 64), (100, 400), (128, 512), (37, 70)]


def _machines(orc, shapes, chains, rng, scale=0.2):
    out, rbms = [], []
    for q, (n, m) in enumerate(shapes):
        rbm = orc.RBM.random(n, m, scale, rng)
        h = 0.5 + q / max(len(shapes) - 1, 1)
        out.append(dict(n=n, m=m, chains=chains[q] if isinstance(chains, list) else chains,
                        seed=1000 + 17 * q, ham=(0, dict(J=-1.0, h=-h))))
        rbms.append(rbm)
    return out, rbms


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("kind", [0, 1])
def test_multi_sampler_bit_exact(ctx, orc, prec, kind):
    rng = np.random.default_rng(3 + prec + 2 * kind)
    chains = [9, 16, 5, 12, 8, 7, 11]
    ms, rbms = _machines(orc, SHAPES, chains, rng)
    ctx.multi_create(ms, sweeper=kind, precision=prec)
    for q, rbm in enumerate(rbms):
        ctx.multi_set_params(q, rbm.params())
    s0 = [ctx.multi_state(q)[0] for q in range(len(ms))]
    ctx.multi_sample(3, 2)
    for q, (mc, rbm) in enumerate(zip(ms, rbms)):
        gs, gt = ctx.multi_samples(q)
        fin, ctr, acc = ctx.multi_state(q)
        rfin, rctr, racc, rs, rt = orc.sample_replay(rbm, s0[q], seed=mc["seed"], n_sweeps=3,
                                                     n_discard=2, kind=kind, precision=prec)
        np.testing.assert_array_equal(gs, rs)
        np.testing.assert_array_equal(fin, rfin)
        np.testing.assert_array_equal(ctr, rctr)
        np.testing.assert_array_equal(acc, racc)
        np.testing.assert_array_equal(np.ascontiguousarray(gt).view(np.uint64),
                                      np.ascontiguousarray(rt).view(np.uint64))
        if kind == 1:  # exchange moves conserve the magnetisation of every chain
            sz0 = s0[q].astype(int).sum(axis=1)
            assert (gs.astype(int).sum(axis=1).reshape(3, -1) == sz0[None, :]).all()


@pytest.mark.parametrize("prec", [0, 1])
def test_multi_eloc_matches_replay(ctx, orc, prec):
    rng = np.random.default_rng(5)
    ms, rbms = _machines(orc, SHAPES, 24, rng, scale=0.05)
    ctx.multi_create(ms, sweeper=0, precision=prec)
    for q, rbm in enumerate(rbms):
        ctx.multi_set_params(q, rbm.params())
    ctx.multi_sample(2, 5)
    mean, var = ctx.multi_eloc()
    for q, (mc, rbm) in enumerate(zip(ms, rbms)):
        gs, gt = ctx.multi_samples(q)
        kind, hk = mc["ham"]
        el = orc.eloc_replay(orc.ham(kind, mc["n"], **hk), rbm, gs, gt, precision=prec)
        assert abs(mean[q] - el.mean()) <= 1e-13 * max(1.0, abs(el.mean()))
        assert abs(var[q] - el.var()) <= 1e-9 * max(1.0, el.var())


def test_multi_steps_match_single_contexts(ctx, yq, orc):
    """Epoch bodies of three machines run together equal, bit for bit, the
    same machines run in their own contexts."""
    rng = np.random.default_rng(9)
    shapes = [(20, 40), (48, 96), (32, 32)]
    ms, rbms = _machines(orc, shapes, 256, rng, scale=0.01)
    ctx.multi_create(ms, sweeper=0, precision=1)
    for q, rbm in enumerate(rbms):
        ctx.multi_set_params(q, rbm.params())
    kw = dict(n_sweeps=1, solver=2, precision=1, lam=1e-3, cg_tol=1e-8, lr=0.02)
    multi = [ctx.multi_sr_step(n_discard=10 if i == 0 else 0, **kw) for i in range(3)]
    for q, (mc, rbm) in enumerate(zip(ms, rbms)):
        with yq.Context(0) as c1:
            c1.rbm_create(mc["n"], mc["m"])
            c1.set_params(rbm.params())
            kind, hk = mc["ham"]
            c1.ham_set(kind, mc["n"], **hk)
            c1.sampler_config(0, mc["chains"], precision=1, seed=mc["seed"])
            for i in range(3):
                o = c1.sr_step(n_discard=10 if i == 0 else 0, **kw)
                assert o["energy"] == multi[i][q]["energy"]
                assert o["cg_iters"] == multi[i][q]["cg_iters"]
                assert o["update_norm"] == multi[i][q]["update_norm"]
            np.testing.assert_array_equal(c1.get_params(), ctx.multi_get_params(q))


def test_c5_config(ctx):
    """C5: 32 TFI machines, N ~ U{16, 24, ..., 128}, alpha ~ U{1, 2, 4}, h
    spaced over [0.5, 1.5], 2048 samples each, sampling in float and the
    mixed factored SR; prints the step time."""
    rng = np.random.default_rng(1234)
    Ns = rng.choice(np.arange(16, 129, 8), 32)
    alphas = rng.choice([1, 2, 4], 32)
    hs = np.linspace(0.5, 1.5, 32)
    ms = [dict(n=int(n), m=int(a * n), chains=2048, seed=77 + q,
               ham=(0, dict(J=-1.0, h=-float(h))))
          for q, (n, a, h) in enumerate(zip(Ns, alphas, hs))]
    ctx.multi_create(ms, sweeper=0, precision=1)
    ctx.multi_init_random(11, 0.01)
    kw = dict(n_sweeps=1, solver=2, precision=1, lam=1e-3, cg_tol=1e-8, lr=0.01)
    ctx.multi_sr_step(n_discard=20, **kw)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        outs = ctx.multi_sr_step(**kw)
        times.append(time.perf_counter() - t0)
    for mc, o, h in zip(ms, outs, hs):
        assert np.isfinite(o["energy"]) and o["cg_err"] < 1e-8 and o["n_samples"] == 2048
        # E/N of the TFI chain lies between the classical -1 and the paramagnet -h (J = -1)
        assert -1.7 < o["energy"] / mc["n"] < -0.4
    dims = sum(mc["n"] * mc["m"] + mc["n"] + mc["m"] for mc in ms)
    print(f"\nC5: 32 machines, {dims} parameters in total, step {1e3 * min(times):.1f} ms "
          f"(sampling {outs[0]['sample_ms']:.2f} ms for all machines in one launch), "
          f"CG iterations {sorted(o['cg_iters'] for o in outs)}")


@pytest.mark.parametrize("case", ["default", "jacobi", "autotune", "eager", "budget"])
def test_grouped_solve_matches_per_machine(yq, orc, case):
    """The grouped SR solve (one forward, backward and CG-step launch per
    iteration for all machines, each block running its machine's own tile
    shapes) gives every machine the same solution, bit for bit, as the
    per-machine solves: n from 16 to 128 (1 to 5 k-steps of 32), uneven chain counts,
    with the Jacobi preconditioner, with autotuned tiles (other column-group
    widths and chunks), without graph capture, and with an iteration budget of
    20 (unconverged solves, budgets that run out inside a refinement round)."""
    rng = np.random.default_rng(21)
    shapes = [(120, 240), (16, 16), (100, 400), (37, 70), (64, 128), (128, 128)]
    ms, rbms = _machines(orc, shapes, [512, 256, 1024, 384, 512, 768], rng, scale=0.02)
    kw = dict(n_sweeps=1, solver=2, precision=1, lam=1e-3, cg_tol=1e-8, lr=0.02,
              precond=1 if case == "jacobi" else 0)
    if case == "budget":
        kw.update(cg_max_iter=20, allow_unconverged=True)
    res = []
    for grouped in (True, False):
        with yq.Context(0) as c:
            c.multi_create(ms, sweeper=0, precision=1)
            for q, rbm in enumerate(rbms):
                c.multi_set_params(q, rbm.params())
            c.multi_set_grouped(grouped)
            if case == "autotune":
                c.autotune(True)
            if case == "eager":
                c.graph_mode(0)
            outs = [c.multi_sr_step(n_discard=10 if i == 0 else 0, **kw) for i in range(2)]
            res.append((outs, [c.multi_get_params(q) for q in range(len(ms))]))
    (og, pg), (op, pp) = res
    for i in range(2):
        for q in range(len(ms)):
            for key in ("energy", "cg_iters", "cg_err", "update_norm"):
                assert og[i][q][key] == op[i][q][key], (i, q, key)
            if case == "budget":
                assert og[i][q]["cg_iters"] <= 20
            else:
                assert og[i][q]["cg_err"] < 1e-8
    for q in range(len(ms)):
        np.testing.assert_array_equal(pg[q], pp[q])
