"""The threaded CPU baseline (bl_*, what bench.py times as cpu_baseline) is
the oracle run on several threads: on small batches its local energies equal
the replay rp_eloc bit for bit and the restatement or_eloc to 1e-12, its O
rows, column mean and F equal or_sr_build, its CG solution and iteration
count equal or_cg_solve (all pinned in test_oracle.py against brute force,
finite differences and the Fubini-Study metric), and its sampler equals the
kernel replay rp_sample bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle as orc


@pytest.mark.parametrize("kind,n,m", [(orc.TFI, 10, 20), (orc.XXZ, 12, 12), (orc.J1J2, 10, 30)])
def test_bl_eloc_equals_or_eloc(kind, n, m):
    rng = np.random.default_rng(3 + n)
    rbm = orc.RBM.random(n, m, 0.1, rng)
    H = orc.ham(kind, n, J=-1.0 if kind == orc.TFI else 1.0, h=-0.7, delta=0.6, J2=0.4)
    ns = 97
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    th = orc.theta(rbm, sig)
    ref = orc.eloc(H, rbm, sig, th)
    rep = orc.eloc_replay(H, rbm, sig, th)
    out = np.empty(ns)
    orc.lib().bl_eloc(C.byref(H), ns, n, m, orc._p(rbm.W), orc._p(rbm.a), orc._p(sig),
                      orc._p(th), orc._p(out), 4)
    np.testing.assert_array_equal(out, rep)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_bl_sr_build_and_cg_equal_oracle():
    n, m, ns = 8, 16, 301
    rng = np.random.default_rng(5)
    rbm = orc.RBM.random(n, m, 0.1, rng)
    H = orc.ham(orc.TFI, n)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    th = orc.theta(rbm, sig)
    el = orc.eloc(H, rbm, sig, th)
    O, om, F, _ = orc.sr_build(rbm, sig, th, el, with_S=False)
    dim = rbm.dim
    Ob, omb, Fb = np.empty((ns, dim)), np.empty(dim), np.empty(dim)
    orc.lib().bl_sr_build(ns, n, m, 1, orc._p(sig), orc._p(th), orc._p(el), orc._p(Ob),
                          orc._p(omb), orc._p(Fb), 3)
    # same terms in the same order: the restatement sums samples in index order
    np.testing.assert_allclose(Ob, O, rtol=0, atol=1e-15)
    np.testing.assert_allclose(omb, om, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(Fb, F, rtol=1e-12, atol=1e-15)
    x_ref, it_ref, _, rc_ref = orc.cg_solve(O, 1e-3, F, 1e-10, 10000)
    assert rc_ref == 0
    x = np.empty(dim)
    it, err = C.c_int(), C.c_double()
    rc = orc.lib().bl_cg_solve(ns, dim, orc._p(Ob), 1e-3, orc._p(Fb), 1e-10, 10000, orc._p(x),
                               C.byref(it), C.byref(err), 4)
    assert rc == 0 and err.value < 1e-10
    assert abs(it.value - it_ref) <= 1
    # both stop at |r| < 1e-10 |b|; the solutions agree to the conditioning
    Ax = O.T @ (O @ x) / ns + 1e-3 * x
    assert np.linalg.norm(Ax - F) < 1.01e-10 * np.linalg.norm(F)
    assert np.linalg.norm(x - x_ref) < 1e-6 * np.linalg.norm(x_ref)


def test_bl_sample_equals_replay():
    n, m, c = 12, 24, 40
    rng = np.random.default_rng(8)
    rbm = orc.RBM.random(n, m, 0.2, rng)
    sig0 = orc.randomize(7, 0, np.arange(c), n)
    a = orc.sample_replay(rbm, sig0, seed=7, n_sweeps=3, n_discard=2, threads=1)
    b = orc.sample_replay(rbm, sig0, seed=7, n_sweeps=3, n_discard=2, threads=5)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
