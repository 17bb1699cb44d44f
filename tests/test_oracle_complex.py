"""Pins of the complex-RBM oracle (oracle/oracle_complex.cpp) and of the
complex math the kernels use (yannq_amd/csrc/yq_math.h): numpy's complex
functions, brute force over the 2^N basis, exact diagonalisation, finite
differences of the exact energy and of the normalised state.  Reference
test techniques mirrored: Tests/TestUtility.cpp:13-32 (complex logCosh on
random points of [-10, 10]^2), Tests/TestRBMStates.cpp:94-236 (log-ratios,
theta cache), Tests/TestRBM.cpp:184-231 (log-derivatives by finite
differences)."""
import math

import numpy as np
import pytest


def _wrap(phi):
    """Angle difference wrapped into (-pi, pi]."""
    return (phi + np.pi) % (2 * np.pi) - np.pi


def _close_log(a, b, tol):
    """Complex logarithms agree modulo 2 pi i."""
    return abs(a.real - b.real) <= tol and abs(_wrap(a.imag - b.imag)) <= tol


def test_complex_logcosh(orc):
    L = orc.lib()
    rng = np.random.default_rng(0)
    z = rng.uniform(-10, 10, 20000) + 1j * rng.uniform(-10, 10, 20000)
    ref = np.log(np.cosh(z))
    out = np.empty(2)
    outf = np.empty(2, np.float32)
    for q in range(len(z)):
        x, y = z[q].real, z[q].imag
        got = orc.c_logcosh(z[q])
        assert _close_log(got, ref[q], 1e-7)          # TestUtility.cpp tolerance
        assert _close_log(got, ref[q], 1e-11)
        L.rp_logcosh_c_d(x, y, orc._p(out))
        assert abs(out[0] - ref[q].real) < 1e-12 and abs(_wrap(out[1] - ref[q].imag)) < 1e-12
        # the replayed real part is the same operations as rp_logcosh_c_d's
        assert L.rp_logcosh_re_d(x, y) == out[0]
        xf, yf = np.float32(x), np.float32(y)
        reff = np.log(np.cosh(complex(float(xf), float(yf))))
        L.rp_logcosh_c_f(xf, yf, orc._p(outf))
        # float: a few ulp, also next to the zeros of cosh (no cancellation)
        assert abs(float(outf[0]) - reff.real) < 6e-7 * max(1.0, abs(reff.real))
        assert abs(_wrap(float(outf[1]) - reff.imag)) < 1e-6
        assert L.rp_logcosh_re_f(xf, yf) == outf[0]
    # next to the zeros of cosh (x ~ 0, y ~ pi/2): |log cosh| grows but stays accurate
    zz = 1e-3 * rng.normal(size=2000) + 1j * (np.pi / 2 + 1e-3 * rng.normal(size=2000))
    for q in zz:
        ref = np.log(np.cosh(q))
        L.rp_logcosh_c_d(q.real, q.imag, orc._p(out))
        assert abs(out[0] - ref.real) < 4e-16 * max(1.0, abs(ref.real))
        xf, yf = np.float32(q.real), np.float32(q.imag)
        reff = np.log(np.cosh(complex(float(xf), float(yf))))
        L.rp_logcosh_c_f(xf, yf, orc._p(outf))
        assert abs(float(outf[0]) - reff.real) < 6e-7 * max(1.0, abs(reff.real))
    # the real branch (|x| > 12) of Utility.hpp:18 carries into the complex form
    got = orc.c_logcosh(20.0 + 0.3j)
    assert abs(got.real - (20.0 - math.log(2.0) + 0.5 * math.log(math.cos(0.3) ** 2 +
                                                               math.sin(0.3) ** 2))) < 1e-12


def test_complex_tanh(orc):
    L = orc.lib()
    rng = np.random.default_rng(1)
    z = np.concatenate([rng.uniform(-10, 10, 5000) + 1j * rng.uniform(-10, 10, 5000),
                        rng.uniform(-1e-6, 1e-6, 200) + 1j * rng.uniform(-3, 3, 200),
                        [0.0, 1e-300 + 0.1j, -25.0 + 3.0j, 0.4 + 0j]])
    out = np.empty(2)
    for q in range(len(z)):
        L.rp_tanh_c_d(z[q].real, z[q].imag, orc._p(out))
        t = np.tanh(z[q])
        assert abs(out[0] + 1j * out[1] - t) <= 1e-14 * max(1.0, abs(t)), (z[q], out, t)


def test_complex_theta(orc):
    rng = np.random.default_rng(2)
    rbm = orc.CRBM.random(11, 23, 0.3, rng)
    sig = rng.choice([-1, 1], size=(5, 11)).astype(np.int8)
    np.testing.assert_allclose(orc.c_theta(rbm, sig), sig @ rbm.W + rbm.b, rtol=0, atol=1e-14)


def _c_psi_table(orc, rbm):
    """log psi over the basis from numpy's complex cosh (independent of the
    oracle's logCosh): log psi = a.sigma + sum_j log cosh theta_j."""
    basis = orc.basis_sigma(rbm.n)
    th = basis @ rbm.W + rbm.b
    lp = basis @ rbm.a + np.log(np.cosh(th)).sum(axis=1)
    return basis, th, lp


def test_complex_log_ratio_vector_brute_force(orc):
    """Complex logRatio(vector) against the brute-force log psi table (mod
    2 pi i), lists of 1 to n distinct sites."""
    rng = np.random.default_rng(13)
    n, m = 7, 9
    rbm = orc.CRBM.random(n, m, 0.4, rng)
    basis, th, lp = _c_psi_table(orc, rbm)
    index = {tuple(s): q for q, s in enumerate(basis)}
    for q in rng.choice(len(basis), 25, replace=False):
        s = basis[q]
        for k in (1, 2, 4, n):
            sites = rng.choice(n, k, replace=False)
            s2 = s.copy()
            s2[sites] = -s2[sites]
            want = lp[index[tuple(s2)]] - lp[q]
            assert _close_log(orc.c_log_ratio_vec(rbm, s, th[q], sites), want, 1e-11)


def test_complex_log_coeff_and_ratios_brute_force(orc):
    rng = np.random.default_rng(3)
    n, m = 7, 9
    rbm = orc.CRBM.random(n, m, 0.4, rng)
    basis, th, lp = _c_psi_table(orc, rbm)
    index = {tuple(s): q for q, s in enumerate(basis)}
    for q in rng.choice(len(basis), 25, replace=False):
        s = basis[q]
        assert _close_log(orc.c_log_coeff(rbm, s, th[q]), lp[q], 1e-12)
        for k in range(n):
            s2 = s.copy()
            s2[k] = -s2[k]
            want = lp[index[tuple(s2)]] - lp[q]
            assert _close_log(orc.c_log_ratio1(rbm, s, th[q], k), want, 1e-11)
            l = (k + 2) % n
            s3 = s2.copy()
            s3[l] = -s3[l]
            want2 = lp[index[tuple(s3)]] - lp[q]
            assert _close_log(orc.c_log_ratio2(rbm, s, th[q], k, l), want2, 1e-11)


@pytest.mark.parametrize("name", ["tfi", "xxz", "xxz_sign", "j1j2"])
def test_complex_eloc_brute_force(orc, name):
    """E_loc(sigma) = sum_s' H_{sigma s'} psi(s') / psi(sigma) with H from the
    ED column functors (pinned against the known energies elsewhere)."""
    n, m = 8, 6
    H = {"tfi": orc.ham(orc.TFI, n, J=-1.0, h=-0.7),
         "xxz": orc.ham(orc.XXZ, n, J=1.0, delta=0.6),
         "xxz_sign": orc.ham(orc.XXZ, n, J=1.0, delta=1.0, sign_rule=1),
         "j1j2": orc.ham(orc.J1J2, n, J=1.0, J2=0.4, sign_rule=1)}[name]
    rng = np.random.default_rng(4)
    rbm = orc.CRBM.random(n, m, 0.3, rng)
    basis, th, lp = _c_psi_table(orc, rbm)
    psi = np.exp(lp - lp.real.max())
    Hm = orc.ham_matrix(H).toarray()
    want = (Hm @ psi) / psi
    sel = rng.choice(len(basis), 40, replace=False)
    got = orc.c_eloc(H, rbm, basis[sel], orc.c_theta(rbm, basis[sel]))
    np.testing.assert_allclose(got, want[sel], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("kind,kw", [(0, dict(J=-1.0, h=-0.8)), (1, dict(J=1.0, delta=1.0)),
                                     (2, dict(J=1.0, J2=0.5, sign_rule=1))])
def test_complex_eloc_replay(orc, kind, kw):
    """The kernels' complex E_loc arithmetic (rp_c_eloc) against the libm
    restatement: 1e-10 in FP64; in float (theta rounded to complex64) the
    bound of a float pipeline, which numpy's own complex64 log(cosh) meets
    too (max error 4e-5 of max(|E|, 1), 2e-6 norm-wise on these inputs)."""
    n, m, ns = 24, 48, 300
    rbm = orc.CRBM.random(n, m, 0.2, np.random.default_rng(21))
    sig = np.random.default_rng(kind + 10).choice([-1, 1], size=(ns, n)).astype(np.int8)
    H = orc.ham(kind, n, **kw)
    th = orc.c_theta(rbm, sig)
    want = orc.c_eloc(H, rbm, sig, th)
    got = orc.c_eloc_replay(H, rbm, sig, th)
    assert np.linalg.norm(got - want) < 1e-10 * np.linalg.norm(want)
    th32 = th.astype(np.complex64).astype(np.complex128)
    got32 = orc.c_eloc_replay(H, rbm, sig, th32, precision=orc.MIXED)
    err = np.abs(got32 - want)
    assert (err / np.maximum(np.abs(want), 1.0)).max() < 1e-4
    assert np.linalg.norm(got32 - want) < 1e-5 * np.linalg.norm(want)


def test_complex_logderiv_finite_difference(orc):
    """O = d log psi / d p is holomorphic: a real step gives O, an imaginary
    step gives i O (central differences, steps 1e-6)."""
    rng = np.random.default_rng(5)
    n, m = 6, 5
    rbm = orc.CRBM.random(n, m, 0.3, rng)
    s = rng.choice([-1, 1], size=n).astype(np.int8)
    O = orc.c_logderiv(rbm, s, orc.c_theta(rbm, s)[0])
    p0 = rbm.params()
    h = 1e-6
    for q in range(len(p0)):
        for dirn in (1.0, 1j):
            vals = []
            for sg in (1, -1):
                p = p0.copy()
                p[q] += sg * h * dirn
                rbm.set_params(p)
                vals.append(orc.c_log_coeff(rbm, s))
            rbm.set_params(p0)
            d = vals[0] - vals[1]
            d = complex(d.real, _wrap(d.imag)) / (2 * h)
            assert abs(d - dirn * O[q]) < 1e-7, (q, dirn, d, O[q])


def _exact_c_sr(orc, rbm, H):
    basis, th, lp = _c_psi_table(orc, rbm)
    w = np.exp(2 * (lp.real - lp.real.max()))
    w /= w.sum()
    el = orc.c_eloc(H, rbm, basis, th)
    O, om, F, S = orc.c_sr_build(rbm, basis, th, el, weights=w)
    return basis, w, el, F, S


def test_complex_force_is_energy_gradient(orc):
    """With F = <O_c^* E_loc> (SRMat.hpp:98-102 with the adjoint), the exact
    energy satisfies dE/dRe p = 2 Re F and dE/dIm p = 2 Im F."""
    rng = np.random.default_rng(6)
    n = 7
    rbm = orc.CRBM.random(n, n, 0.2, rng)
    H = orc.ham(orc.XXZ, n, J=1.0, delta=1.0)
    Hm = orc.ham_matrix(H)

    def energy(r):
        _, _, lp = _c_psi_table(orc, r)
        psi = np.exp(lp - lp.real.max())
        return (np.vdot(psi, Hm @ psi) / np.vdot(psi, psi)).real

    _, w, el, F, S = _exact_c_sr(orc, rbm, H)
    assert abs(w @ el.real - energy(rbm)) < 1e-11 and abs(w @ el.imag) < 1e-11
    p0 = rbm.params()
    h = 1e-5
    for q in rng.choice(len(p0), 10, replace=False):
        for dirn, part in ((1.0, F[q].real), (1j, F[q].imag)):
            vals = []
            for sg in (1, -1):
                p = p0.copy()
                p[q] += sg * h * dirn
                rbm.set_params(p)
                vals.append(energy(rbm))
            rbm.set_params(p0)
            assert abs((vals[0] - vals[1]) / (2 * h) - 2 * part) < 1e-7


def test_complex_S_is_quantum_geometric_tensor(orc):
    """v^H S v = |d_v psi|^2/|psi|^2 - |<psi, d_v psi>|^2/|psi|^4 with d_v psi
    a central difference of the unnormalised state vector."""
    rng = np.random.default_rng(7)
    n = 7
    rbm = orc.CRBM.random(n, n, 0.2, rng)
    H = orc.ham(orc.XXZ, n, J=1.0, delta=1.0)
    _, w, el, F, S = _exact_c_sr(orc, rbm, H)
    assert np.allclose(S, S.conj().T, atol=1e-14)
    assert np.linalg.eigvalsh(S).min() > -1e-12
    p0 = rbm.params()
    _, _, lp0 = _c_psi_table(orc, rbm)
    shift = lp0.real.max()

    def psi_of(p):
        rbm.set_params(p)
        _, _, lp = _c_psi_table(orc, rbm)
        return np.exp(lp - shift)

    psi0 = psi_of(p0)
    h = 1e-5
    for _ in range(4):
        v = rng.normal(size=len(p0)) + 1j * rng.normal(size=len(p0))
        dpsi = (psi_of(p0 + h * v) - psi_of(p0 - h * v)) / (2 * h)
        nn = np.vdot(psi0, psi0).real
        want = np.vdot(dpsi, dpsi).real / nn - abs(np.vdot(psi0, dpsi)) ** 2 / nn ** 2
        got = np.vdot(v, S @ v)
        assert abs(got.imag) < 1e-12 * abs(got)
        assert abs(got.real - want) < 1e-7 * want
    rbm.set_params(p0)


def test_complex_sr_build_uniform_weights(orc):
    rng = np.random.default_rng(8)
    n, m, ns = 6, 8, 30
    rbm = orc.CRBM.random(n, m, 0.3, rng)
    H = orc.ham(orc.XXZ, n, J=1.0, delta=1.0)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    th = orc.c_theta(rbm, sig)
    el = orc.c_eloc(H, rbm, sig, th)
    O, om, F, S = orc.c_sr_build(rbm, sig, th, el)
    O2, om2, F2, S2 = orc.c_sr_build(rbm, sig, th, el, weights=np.full(ns, 1.0 / ns))
    np.testing.assert_allclose(F, F2, atol=1e-14)
    np.testing.assert_allclose(S, S2, atol=1e-14)
    raw = np.array([orc.c_logderiv(rbm, sig[s], th[s]) for s in range(ns)])
    np.testing.assert_allclose(raw.mean(axis=0), om, atol=1e-15)
    np.testing.assert_allclose(O.mean(axis=0), 0, atol=1e-15)
    assert np.linalg.matrix_rank(S) <= ns - 1


def _python_c_sweep(orc, rbm, sig, seed, stream, ctr, kind, beta=1.0):
    """One sweep with the libm restatement (Re of or_c_log_ratio*) and the
    Philox draws (LocalSweeper.hpp:32-53 / SwapSweeper.hpp:28-56 acceptance
    exp(2 beta Re dlogpsi))."""
    n = rbm.n
    sig = sig.copy()
    for _ in range(n):
        x = orc.philox((ctr & 0xffffffff, ctr >> 32, stream, 0), (seed & 0xffffffff, seed >> 32))
        ctr += 1
        k = (int(x[0]) * n) >> 32
        th = orc.c_theta(rbm, sig)[0]
        if kind == orc.SWAP:
            l = (int(x[1]) * n) >> 32
            if sig[k] == sig[l]:
                continue
            d = orc.c_log_ratio2(rbm, sig, th, k, l).real
        else:
            d = orc.c_log_ratio1(rbm, sig, th, k).real
        u = (((int(x[2]) >> 5) << 26) | (int(x[3]) >> 6)) / 2.0 ** 53
        if u < min(1.0, math.exp(2 * beta * d)):
            sig[k] = -sig[k]
            if kind == orc.SWAP:
                sig[l] = -sig[l]
    return sig


@pytest.mark.parametrize("kind", [0, 1])
def test_complex_sampler_replay_matches_restatement(orc, kind):
    rng = np.random.default_rng(9)
    n, m, Cn = 10, 20, 6
    rbm = orc.CRBM.random(n, m, 0.4, rng)
    sig0 = orc.randomize(5, 0, np.arange(Cn), n, nup=n // 2 if kind else -1)
    seed = 0x2468_ace0_1357
    fin, ctr, acc, ss, st = orc.c_sample_replay(rbm, sig0, seed=seed, n_sweeps=3, n_discard=0,
                                                kind=kind)
    assert (ctr == 3 * n).all() and (acc > 0).all()
    for c in range(Cn):
        s = sig0[c]
        for sw in range(3):
            s = _python_c_sweep(orc, rbm, s, seed, c, sw * n, kind)
            np.testing.assert_array_equal(ss[sw * Cn + c], s)
    np.testing.assert_array_equal(ss[2 * Cn:], fin)
    np.testing.assert_allclose(st, orc.c_theta(rbm, ss), atol=1e-13)
    if kind == orc.SWAP:
        assert (ss.astype(int).sum(axis=1) == 0).all()
    # mixed: float theta refreshed from double every sweep
    _, _, _, ss2, st2 = orc.c_sample_replay(rbm, sig0, seed=seed, n_sweeps=4, n_discard=0,
                                            kind=kind, precision=orc.MIXED)
    np.testing.assert_allclose(st2, orc.c_theta(rbm, ss2), atol=2e-6)


def test_complex_sampler_distribution_matches_psi2(orc):
    """Chi-square of exchange-move samples against |psi|^2 restricted to the
    Sz = 0 sector (N=6: 20 configurations), over the bins expected to hold at
    least 20 samples (correlated samples inflate the rare bins), and the total
    variation distance."""
    rng = np.random.default_rng(11)
    n = 6
    rbm = orc.CRBM.random(n, 2 * n, 0.25, rng)
    basis, th, lp = _c_psi_table(orc, rbm)
    sector = basis.astype(int).sum(axis=1) == 0
    p = np.exp(2 * (lp.real - lp.real.max())) * sector
    p /= p.sum()
    Cn = 64
    sig0 = orc.randomize(3, 0, np.arange(Cn), n, nup=n // 2)
    _, _, _, ss, _ = orc.c_sample_replay(rbm, sig0, seed=91, n_sweeps=300, n_discard=20,
                                         kind=orc.SWAP)
    idx = (((1 - ss.astype(np.int64)) // 2) << np.arange(n)).sum(axis=1)
    cnt = np.bincount(idx, minlength=1 << n)
    assert cnt[~sector].sum() == 0
    N = cnt.sum()
    big = sector & (N * p >= 20)
    chi2 = ((cnt[big] - N * p[big]) ** 2 / (N * p[big])).sum()
    assert big.sum() >= 12 and chi2 < 3 * (big.sum() - 1), chi2
    assert 0.5 * np.abs(cnt / N - p).sum() < 0.02


@pytest.mark.parametrize("kind", [0, 1])
def test_complex_tempering_replay_matches_restatement(orc, kind):
    """Complex parallel tempering (SamplerMT.hpp:120-153): sweeps at
    beta_t = (T-t)/T with the libm restatement, then the two swap passes with
    p = exp((beta_{t+1} - beta_t) 2 Re(log psi_t - log psi_{t+1})) from numpy's
    complex log cosh and the lower chain's mix-domain Philox draw."""
    rng = np.random.default_rng(21)
    n, m, T, G = 10, 20, 4, 2
    rbm = orc.CRBM.random(n, m, 0.4, rng)
    Cn = T * G
    sig0 = orc.randomize(8, 0, np.arange(Cn), n, nup=n // 2 if kind else -1)
    seed = 0x1357_2468_ace0
    sweeps = 4

    def lpsi(s):
        th = s.astype(float) @ rbm.W + rbm.b
        return (s.astype(float) @ rbm.a + np.log(np.cosh(th)).sum()).real

    fin, ctr, acc, ss, st = orc.c_sample_replay_pt(rbm, sig0, n_temps=T, seed=seed,
                                                   n_sweeps=sweeps, n_discard=0, kind=kind)
    assert (ctr == sweeps * n).all()
    swaps = 0
    for g in range(G):
        state = [sig0[g * T + t].copy() for t in range(T)]
        for sw in range(sweeps):
            for t in range(T):
                state[t] = _python_c_sweep(orc, rbm, state[t], seed, g * T + t, sw * n, kind,
                                           beta=(T - t) / T)
            lp = [lpsi(state[t]) for t in range(T)]
            for pas in (0, 1):
                for t in range(pas, T - 1 - pas, 2):
                    c = g * T + t
                    ev = (sw + 1) * n
                    x = orc.philox((ev & 0xffffffff, ev >> 32, c, 5),
                                   (seed & 0xffffffff, seed >> 32))
                    u = (((int(x[2]) >> 5) << 26) | (int(x[3]) >> 6)) / 2.0 ** 53
                    dbeta = (T - t - 1) / T - (T - t) / T
                    if u < min(1.0, math.exp(dbeta * 2 * (lp[t] - lp[t + 1]))):
                        state[t], state[t + 1] = state[t + 1], state[t]
                        lp[t], lp[t + 1] = lp[t + 1], lp[t]
                        swaps += 1
            np.testing.assert_array_equal(ss[sw * G + g], state[0])
        for t in range(T):
            np.testing.assert_array_equal(fin[g * T + t], state[t])
    assert swaps > 0
    np.testing.assert_allclose(st, orc.c_theta(rbm, ss), atol=1e-13)
    _, _, _, ss2, st2 = orc.c_sample_replay_pt(rbm, sig0, n_temps=T, seed=seed, n_sweeps=4,
                                               n_discard=0, kind=kind, precision=orc.MIXED)
    np.testing.assert_allclose(st2, orc.c_theta(rbm, ss2), atol=2e-6)


@pytest.mark.parametrize("prec", [0, 1])
def test_complex_tempering_real_parameters_equal_real_replay(orc, prec):
    """With Im W = Im a = Im b = 0 the complex tempering replay makes the real
    tempering replay's decisions: Re logcosh of a real argument is logcosh."""
    rng = np.random.default_rng(22)
    n, m, T, G = 12, 24, 4, 3
    r = orc.RBM.random(n, m, 0.5, rng)
    c = orc.CRBM(n, m, r.W.astype(complex), r.a.astype(complex), r.b.astype(complex))
    sig0 = orc.randomize(9, 0, np.arange(T * G), n)
    kw = dict(n_temps=T, seed=77, n_sweeps=5, n_discard=2, precision=prec)
    a = orc.sample_replay_pt(r, sig0, **kw)
    b = orc.c_sample_replay_pt(c, sig0, **kw)
    for u, v in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(b[4].imag, 0)
    np.testing.assert_allclose(b[4].real, a[4], rtol=0, atol=2e-7 if prec else 1e-15)


def test_complex_tempering_distribution_matches_psi2(orc):
    """The beta = 1 slot of tempered exchange-move chains samples |psi|^2 on
    the Sz = 0 sector (N = 6), with swaps between the temperatures."""
    rng = np.random.default_rng(23)
    n = 6
    rbm = orc.CRBM.random(n, 2 * n, 0.15, rng)
    basis, th, lp = _c_psi_table(orc, rbm)
    sector = basis.astype(int).sum(axis=1) == 0
    p = np.exp(2 * (lp.real - lp.real.max())) * sector
    p /= p.sum()
    T, G = 4, 48
    sig0 = orc.randomize(4, 0, np.arange(T * G), n, nup=n // 2)
    _, _, _, ss, _ = orc.c_sample_replay_pt(rbm, sig0, n_temps=T, seed=93, n_sweeps=400,
                                            n_discard=20, kind=orc.SWAP)
    idx = (((1 - ss.astype(np.int64)) // 2) << np.arange(n)).sum(axis=1)
    cnt = np.bincount(idx, minlength=1 << n)
    assert cnt[~sector].sum() == 0
    N = cnt.sum()
    big = sector & (N * p >= 20)
    chi2 = ((cnt[big] - N * p[big]) ** 2 / (N * p[big])).sum()
    assert big.sum() >= 12 and chi2 < 3 * (big.sum() - 1), chi2
    assert 0.5 * np.abs(cnt / N - p).sum() < 0.03
