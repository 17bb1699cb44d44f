"""Pins of the CPU oracle (oracle/oracle.cpp) against known answers, brute
force over the 2^N basis, finite differences, exact diagonalisation and
closed forms.  Reference test techniques mirrored here:
Tests/TestUtility.cpp:13-32 (logCosh), Tests/TestRBMStates.cpp:94-236
(brute-force log-ratios, theta-cache invariant), Tests/TestRBM.cpp:184-231
(finite-difference log-derivatives), Tests/TestHamTFI.cpp:7-20 and
Tests/HamiltonianHelper.hpp (E_loc functor vs ED column functor)."""
import math

import numpy as np
import pytest


def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32-10
    assert list(orc.philox((0, 0, 0, 0), (0, 0))) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(orc.philox((0xffffffff,) * 4, (0xffffffff,) * 2)) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
                           (0xa4093822, 0x299f31d0))) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_logcosh_restatement_and_replay(orc):
    L = orc.lib()
    xs = np.concatenate([np.linspace(-12, 12, 20001), [0.0, 1e-8, -3e-5, 11.999999, 12.0]])
    for x in xs:
        ref = math.log(math.cosh(abs(x)))
        assert L.or_logcosh(x) == ref
        # replayable form: |x| - ln2 + log1p(e^{-2|x|}); absolute error ~1 ulp of |x|+ln2
        assert abs(L.rp_logcosh_d(x) - ref) <= 4e-16 * (1 + abs(x))
        assert abs(L.rp_logcosh_f(np.float32(x)) - math.log(math.cosh(float(np.float32(x))))) \
            <= 3e-7 * (1 + abs(x))
    # branch point of Utility.hpp:18: beyond 12 both return |x| - ln 2
    for x in (12.5, -20.0, 700.0):
        assert L.or_logcosh(x) == abs(x) - math.log(2.0)
        assert L.rp_logcosh_d(x) == abs(x) - math.log(2.0)
    assert math.isnan(L.rp_logcosh_d(float("nan")))


def test_exp_replay(orc):
    L = orc.lib()
    for x in np.linspace(-707, 709, 50001):
        assert abs(L.rp_exp_d(x) - math.exp(x)) <= 3e-16 * math.exp(x)
    for x in np.linspace(-86, 88, 20001, dtype=np.float32):
        assert abs(L.rp_exp_f(x) - math.exp(float(x))) <= 2.5e-7 * math.exp(float(x))
    assert L.rp_exp_d(-800.0) == 0.0 and L.rp_exp_d(800.0) == math.inf
    assert L.rp_exp_d(0.0) == 1.0


def test_theta(orc):
    rng = np.random.default_rng(1)
    rbm = orc.RBM.random(13, 29, 0.3, rng)
    sig = rng.choice([-1, 1], size=(5, 13)).astype(np.int8)
    want = sig @ rbm.W + rbm.b
    np.testing.assert_allclose(orc.theta(rbm, sig), want, rtol=0, atol=1e-14)
    np.testing.assert_allclose(orc.theta_replay(rbm, sig), want, rtol=0, atol=1e-14)


def _psi_table(orc, rbm):
    basis = orc.basis_sigma(rbm.n)
    th = orc.theta(rbm, basis)
    return basis, th, np.array([orc.log_coeff(rbm, basis[s], th[s]) for s in range(len(basis))])


def test_log_ratios_brute_force(orc):
    rng = np.random.default_rng(2)
    n = 10
    rbm = orc.RBM.random(n, n, 1.0, rng)  # TestRBMStates.cpp:104-107 uses sigma=1
    basis, th, logpsi = _psi_table(orc, rbm)
    for _ in range(20):
        v = int(rng.integers(0, 1 << n))
        for _ in range(3):
            k, l = rng.choice(n, 2, replace=False)
            want1 = logpsi[v ^ (1 << k)] - logpsi[v]
            want2 = logpsi[v ^ (1 << k) ^ (1 << l)] - logpsi[v]
            assert abs(orc.log_ratio1(rbm, basis[v], th[v], int(k)) - want1) < 1e-11
            assert abs(orc.log_ratio2(rbm, basis[v], th[v], int(k), int(l)) - want2) < 1e-11


def test_logderiv_finite_difference(orc):
    rng = np.random.default_rng(3)
    n, m = 7, 11
    rbm = orc.RBM.random(n, m, 0.5, rng)
    sig = rng.choice([-1, 1], n).astype(np.int8)
    O = orc.logderiv(rbm, sig, orc.theta(rbm, sig)[0])
    p0 = rbm.params()
    h = 1e-6
    fd = np.empty_like(p0)
    for q in range(len(p0)):
        vals = []
        for sgn in (1, -1):
            p = p0.copy()
            p[q] += sgn * h
            rbm.set_params(p)
            vals.append(orc.log_coeff(rbm, sig))
        fd[q] = (vals[0] - vals[1]) / (2 * h)
    rbm.set_params(p0)
    np.testing.assert_allclose(O, fd, atol=2e-9)
    # no-bias layout keeps only the W block (RBM.hpp:401-402)
    rbm2 = orc.RBM(n, m, rbm.W, use_bias=False)
    O2 = orc.logderiv(rbm2, sig, orc.theta(rbm2, sig)[0])
    assert O2.shape == (n * m,)


HAMS = [
    ("tfi", lambda orc, n: orc.ham(orc.TFI, n, J=-1.0, h=-1.0)),
    ("xxz", lambda orc, n: orc.ham(orc.XXZ, n, J=1.0, delta=1.0, sign_rule=1)),
    ("j1j2", lambda orc, n: orc.ham(orc.J1J2, n, J=1.0, J2=0.5, sign_rule=1)),
]


def test_log_ratio_vector_brute_force(orc):
    """logRatio(vector) (RBMState.hpp:108-126): flipping every listed site,
    for lists of 1 to n distinct sites, against the brute-force psi table;
    one- and two-site lists equal logRatio(k) and logRatio(k, l)."""
    rng = np.random.default_rng(12)
    n = 10
    rbm = orc.RBM.random(n, 14, 0.7, rng)
    basis, th, logpsi = _psi_table(orc, rbm)
    for _ in range(25):
        v = int(rng.integers(0, 1 << n))
        for k in (1, 2, 3, 5, n):
            sites = rng.choice(n, k, replace=False)
            w = v
            for s in sites:
                w ^= 1 << int(s)
            got = orc.log_ratio_vec(rbm, basis[v], th[v], sites)
            assert abs(got - (logpsi[w] - logpsi[v])) < 1e-11
        k, l = (int(x) for x in rng.choice(n, 2, replace=False))
        assert orc.log_ratio_vec(rbm, basis[v], th[v], [k]) == \
            orc.log_ratio1(rbm, basis[v], th[v], k)
        assert orc.log_ratio_vec(rbm, basis[v], th[v], [k, l]) == \
            orc.log_ratio2(rbm, basis[v], th[v], k, l)


@pytest.mark.parametrize("name,mk", HAMS)
def test_eloc_brute_force(orc, name, mk):
    rng = np.random.default_rng(4)
    n = 10
    rbm = orc.RBM.random(n, 2 * n, 0.3, rng)
    H = mk(orc, n)
    Hm = orc.ham_matrix(H)
    basis, th, logpsi = _psi_table(orc, rbm)
    psi = np.exp(logpsi - logpsi.max())
    Hpsi = Hm @ psi
    for v in rng.integers(0, 1 << n, 25):
        got = orc.eloc(H, rbm, basis[v], th[v])[0]
        want = Hpsi[v] / psi[v]
        assert abs(got - want) < 1e-10 * max(1, abs(want))
        rep = orc.eloc_replay(H, rbm, basis[v], th[v])[0]
        assert abs(rep - got) < 1e-12 * max(1, abs(got))
        rep32 = orc.eloc_replay(H, rbm, basis[v], th[v], precision=orc.MIXED)[0]
        assert abs(rep32 - got) < 2e-5 * max(1, abs(got))


def test_ed_known_energies(orc):
    import scipy.sparse.linalg as sla
    # TFI PBC, |J| = |h| = 1: E0 = -2 / sin(pi / 2N); sign of h does not matter
    for n, h in ((12, -1.0), (12, 1.0), (10, -1.0)):
        e = sla.eigsh(orc.ham_matrix(orc.ham(orc.TFI, n, J=-1.0, h=h)), k=1, which="SA")[0][0]
        assert abs(e - (-2 / math.sin(math.pi / (2 * n)))) < 1e-9
    # J1-J2 at the Majumdar-Ghosh point J2/J1 = 1/2 (Pauli units): E0 = -3N/2
    e = sla.eigsh(orc.ham_matrix(orc.ham(orc.J1J2, 10, J=1.0, J2=0.5, sign_rule=1)), k=1,
                  which="SA")[0][0]
    assert abs(e + 15.0) < 1e-9
    # Heisenberg ring N=10 (Pauli units: 4x the spin-1/2 value -4.515446354...)
    e = sla.eigsh(orc.ham_matrix(orc.ham(orc.XXZ, 10, J=1.0, delta=1.0, sign_rule=1)), k=1,
                  which="SA")[0][0]
    assert abs(e - 4 * (-4.515446354)) < 1e-7


def test_ham_column_matches_eloc_functor(orc):
    """HamiltonianHelper.hpp technique: with psi == 1 on every state, E_loc(sigma)
    is the column sum of H; with a one-hot psi it isolates single elements."""
    for name, mk in HAMS:
        n = 8
        H = mk(orc, n)
        Hm = orc.ham_matrix(H).toarray()
        assert np.allclose(Hm, Hm.T)
        rbm = orc.RBM(n, 1)  # W = a = b = 0: every ratio is 1
        basis = orc.basis_sigma(n)
        th = orc.theta(rbm, basis)
        el = orc.eloc(H, rbm, basis, th)
        np.testing.assert_allclose(el, Hm.sum(axis=1), atol=1e-12)


def _exact_sr(orc, rbm, H):
    basis, th, logpsi = _psi_table(orc, rbm)
    w = np.exp(2 * (logpsi - logpsi.max()))
    w /= w.sum()
    el = orc.eloc(H, rbm, basis, th)
    O, om, F, S = orc.sr_build(rbm, basis, th, el, weights=w)
    return basis, w, el, F, S


def test_force_is_half_energy_gradient(orc):
    """F = <O_c^T E_loc> (SRMat.hpp:98-102) equals dE/dtheta / 2 for a real RBM."""
    rng = np.random.default_rng(5)
    n = 8
    rbm = orc.RBM.random(n, n, 0.2, rng)
    H = orc.ham(orc.TFI, n, J=-1.0, h=-1.0)
    Hm = orc.ham_matrix(H)

    def energy(r):
        basis = orc.basis_sigma(n)
        th = orc.theta(r, basis)
        lp = np.array([orc.log_coeff(r, basis[s], th[s]) for s in range(len(basis))])
        psi = np.exp(lp - lp.max())
        return psi @ (Hm @ psi) / (psi @ psi)

    _, w, el, F, S = _exact_sr(orc, rbm, H)
    assert abs(w @ el - energy(rbm)) < 1e-11
    p0 = rbm.params()
    h = 1e-5
    for q in rng.choice(len(p0), 12, replace=False):
        vals = []
        for sg in (1, -1):
            p = p0.copy()
            p[q] += sg * h
            rbm.set_params(p)
            vals.append(energy(rbm))
        rbm.set_params(p0)
        assert abs((vals[0] - vals[1]) / (2 * h) - 2 * F[q]) < 1e-7


def test_S_is_fubini_study_metric(orc):
    rng = np.random.default_rng(6)
    n = 8
    rbm = orc.RBM.random(n, n, 0.2, rng)
    H = orc.ham(orc.TFI, n, J=-1.0, h=-1.0)
    _, w, el, F, S = _exact_sr(orc, rbm, H)
    assert np.allclose(S, S.T) and np.linalg.eigvalsh(S).min() > -1e-12
    basis = orc.basis_sigma(n)

    def psi_of(r):
        th = orc.theta(r, basis)
        lp = np.array([orc.log_coeff(r, basis[s], th[s]) for s in range(len(basis))])
        ps = np.exp(lp - lp.max())
        return ps / np.linalg.norm(ps)

    p0 = rbm.params()
    psi0 = psi_of(rbm)
    eps = 1e-4
    for _ in range(4):
        v = rng.normal(size=len(p0))
        rbm.set_params(p0 + eps * v)
        d2 = 1 - (psi0 @ psi_of(rbm)) ** 2
        rbm.set_params(p0)
        want = eps ** 2 * (v @ S @ v)
        assert abs(d2 - want) < 2e-3 * want


def test_mc_sr_build_and_solvers(orc):
    rng = np.random.default_rng(7)
    n, m, ns = 6, 12, 40
    rbm = orc.RBM.random(n, m, 0.3, rng)
    H = orc.ham(orc.TFI, n, J=-1.0, h=-1.0)
    sig = rng.choice([-1, 1], size=(ns, n)).astype(np.int8)
    th = orc.theta(rbm, sig)
    el = orc.eloc(H, rbm, sig, th)
    O, om, F, S = orc.sr_build(rbm, sig, th, el)
    # unweighted == weighted with uniform weights 1/Ns
    O2, om2, F2, S2 = orc.sr_build(rbm, sig, th, el, weights=np.full(ns, 1.0 / ns))
    np.testing.assert_allclose(F, F2, atol=1e-14)
    np.testing.assert_allclose(S, S2, atol=1e-14)
    # rows of O are the log-derivatives, columns of O_c have zero mean
    raw = np.array([orc.logderiv(rbm, sig[s], th[s]) for s in range(ns)])
    np.testing.assert_allclose(raw.mean(axis=0), om, atol=1e-15)
    np.testing.assert_allclose(O.mean(axis=0), 0, atol=1e-15)
    # S is rank-deficient (rank <= Ns-1 < dim): SPD only through the shift
    assert np.linalg.matrix_rank(S) <= ns - 1
    lam = 1e-3
    x = orc.llt_solve(S, lam, F)
    np.testing.assert_allclose((S + lam * np.eye(len(F))) @ x, F, atol=1e-12)
    with pytest.raises(np.linalg.LinAlgError):
        orc.llt_solve(S, 0.0, F)
    xc, it, err, rc = orc.cg_solve(O, lam, F, 1e-12, 10 * len(F))
    assert rc == 0 and err < 1e-12
    np.testing.assert_allclose(xc, x, rtol=1e-7, atol=1e-9)
    # max-iteration exit is reported
    _, it2, err2, rc2 = orc.cg_solve(O, lam, F, 1e-14, 3)
    assert rc2 == 1 and it2 == 3


def test_sgd(orc):
    v = np.arange(5.0)
    np.testing.assert_allclose(orc.sgd_update(v, 0.01, 0.0, 1), -0.01 * v)
    np.testing.assert_allclose(orc.sgd_update(v, 0.01, 1.0, 1000), -1e-4 * v)  # floor


def test_randomize(orc):
    s = orc.randomize(11, 0, np.arange(64), 64, nup=32)
    assert (s.sum(axis=1) == 0).all()
    s1 = orc.randomize(11, 0, np.arange(2000), 37)
    assert abs(s1.mean()) < 0.02 and set(np.unique(s1)) == {-1, 1}
    assert not np.array_equal(orc.randomize(11, 1, np.arange(4), 37), s1[:4])


def _python_sweep(orc, rbm, sig, seed, stream, ctr, kind, beta=1.0):
    """One sweep replayed with the libm restatement (or_log_ratio*) and the
    Philox draws; decisions must coincide with the kernel-arithmetic replay."""
    n = rbm.n
    sig = sig.copy()
    for _ in range(n):
        x = orc.philox((ctr & 0xffffffff, ctr >> 32, stream, 0), (seed & 0xffffffff, seed >> 32))
        ctr += 1
        k = (int(x[0]) * n) >> 32
        th = orc.theta(rbm, sig)[0]
        if kind == orc.SWAP:
            l = (int(x[1]) * n) >> 32
            if sig[k] == sig[l]:
                continue
            d = orc.log_ratio2(rbm, sig, th, k, l)
        else:
            d = orc.log_ratio1(rbm, sig, th, k)
        u = (((int(x[2]) >> 5) << 26) | (int(x[3]) >> 6)) / 2.0 ** 53
        if u < min(1.0, math.exp(beta * 2 * d)):
            sig[k] = -sig[k]
            if kind == orc.SWAP:
                sig[l] = -sig[l]
    return sig


@pytest.mark.parametrize("kind", [0, 1])
def test_sampler_replay_matches_restatement(orc, kind):
    rng = np.random.default_rng(8)
    n, m, Cn = 12, 24, 6
    rbm = orc.RBM.random(n, m, 0.4, rng)
    sig0 = orc.randomize(5, 0, np.arange(Cn), n, nup=n // 2 if kind else -1)
    seed = 0x1234_5678_9abc
    fin, ctr, acc, ss, st = orc.sample_replay(rbm, sig0, seed=seed, n_sweeps=3, n_discard=0,
                                              kind=kind)
    assert (ctr == 3 * n).all() and (acc > 0).all()
    for c in range(Cn):
        s = sig0[c]
        for sw in range(3):
            s = _python_sweep(orc, rbm, s, seed, c, sw * n, kind)
            np.testing.assert_array_equal(ss[sw * Cn + c], s)
    np.testing.assert_array_equal(ss[2 * Cn:], fin)
    # theta cache invariant (TestRBMStates.cpp:214-236)
    np.testing.assert_allclose(st, orc.theta(rbm, ss), atol=1e-13)
    if kind == orc.SWAP:
        assert (ss.astype(int).sum(axis=1) == 0).all()
    # lanes change only the reduction order, so decisions almost never move
    fin16 = orc.sample_replay(rbm, sig0, seed=seed, n_sweeps=3, n_discard=0, kind=kind,
                              lanes=16)[0]
    assert (fin16 == fin).mean() > 0.99


def test_sampler_distribution_matches_psi2(orc):
    """Chi-square of the sampled histogram against |psi|^2 (N=6)."""
    rng = np.random.default_rng(9)
    n = 6
    rbm = orc.RBM.random(n, 2 * n, 0.5, rng)
    basis, th, logpsi = _psi_table(orc, rbm)
    p = np.exp(2 * (logpsi - logpsi.max()))
    p /= p.sum()
    Cn = 64
    sig0 = orc.randomize(3, 0, np.arange(Cn), n)
    _, _, _, ss, _ = orc.sample_replay(rbm, sig0, seed=77, n_sweeps=400, n_discard=20,
                                       threads=4)
    idx = (((1 - ss.astype(np.int64)) // 2) << np.arange(n)).sum(axis=1)
    cnt = np.bincount(idx, minlength=1 << n)
    N = cnt.sum()
    chi2 = ((cnt - N * p) ** 2 / (N * p)).sum()
    # 63 degrees of freedom; correlated samples inflate chi2 somewhat
    assert chi2 < 3 * 63, chi2
    for precision in (orc.MIXED,):
        fin, _, _, ss2, st2 = orc.sample_replay(rbm, sig0, seed=77, n_sweeps=20, n_discard=0,
                                                precision=precision)
        np.testing.assert_allclose(st2, orc.theta(rbm, ss2), atol=2e-7)


@pytest.mark.parametrize("kind", [0, 1])
def test_tempering_replay_matches_restatement(orc, kind):
    """Parallel tempering (SamplerMT.hpp:120-153): sweeps at beta_t = (T-t)/T
    with the libm restatement, then the two swap passes with
    p = exp((beta_{t+1} - beta_t) 2 (log psi_t - log psi_{t+1})) from
    or_log_coeff and the lower chain's mix-domain Philox draw."""
    rng = np.random.default_rng(18)
    n, m, T, G = 12, 24, 4, 2
    rbm = orc.RBM.random(n, m, 0.4, rng)
    Cn = T * G
    sig0 = orc.randomize(6, 0, np.arange(Cn), n, nup=n // 2 if kind else -1)
    seed = 0x2468_ace0_1357
    sweeps = 4
    fin, ctr, acc, ss, st = orc.sample_replay_pt(rbm, sig0, n_temps=T, seed=seed,
                                                 n_sweeps=sweeps, n_discard=0, kind=kind)
    assert (ctr == sweeps * n).all()
    swaps = 0
    for g in range(G):
        state = [sig0[g * T + t].copy() for t in range(T)]
        for sw in range(sweeps):
            for t in range(T):
                c = g * T + t
                state[t] = _python_sweep(orc, rbm, state[t], seed, c, sw * n, kind,
                                         beta=(T - t) / T)
            lp = [orc.log_coeff(rbm, state[t]) for t in range(T)]
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
    np.testing.assert_allclose(st, orc.theta(rbm, ss), atol=1e-13)


def test_tempering_distribution_matches_psi2(orc):
    """The beta = 1 slot of a tempered chain samples |psi|^2 (chi^2, N = 6),
    with swaps that move states between the temperatures."""
    rng = np.random.default_rng(19)
    n = 6
    rbm = orc.RBM.random(n, 2 * n, 0.8, rng)
    basis, th, logpsi = _psi_table(orc, rbm)
    p = np.exp(2 * (logpsi - logpsi.max()))
    p /= p.sum()
    T, G = 4, 32
    sig0 = orc.randomize(3, 0, np.arange(T * G), n)
    _, _, _, ss, st = orc.sample_replay_pt(rbm, sig0, n_temps=T, seed=91, n_sweeps=400,
                                           n_discard=20)
    idx = (((1 - ss.astype(np.int64)) // 2) << np.arange(n)).sum(axis=1)
    cnt = np.bincount(idx, minlength=1 << n)
    N = cnt.sum()
    chi2 = ((cnt - N * p) ** 2 / (N * p)).sum()
    assert chi2 < 3 * 63, chi2
    _, _, _, ss2, st2 = orc.sample_replay_pt(rbm, sig0, n_temps=T, seed=91, n_sweeps=20,
                                             n_discard=0, precision=orc.MIXED)
    np.testing.assert_allclose(st2, orc.theta(rbm, ss2), atol=2e-7)
