"""Pins of the AmplitudePhase restatement (oracle/oracle_ap.cpp):
psi = sqrt(RBM) exp(i pi FF), Machines/AmplitudePhase.hpp,
States/AmplitudePhaseState.hpp, GroundState/NGDExact.hpp.

The reference's own test (Tests/TestAmplitudePhase.cpp) includes files that
do not exist and holds no vectors, so every function here is pinned by what
the mathematics fixes:
  * E_loc(s) = (H psi)_s / psi_s over the full basis, H from exact
    diagonalisation's column functors;
  * the log-derivative rows are the finite differences of log psi (Re for
    the amplitude block, Im for the phase block, zero across);
  * NGDExact's energy is <psi|H|psi>/<psi|psi>, its gradient the finite
    difference of that energy, and its metric blocks the Fubini-Study metric
    Re[<d_k psi|d_l psi>/N - <d_k psi|psi><psi|d_l psi>/N^2] of the state
    vector, whose amplitude-phase block vanishes;
  * Metropolis at beta = 1/2 on the RBM samples |psi_AP|^2 = RBM(sigma).
"""
import numpy as np
import pytest


def _machine(orc, n, m, rng, amp=0.15, phase=0.4, use_bias=True, C=2, K=3, H=3):
    rbm = orc.RBM.random(n, m, amp, rng, use_bias=use_bias)
    desc = orc.FFDesc(n, C, K, H)
    ffp = rng.normal(0, phase, orc.ff_npar(desc))
    return rbm, desc, ffp


def _hams(orc, n):
    return {
        "tfi": orc.ham(orc.TFI, n, J=-1.0, h=-0.7),
        "xxz": orc.ham(orc.XXZ, n, J=1.0, delta=0.6, sign_rule=1),
        "j1j2": orc.ham(orc.J1J2, n, J=1.0, J2=0.44, sign_rule=0),
    }


@pytest.mark.parametrize("name", ["tfi", "xxz", "j1j2"])
def test_ap_eloc_is_H_psi_over_psi(orc, name):
    rng = np.random.default_rng(31)
    n = 8
    rbm, desc, ffp = _machine(orc, n, 2 * n, rng)
    H = _hams(orc, n)[name]
    sig = orc.basis_sigma(n)
    lp = orc.ap_log_psi(rbm, desc, ffp, sig)
    psi = np.exp(lp)
    Hpsi = orc.ham_matrix(H) @ psi
    el = orc.ap_eloc(H, rbm, desc, ffp, sig, orc.theta(rbm, sig))
    np.testing.assert_allclose(el, Hpsi / psi, rtol=1e-11, atol=1e-11)
    # log-ratio of one flip and of a pair: differences of log psi
    th = orc.theta(rbm, sig[5:6])[0]
    want1 = orc.ap_log_psi(rbm, desc, ffp, np.where(np.arange(n) == 2, -sig[5], sig[5]))[0] - lp[5]
    got1 = orc.ap_log_ratio(rbm, desc, ffp, sig[5], th, 2)
    assert abs(got1.real - want1.real) < 1e-12
    assert abs(np.exp(1j * got1.imag) - np.exp(1j * want1.imag)) < 1e-12


def test_ap_logderiv_finite_differences(orc):
    rng = np.random.default_rng(32)
    n = 6
    rbm, desc, ffp = _machine(orc, n, n, rng, amp=0.3, phase=0.5)
    dA, dP = rbm.dim, orc.ff_npar(desc)
    sig = orc.basis_sigma(n)[[3, 17, 40]]
    rows = orc.ap_logderiv(rbm, desc, ffp, sig, orc.theta(rbm, sig))
    h = 1e-5
    p0 = rbm.params().copy()
    fd = np.zeros((len(sig), dA + dP), dtype=complex)
    for k in range(dA):
        vals = []
        for sgn in (1, -1):
            q = p0.copy()
            q[k] += sgn * h
            r = orc.RBM(n, rbm.m, use_bias=rbm.use_bias)
            r.set_params(q)
            vals.append(orc.ap_log_psi(r, desc, ffp, sig))
        fd[:, k] = (vals[0] - vals[1]) / (2 * h)
    for k in range(dP):
        vals = []
        for sgn in (1, -1):
            q = ffp.copy()
            q[k] += sgn * h
            vals.append(orc.ap_log_psi(rbm, desc, q, sig))
        fd[:, dA + k] = (vals[0] - vals[1]) / (2 * h)
    # amplitude block: d Re log psi; phase block: d Im log psi; zero across
    np.testing.assert_allclose(rows[:, :dA], fd[:, :dA].real, rtol=0, atol=2e-8)
    np.testing.assert_allclose(rows[:, dA:], fd[:, dA:].imag, rtol=0, atol=2e-8)
    assert np.abs(fd[:, :dA].imag).max() < 1e-9
    assert np.abs(fd[:, dA:].real).max() < 1e-9
    # the halving and the factor pi are there (a dropped factor fails)
    r = orc.logderiv(rbm, sig[0], orc.theta(rbm, sig[:1])[0])
    np.testing.assert_allclose(rows[0, :dA], r / 2, rtol=1e-15)


def _energy(orc, H, rbm, desc, ffp, basis):
    sig = orc.basis_to_sigma(basis, rbm.n)
    psi = np.exp(orc.ap_log_psi(rbm, desc, ffp, sig))
    full = orc.ham_matrix(H)
    Hs = full[basis][:, basis]
    return float(np.real(np.vdot(psi, Hs @ psi)) / np.vdot(psi, psi).real)


@pytest.mark.parametrize("name,basis", [("tfi", "full"), ("j1j2", "jz"), ("xxz", "jz")])
def test_ap_ngd_exact_energy_and_gradient(orc, name, basis):
    rng = np.random.default_rng(33)
    n = 6
    rbm, desc, ffp = _machine(orc, n, 2 * n, rng, amp=0.25, phase=0.5)
    H = _hams(orc, n)[name]
    B = np.arange(1 << n, dtype=np.uint32) if basis == "full" else orc.basis_jz(n, n // 2)
    r = orc.ap_ngd_exact(H, rbm, desc, ffp, B, with_S=False)
    E0 = _energy(orc, H, rbm, desc, ffp, B)
    assert abs(r["E"] - E0) < 1e-12 * max(1.0, abs(E0))
    dA = rbm.dim
    p0 = np.concatenate([rbm.params(), ffp])
    h = 1e-5

    def energy_at(p):
        rr = orc.RBM(n, rbm.m, use_bias=rbm.use_bias)
        rr.set_params(p[:dA])
        return orc.ap_ngd_exact(H, rr, desc, p[dA:], B, with_S=False)["E"]

    dirs = [rng.normal(size=len(p0)) for _ in range(4)]
    for blk in (slice(0, dA), slice(dA, len(p0))):  # one direction inside each block
        u = np.zeros(len(p0))
        u[blk] = rng.normal(size=u[blk].shape)
        dirs.append(u)
    for u in dirs:
        fd = (energy_at(p0 + h * u) - energy_at(p0 - h * u)) / (2 * h)
        assert abs(fd - r["grad"] @ u) < 1e-7 * max(1.0, abs(fd)), (fd, r["grad"] @ u)


def test_ap_metric_blocks_are_fubini_study(orc):
    rng = np.random.default_rng(34)
    n = 6
    rbm, desc, ffp = _machine(orc, n, n, rng, amp=0.3, phase=0.5)
    H = _hams(orc, n)["j1j2"]
    B = orc.basis_jz(n, 3)
    r = orc.ap_ngd_exact(H, rbm, desc, ffp, B)
    dA, dP = rbm.dim, orc.ff_npar(desc)
    sig = orc.basis_to_sigma(B, n)
    p0 = np.concatenate([rbm.params(), ffp])

    def psi_at(p):
        rr = orc.RBM(n, rbm.m, use_bias=rbm.use_bias)
        rr.set_params(p[:dA])
        return np.exp(orc.ap_log_psi(rr, desc, p[dA:], sig))

    idx = np.concatenate([rng.choice(dA, 8, replace=False), dA + rng.choice(dP, 8, replace=False)])
    psi = psi_at(p0)
    N = np.vdot(psi, psi).real
    h = 1e-5
    dpsi = []
    for k in idx:
        e = np.zeros(len(p0))
        e[k] = h
        dpsi.append((psi_at(p0 + e) - psi_at(p0 - e)) / (2 * h))
    dpsi = np.array(dpsi)
    G = (dpsi.conj() @ dpsi.T / N - np.outer(dpsi.conj() @ psi, psi.conj() @ dpsi.T) / N**2).real
    S = np.zeros((dA + dP, dA + dP))
    S[:dA, :dA] = r["S_amp"]
    S[dA:, dA:] = r["S_phase"]
    np.testing.assert_allclose(S[np.ix_(idx, idx)], G, rtol=0, atol=1e-8 * np.abs(G).max())
    # the amplitude-phase block of the metric is zero
    assert np.abs(G[:8, 8:]).max() < 1e-8 * np.abs(G).max()
    # and the energy gradient of the tangent vectors: F = Re <d_k psi|(H - E)|psi> / N
    Hs = orc.ham_matrix(H)[B][:, B]
    F = (dpsi.conj() @ (Hs @ psi) / N - (dpsi.conj() @ psi) / N * r["E"]).real
    np.testing.assert_allclose(r["grad"][idx] / 2, F, rtol=0, atol=1e-8 * np.abs(F).max())


def test_half_beta_chain_samples_ap_distribution(orc):
    """Metropolis on the RBM at beta = 1/2 (logRatioRe = logRatio / 2) samples
    |psi_AP|^2 = RBM(sigma) / sum: chi^2 over the N = 6 basis."""
    rng = np.random.default_rng(35)
    n = 6
    rbm = orc.RBM.random(n, 2 * n, 0.8, rng)
    basis = orc.basis_sigma(n)
    th = orc.theta(rbm, basis)
    lc = np.array([orc.log_coeff(rbm, basis[s], th[s]) for s in range(len(basis))])
    p = np.exp(lc - lc.max())
    p /= p.sum()
    Cn = 64
    sig0 = orc.randomize(5, 0, np.arange(Cn), n)
    _, _, _, ss, _ = orc.sample_replay(rbm, sig0, seed=78, n_sweeps=400, n_discard=20,
                                       beta=0.5, threads=4)
    idx = (((1 - ss.astype(np.int64)) // 2) << np.arange(n)).sum(axis=1)
    cnt = np.bincount(idx, minlength=1 << n)
    Nn = cnt.sum()
    chi2 = ((cnt - Nn * p) ** 2 / (Nn * p)).sum()
    assert chi2 < 3 * 63, chi2
    # and not |RBM|^2 (the beta = 1 distribution), which differs at this scale
    p1 = np.exp(2 * (lc - lc.max()))
    p1 /= p1.sum()
    chi2_1 = ((cnt - Nn * p1) ** 2 / (Nn * p1)).sum()
    assert chi2_1 > 10 * chi2, (chi2, chi2_1)
