"""CPU oracle for the VMC / SR step.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
always as the checker or the timed CPU baseline; the library never imports
it.  The C++ sources are ``oracle/oracle.cpp`` (restatement ``or_*``, kernel
replay ``rp_*``, threaded baseline ``bl_*``); this module is a ctypes binding
plus small numpy conveniences.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

TFI, XXZ, J1J2 = 0, 1, 2
LOCAL, SWAP = 0, 1
FP64, MIXED = 0, 1


class Ham(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("J", C.c_double),
                ("h", C.c_double), ("delta", C.c_double), ("J2", C.c_double),
                ("sign_rule", C.c_int)]


class RpCfg(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("use_bias", C.c_int),
                ("kind", C.c_int), ("precision", C.c_int), ("lanes", C.c_int),
                ("sweep_factor", C.c_int), ("beta", C.c_double),
                ("seed", C.c_uint64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        d, i, u64, u32, p = C.c_double, C.c_int, C.c_uint64, C.c_uint32, C.c_void_p
        sig = {
            "or_logcosh": (d, [d]),
            "or_theta": (None, [i, i, p, p, p, p]),
            "or_log_ratio1": (d, [i, i, p, p, p, p, i]),
            "or_log_ratio2": (d, [i, i, p, p, p, p, i, i]),
            "or_log_ratio_vec": (d, [i, i, p, p, p, p, i, p]),
            "or_log_coeff": (d, [i, i, p, p, p]),
            "or_logderiv": (None, [i, i, i, p, p, p]),
            "or_eloc": (d, [p, i, i, p, p, p, p]),
            "or_ham_column": (i, [p, u32, p, p, i]),
            "or_sr_build": (None, [i, i, i, i, p, p, p, p, p, p, p, p]),
            "or_llt_solve": (i, [i, p, d, p, p]),
            "or_cg_solve": (i, [i, i, p, p, d, p, d, i, p, p, p]),
            "or_sgd_update": (None, [i, d, d, i, p, p]),
            "rp_theta": (None, [i, i, p, p, p, p]),
            "rp_logcosh_d": (d, [d]),
            "rp_logcosh_f": (C.c_float, [C.c_float]),
            "rp_exp_d": (d, [d]),
            "rp_exp_f": (C.c_float, [C.c_float]),
            "rp_philox": (None, [u32, u32, u32, u32, u32, u32, p]),
            "rp_sample": (i, [p, i, p, p, p, p, p, p, p, i, i, p, p]),
            "rp_sample_pt": (i, [p, i, i, p, p, p, p, p, p, p, i, i, p, p]),
            "rp_eloc": (None, [p, i, i, i, i, p, p, p, p, p]),
            "rp_randomize": (None, [u64, u64, i, p, i, i, p]),
            "or_ff_npar": (i, [p]),
            "or_ff_forward": (d, [p, p, p]),
            "or_ff_jacobian": (d, [p, p, p, p]),
            "or_ffs_check": (i, [i, i, p]),
            "bl_threads": (i, []),
            "bl_sample": (i, [p, i, p, p, p, p, p, p, p, i, i, p, p, i]),
            "bl_eloc": (None, [p, i, i, i, p, p, p, p, p, i]),
            "bl_sr_build": (None, [i, i, i, i, p, p, p, p, p, p, i]),
            "bl_cg_solve": (i, [i, i, p, d, p, d, i, p, p, p, i]),
            "or_c_logcosh": (None, [d, d, p]),
            "or_c_theta": (None, [i, i, p, p, p, p]),
            "or_c_log_ratio1": (None, [i, i, p, p, p, p, i, p]),
            "or_c_log_ratio2": (None, [i, i, p, p, p, p, i, i, p]),
            "or_c_log_ratio_vec": (None, [i, i, p, p, p, p, i, p, p]),
            "or_c_log_coeff": (None, [i, i, p, p, p, p]),
            "or_c_logderiv": (None, [i, i, i, p, p, p]),
            "or_c_eloc": (None, [p, i, i, p, p, p, p, p]),
            "or_c_sr_build": (None, [i, i, i, i, p, p, p, p, p, p, p, p]),
            "rp_c_sample": (i, [p, i, p, p, p, p, p, p, p, i, i, p, p]),
            "rp_c_sample_pt": (i, [p, i, i, p, p, p, p, p, p, p, i, i, p, p]),
            "rp_c_eloc": (None, [p, i, i, i, i, p, p, p, p, p]),
            "rp_logcosh_re_d": (d, [d, d]),
            "rp_logcosh_re_f": (C.c_float, [C.c_float, C.c_float]),
            "rp_logcosh_c_d": (None, [d, d, p]),
            "rp_logcosh_c_f": (None, [C.c_float, C.c_float, p]),
            "rp_tanh_c_d": (None, [d, d, p]),
            "or_ap_log_psi": (None, [i, i, p, p, p, p, p, p, p]),
            "or_ap_log_ratio": (None, [i, i, p, p, p, p, p, p, i, i, p]),
            "or_ap_eloc": (None, [p, i, i, p, p, p, p, p, p, p]),
            "or_ap_logderiv": (None, [i, i, i, p, p, p, p, p]),
            "or_ap_ngd_exact": (i, [p, i, i, i, p, p, p, p, i, p, p, p, p, p, p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i8(a):
    return np.ascontiguousarray(a, dtype=np.int8)


def ham(kind, n, J=-1.0, h=-1.0, delta=1.0, J2=0.0, sign_rule=0):
    return Ham(kind, n, J, h, delta, J2, sign_rule)


# ---------------------------------------------------------------- RBM helpers
class RBM:
    """Real double-precision RBM parameters in the [W col-major | a | b]
    layout of Machines/RBM.hpp:271-281."""

    def __init__(self, n, m, W=None, a=None, b=None, use_bias=True):
        self.n, self.m, self.use_bias = n, m, use_bias
        self.W = f64(np.zeros((n, m)) if W is None else W)  # W[i, j] = W_ji
        self.a = f64(np.zeros(n) if a is None else a)
        self.b = f64(np.zeros(m) if b is None else b)

    @property
    def dim(self):
        return self.n * self.m + (self.n + self.m if self.use_bias else 0)

    @classmethod
    def random(cls, n, m, scale, rng, use_bias=True):
        W = rng.normal(0, scale, (n, m))
        a = rng.normal(0, scale, n) if use_bias else np.zeros(n)
        b = rng.normal(0, scale, m) if use_bias else np.zeros(m)
        return cls(n, m, W, a, b, use_bias)

    def params(self):
        v = [self.W.reshape(-1)]
        if self.use_bias:
            v += [self.a, self.b]
        return np.concatenate(v)

    def set_params(self, v):
        v = f64(v)
        self.W = f64(v[: self.n * self.m].reshape(self.n, self.m))
        if self.use_bias:
            self.a = f64(v[self.n * self.m: self.n * self.m + self.n])
            self.b = f64(v[self.n * self.m + self.n:])


def theta(rbm, sigma):
    sigma = i8(np.atleast_2d(sigma))
    out = np.empty((sigma.shape[0], rbm.m))
    L = lib()
    for s in range(sigma.shape[0]):
        L.or_theta(rbm.n, rbm.m, _p(rbm.W), _p(rbm.b), _p(sigma[s]), _p(out[s]))
    return out


def theta_replay(rbm, sigma):
    sigma = i8(np.atleast_2d(sigma))
    out = np.empty((sigma.shape[0], rbm.m))
    L = lib()
    for s in range(sigma.shape[0]):
        L.rp_theta(rbm.n, rbm.m, _p(rbm.W), _p(rbm.b), _p(sigma[s]), _p(out[s]))
    return out


def log_ratio1(rbm, sigma, th, k):
    sigma, th = i8(sigma), f64(th)
    return lib().or_log_ratio1(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(sigma), _p(th), k)


def log_ratio2(rbm, sigma, th, k, l):
    sigma, th = i8(sigma), f64(th)
    return lib().or_log_ratio2(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(sigma), _p(th), k, l)


def log_ratio_vec(rbm, sigma, th, v):
    """logRatio(vector): flip every listed site (RBMState.hpp:108-126)."""
    sigma, th = i8(sigma), f64(th)
    v = np.ascontiguousarray(v, dtype=np.int32)
    return lib().or_log_ratio_vec(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(sigma), _p(th), len(v),
                                  _p(v))


def log_coeff(rbm, sigma, th=None):
    sigma = i8(sigma)
    th = theta(rbm, sigma)[0] if th is None else f64(th)
    return lib().or_log_coeff(rbm.n, rbm.m, _p(rbm.a), _p(sigma), _p(th))


def logderiv(rbm, sigma, th):
    sigma, th = i8(sigma), f64(th)
    out = np.empty(rbm.dim)
    lib().or_logderiv(rbm.n, rbm.m, int(rbm.use_bias), _p(sigma), _p(th), _p(out))
    return out


def eloc(H, rbm, sigma, th):
    sigma, th = i8(np.atleast_2d(sigma)), f64(np.atleast_2d(th))
    L = lib()
    return np.array([L.or_eloc(C.byref(H), rbm.n, rbm.m, _p(rbm.W), _p(rbm.a),
                               _p(sigma[s]), _p(th[s])) for s in range(sigma.shape[0])])


def eloc_replay(H, rbm, sigma, th, precision=FP64):
    sigma, th = i8(np.atleast_2d(sigma)), f64(np.atleast_2d(th))
    out = np.empty(sigma.shape[0])
    lib().rp_eloc(C.byref(H), precision, sigma.shape[0], rbm.n, rbm.m, _p(rbm.W),
                  _p(rbm.a), _p(sigma), _p(th), _p(out))
    return out


def ham_matrix(H):
    """Dense/sparse H from the ED column functors (small n only)."""
    import scipy.sparse as sp
    n = H.n
    N = 1 << n
    rows_buf = np.empty(8 * n + 8, dtype=np.uint32)
    vals_buf = np.empty(8 * n + 8)
    L = lib()
    r, c, v = [], [], []
    for col in range(N):
        cnt = L.or_ham_column(C.byref(H), col, _p(rows_buf), _p(vals_buf), len(vals_buf))
        r.append(rows_buf[:cnt].copy())
        v.append(vals_buf[:cnt].copy())
        c.append(np.full(cnt, col))
    return sp.csr_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=(N, N))


def basis_sigma(n):
    """All 2^n configurations; bit i set <=> sigma_i = -1 (Utility.hpp:91-99)."""
    idx = np.arange(1 << n, dtype=np.uint64)[:, None]
    bits = (idx >> np.arange(n, dtype=np.uint64)[None, :]) & 1
    return (1 - 2 * bits).astype(np.int8)


def sr_build(rbm, sigma, th, el, weights=None, with_S=True):
    sigma, th, el = i8(sigma), f64(th), f64(el)
    ns = sigma.shape[0]
    O = np.empty((ns, rbm.dim))
    om = np.empty(rbm.dim)
    F = np.empty(rbm.dim)
    S = np.empty((rbm.dim, rbm.dim)) if with_S else None
    w = f64(weights) if weights is not None else None
    lib().or_sr_build(ns, rbm.n, rbm.m, int(rbm.use_bias), _p(sigma), _p(th), _p(el),
                      _p(w), _p(O), _p(om), _p(F), _p(S))
    return O, om, F, S


def llt_solve(S, shift, F):
    S, F = f64(S), f64(F)
    x = np.empty_like(F)
    info = lib().or_llt_solve(len(F), _p(S), shift, _p(F), _p(x))
    if info:
        raise np.linalg.LinAlgError(f"not SPD at pivot {info - 1}")
    return x


def cg_solve(Oc, shift, b, tol, max_iter, weights=None):
    Oc, b = f64(Oc), f64(b)
    x = np.empty_like(b)
    it = C.c_int()
    err = C.c_double()
    w = f64(weights) if weights is not None else None
    rc = lib().or_cg_solve(Oc.shape[0], Oc.shape[1], _p(Oc), _p(w), shift, _p(b), tol,
                           max_iter, _p(x), C.byref(it), C.byref(err))
    return x, it.value, err.value, rc


def sgd_update(v, alpha, p, t):
    v = f64(v)
    out = np.empty_like(v)
    lib().or_sgd_update(len(v), alpha, p, t, _p(v), _p(out))
    return out


def philox(ctr, key):
    out = np.empty(4, dtype=np.uint32)
    lib().rp_philox(*[int(c) for c in ctr], *[int(k) for k in key], _p(out))
    return out


def randomize(seed, epoch, streams, n, nup=-1):
    streams = np.ascontiguousarray(streams, dtype=np.uint32)
    out = np.empty((len(streams), n), dtype=np.int8)
    lib().rp_randomize(seed, epoch, len(streams), _p(streams), n, nup, _p(out))
    return out


def sample_replay(rbm, sigma, *, seed, n_sweeps, n_discard, kind=LOCAL, precision=FP64,
                  lanes=64, sweep_factor=1, beta=1.0, streams=None, ctr=None, acc=None,
                  threads=1):
    """Replay of the sampler kernel.  Returns (sigma_final, ctr, acc,
    samples_sigma (n_sweeps*C x n), samples_theta (n_sweeps*C x m))."""
    sigma = i8(sigma).copy()
    Cn = sigma.shape[0]
    streams = np.arange(Cn, dtype=np.uint32) if streams is None else \
        np.ascontiguousarray(streams, dtype=np.uint32)
    ctr = np.zeros(Cn, np.uint64) if ctr is None else np.array(ctr, dtype=np.uint64)
    acc = np.zeros(Cn, np.uint64) if acc is None else np.array(acc, dtype=np.uint64)
    cfg = RpCfg(rbm.n, rbm.m, int(rbm.use_bias), kind, precision, lanes, sweep_factor,
                beta, seed)
    out_s = np.empty((n_sweeps * Cn, rbm.n), dtype=np.int8)
    out_t = np.empty((n_sweeps * Cn, rbm.m))
    L = lib()
    if threads > 1:
        rc = L.bl_sample(C.byref(cfg), Cn, _p(streams), _p(rbm.W), _p(rbm.a), _p(rbm.b),
                         _p(sigma), _p(ctr), _p(acc), n_sweeps, n_discard, _p(out_s),
                         _p(out_t), threads)
    else:
        rc = L.rp_sample(C.byref(cfg), Cn, _p(streams), _p(rbm.W), _p(rbm.a), _p(rbm.b),
                         _p(sigma), _p(ctr), _p(acc), n_sweeps, n_discard, _p(out_s),
                         _p(out_t))
    if rc:
        raise ValueError("bad replay configuration")
    return sigma, ctr, acc, out_s, out_t


def sample_replay_pt(rbm, sigma, *, n_temps, seed, n_sweeps, n_discard, kind=LOCAL,
                     precision=FP64, lanes=64, sweep_factor=1, ctr=None, acc=None):
    """Replay of the parallel-tempering kernel: sigma holds all n_temps x groups
    chains (chain g * n_temps + t at beta (n_temps - t) / n_temps).  Returns
    (sigma_final, ctr, acc, samples_sigma (n_sweeps*groups x n), samples_theta)."""
    sigma = i8(sigma).copy()
    Cn = sigma.shape[0]
    groups = Cn // n_temps
    streams = np.arange(Cn, dtype=np.uint32)
    ctr = np.zeros(Cn, np.uint64) if ctr is None else np.array(ctr, dtype=np.uint64)
    acc = np.zeros(Cn, np.uint64) if acc is None else np.array(acc, dtype=np.uint64)
    cfg = RpCfg(rbm.n, rbm.m, int(rbm.use_bias), kind, precision, lanes, sweep_factor, 1.0, seed)
    out_s = np.empty((n_sweeps * groups, rbm.n), dtype=np.int8)
    out_t = np.empty((n_sweeps * groups, rbm.m))
    rc = lib().rp_sample_pt(C.byref(cfg), n_temps, groups, _p(streams), _p(rbm.W), _p(rbm.a),
                            _p(rbm.b), _p(sigma), _p(ctr), _p(acc), n_sweeps, n_discard,
                            _p(out_s), _p(out_t))
    if rc:
        raise ValueError("bad replay configuration")
    return sigma, ctr, acc, out_s, out_t


# ---------------------------------------------------------------- FeedForward
class FFLayer(C.Structure):
    """One layer of a general stack: kind 1 Conv1D(a in, b out, c kernel,
    d stride), 2 FullyConnected(a in, b out), 3 activation a with param,
    4 Mean({a, b}, axis c); the numbering of yq_ff_layer."""
    _fields_ = [("kind", C.c_int), ("a", C.c_int), ("b", C.c_int), ("c", C.c_int),
                ("d", C.c_int), ("use_bias", C.c_int), ("param", C.c_double)]


class FFDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("channels", C.c_int), ("kernel", C.c_int),
                ("hidden", C.c_int), ("nl", C.c_int), ("layers", C.POINTER(FFLayer))]


CONV1D, FC, ACT, MEAN = 1, 2, 3, 4
(A_IDENTITY, A_LNCOSH, A_TANH, A_WEAKTANH, A_SIGMOID, A_RELU, A_LEAKYRELU, A_HARDTANH,
 A_SOFTSHRINK, A_LEAKYHARDTANH, A_SOFTSIGN, A_COS) = range(12)


def ff_stack(n, layers):
    """FFDesc of a general stack on n inputs; layers are tuples
    (kind, a, b, c, d, use_bias, param).  The layer array lives as long as
    the descriptor."""
    arr = (FFLayer * len(layers))(*[FFLayer(*[int(v) for v in L[:6]], float(L[6]))
                                    for L in layers])
    d = FFDesc(n, 0, 0, 0, len(layers), arr)
    d._keep = arr
    if lib().or_ffs_check(n, len(layers), C.byref(arr)):
        raise ValueError("layer shapes do not chain to one output")
    return d


def ff_npar(desc):
    return lib().or_ff_npar(C.byref(desc))


def ff_forward(desc, params, sigma):
    params, sigma = f64(params), i8(np.atleast_2d(sigma))
    L = lib()
    return np.array([L.or_ff_forward(C.byref(desc), _p(params), _p(sigma[s]))
                     for s in range(sigma.shape[0])])


def ff_jacobian(desc, params, sigma):
    """Rows d log psi / d params and the outputs log psi."""
    params, sigma = f64(params), i8(np.atleast_2d(sigma))
    L = lib()
    J = np.empty((sigma.shape[0], ff_npar(desc)))
    out = np.empty(sigma.shape[0])
    for s in range(sigma.shape[0]):
        out[s] = L.or_ff_jacobian(C.byref(desc), _p(params), _p(sigma[s]), _p(J[s]))
    return J, out


# ---------------------------------------------------------------- complex RBM
def c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class CRBM:
    """Complex RBM (RBM<std::complex<double>>) in the [W col-major | a | b]
    layout; arrays are complex128 (interleaved re, im in memory)."""

    def __init__(self, n, m, W=None, a=None, b=None, use_bias=True):
        self.n, self.m, self.use_bias = n, m, use_bias
        self.W = c128(np.zeros((n, m)) if W is None else W)
        self.a = c128(np.zeros(n) if a is None else a)
        self.b = c128(np.zeros(m) if b is None else b)

    @property
    def dim(self):
        return self.n * self.m + (self.n + self.m if self.use_bias else 0)

    @classmethod
    def random(cls, n, m, scale, rng, use_bias=True):
        def cn(*shape):
            return rng.normal(0, scale, shape) + 1j * rng.normal(0, scale, shape)
        a = cn(n) if use_bias else np.zeros(n)
        b = cn(m) if use_bias else np.zeros(m)
        return cls(n, m, cn(n, m), a, b, use_bias)

    def params(self):
        v = [self.W.reshape(-1)]
        if self.use_bias:
            v += [self.a, self.b]
        return np.concatenate(v)

    def set_params(self, v):
        v = c128(v)
        nm = self.n * self.m
        self.W = c128(v[:nm].reshape(self.n, self.m))
        if self.use_bias:
            self.a = c128(v[nm:nm + self.n])
            self.b = c128(v[nm + self.n:])


def c_logcosh(z):
    out = np.empty(2)
    lib().or_c_logcosh(float(np.real(z)), float(np.imag(z)), _p(out))
    return out[0] + 1j * out[1]


def c_theta(rbm, sigma):
    sigma = i8(np.atleast_2d(sigma))
    out = np.empty((sigma.shape[0], rbm.m), np.complex128)
    L = lib()
    for s in range(sigma.shape[0]):
        L.or_c_theta(rbm.n, rbm.m, _p(rbm.W), _p(rbm.b), _p(sigma[s]), _p(out[s]))
    return out


def c_log_ratio1(rbm, sigma, th, k):
    out = np.empty(2)
    lib().or_c_log_ratio1(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(i8(sigma)), _p(c128(th)), k,
                          _p(out))
    return out[0] + 1j * out[1]


def c_log_ratio_vec(rbm, sigma, th, v):
    v = np.ascontiguousarray(v, dtype=np.int32)
    out = np.empty(2)
    lib().or_c_log_ratio_vec(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(i8(sigma)), _p(c128(th)),
                             len(v), _p(v), _p(out))
    return out[0] + 1j * out[1]


def c_log_ratio2(rbm, sigma, th, k, l):
    out = np.empty(2)
    lib().or_c_log_ratio2(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(i8(sigma)), _p(c128(th)), k, l,
                          _p(out))
    return out[0] + 1j * out[1]


def c_log_coeff(rbm, sigma, th=None):
    sigma = i8(sigma)
    th = c_theta(rbm, sigma)[0] if th is None else c128(th)
    out = np.empty(2)
    lib().or_c_log_coeff(rbm.n, rbm.m, _p(rbm.a), _p(sigma), _p(th), _p(out))
    return out[0] + 1j * out[1]


def c_logderiv(rbm, sigma, th):
    out = np.empty(rbm.dim, np.complex128)
    lib().or_c_logderiv(rbm.n, rbm.m, int(rbm.use_bias), _p(i8(sigma)), _p(c128(th)), _p(out))
    return out


def c_eloc(H, rbm, sigma, th):
    sigma, th = i8(np.atleast_2d(sigma)), c128(np.atleast_2d(th))
    L = lib()
    out = np.empty(2)
    res = np.empty(sigma.shape[0], np.complex128)
    for s in range(sigma.shape[0]):
        L.or_c_eloc(C.byref(H), rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(sigma[s]), _p(th[s]),
                    _p(out))
        res[s] = out[0] + 1j * out[1]
    return res


def c_eloc_replay(H, rbm, sigma, th, precision=FP64):
    """Complex E_loc with the kernels' arithmetic (theta rounded to float
    for precision MIXED)."""
    sigma, th = i8(np.atleast_2d(sigma)), c128(np.atleast_2d(th))
    out = np.empty(sigma.shape[0], np.complex128)
    lib().rp_c_eloc(C.byref(H), precision, sigma.shape[0], rbm.n, rbm.m, _p(rbm.W), _p(rbm.a),
                    _p(sigma), _p(th), _p(out))
    return out


def c_sr_build(rbm, sigma, th, el, weights=None, with_S=True):
    sigma, th, el = i8(sigma), c128(th), c128(el)
    ns = sigma.shape[0]
    O = np.empty((ns, rbm.dim), np.complex128)
    om = np.empty(rbm.dim, np.complex128)
    F = np.empty(rbm.dim, np.complex128)
    S = np.empty((rbm.dim, rbm.dim), np.complex128) if with_S else None
    w = f64(weights) if weights is not None else None
    lib().or_c_sr_build(ns, rbm.n, rbm.m, int(rbm.use_bias), _p(sigma), _p(th), _p(el), _p(w),
                        _p(O), _p(om), _p(F), _p(S))
    return O, om, F, S


def c_sample_replay(rbm, sigma, *, seed, n_sweeps, n_discard, kind=LOCAL, precision=FP64,
                    lanes=64, sweep_factor=1, beta=1.0, streams=None, ctr=None, acc=None):
    """Replay of the complex sampler kernel; theta samples are complex128."""
    sigma = i8(sigma).copy()
    Cn = sigma.shape[0]
    streams = np.arange(Cn, dtype=np.uint32) if streams is None else \
        np.ascontiguousarray(streams, dtype=np.uint32)
    ctr = np.zeros(Cn, np.uint64) if ctr is None else np.array(ctr, dtype=np.uint64)
    acc = np.zeros(Cn, np.uint64) if acc is None else np.array(acc, dtype=np.uint64)
    cfg = RpCfg(rbm.n, rbm.m, int(rbm.use_bias), kind, precision, lanes, sweep_factor,
                beta, seed)
    out_s = np.empty((n_sweeps * Cn, rbm.n), dtype=np.int8)
    out_t = np.empty((n_sweeps * Cn, rbm.m), np.complex128)
    rc = lib().rp_c_sample(C.byref(cfg), Cn, _p(streams), _p(rbm.W), _p(rbm.a), _p(rbm.b),
                           _p(sigma), _p(ctr), _p(acc), n_sweeps, n_discard, _p(out_s),
                           _p(out_t))
    if rc:
        raise ValueError("bad replay configuration")
    return sigma, ctr, acc, out_s, out_t


def c_sample_replay_pt(rbm, sigma, *, n_temps, seed, n_sweeps, n_discard, kind=LOCAL,
                       precision=FP64, lanes=64, sweep_factor=1, ctr=None, acc=None):
    """Replay of the complex parallel-tempering kernel (sample_replay_pt's
    arrangement; theta samples complex128)."""
    sigma = i8(sigma).copy()
    Cn = sigma.shape[0]
    groups = Cn // n_temps
    streams = np.arange(Cn, dtype=np.uint32)
    ctr = np.zeros(Cn, np.uint64) if ctr is None else np.array(ctr, dtype=np.uint64)
    acc = np.zeros(Cn, np.uint64) if acc is None else np.array(acc, dtype=np.uint64)
    cfg = RpCfg(rbm.n, rbm.m, int(rbm.use_bias), kind, precision, lanes, sweep_factor, 1.0, seed)
    out_s = np.empty((n_sweeps * groups, rbm.n), dtype=np.int8)
    out_t = np.empty((n_sweeps * groups, rbm.m), np.complex128)
    rc = lib().rp_c_sample_pt(C.byref(cfg), n_temps, groups, _p(streams), _p(rbm.W),
                              _p(rbm.a), _p(rbm.b), _p(sigma), _p(ctr), _p(acc), n_sweeps,
                              n_discard, _p(out_s), _p(out_t))
    if rc:
        raise ValueError("bad replay configuration")
    return sigma, ctr, acc, out_s, out_t


# ---------------------------------------------------------------- AmplitudePhase
# psi = sqrt(RBM) exp(i pi FF): a real RBM (oracle.RBM) as the amplitude and
# the FeedForward network (FFDesc, params) as the phase.
def ap_log_psi(rbm, desc, ffp, sigma):
    """log coeff = log RBM coeff / 2 + i pi FF(sigma), per configuration."""
    sigma, ffp = i8(np.atleast_2d(sigma)), f64(ffp)
    L = lib()
    out = np.empty((sigma.shape[0], 2))
    for s in range(sigma.shape[0]):
        L.or_ap_log_psi(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), _p(rbm.b), C.byref(desc),
                        _p(ffp), _p(sigma[s]), _p(out[s]))
    return out[:, 0] + 1j * out[:, 1]


def ap_log_ratio(rbm, desc, ffp, sigma, th, k, l=-1):
    sigma, th, ffp = i8(sigma), f64(th), f64(ffp)
    out = np.empty(2)
    lib().or_ap_log_ratio(rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), C.byref(desc), _p(ffp),
                          _p(sigma), _p(th), k, l, _p(out))
    return out[0] + 1j * out[1]


def ap_eloc(H, rbm, desc, ffp, sigma, th):
    sigma, th, ffp = i8(np.atleast_2d(sigma)), f64(np.atleast_2d(th)), f64(ffp)
    L = lib()
    out = np.empty((sigma.shape[0], 2))
    for s in range(sigma.shape[0]):
        L.or_ap_eloc(C.byref(H), rbm.n, rbm.m, _p(rbm.W), _p(rbm.a), C.byref(desc), _p(ffp),
                     _p(sigma[s]), _p(th[s]), _p(out[s]))
    return out[:, 0] + 1j * out[:, 1]


def ap_logderiv(rbm, desc, ffp, sigma, th):
    """Real rows [RBM O / 2 | pi dFF] (the phase block multiplies i)."""
    sigma, th, ffp = i8(np.atleast_2d(sigma)), f64(np.atleast_2d(th)), f64(ffp)
    L = lib()
    out = np.empty((sigma.shape[0], rbm.dim + ff_npar(desc)))
    for s in range(sigma.shape[0]):
        L.or_ap_logderiv(rbm.n, rbm.m, int(rbm.use_bias), C.byref(desc), _p(ffp),
                         _p(sigma[s]), _p(th[s]), _p(out[s]))
    return out


def ap_ngd_exact(H, rbm, desc, ffp, basis, with_S=True):
    """NGDExact over a sorted basis (integers, bit i set <=> sigma_i = -1)."""
    basis = np.ascontiguousarray(basis, dtype=np.uint32)
    ffp = f64(ffp)
    dA, dP = rbm.dim, ff_npar(desc)
    E = C.c_double()
    oa, op, g = np.empty(dA), np.empty(dP), np.empty(dA + dP)
    Sa = np.empty((dA, dA)) if with_S else None
    Sp = np.empty((dP, dP)) if with_S else None
    rc = lib().or_ap_ngd_exact(C.byref(H), rbm.n, rbm.m, int(rbm.use_bias), _p(f64(rbm.params())),
                               C.byref(desc), _p(ffp), _p(basis), len(basis), C.byref(E),
                               _p(oa), _p(op), _p(g), _p(Sa), _p(Sp))
    if rc:
        raise ValueError("Hamiltonian leaves the basis")
    # column-major outputs: transpose back (the blocks are symmetric anyway)
    return {"E": E.value, "oloc_amp": oa, "oloc_phase": op, "grad": g,
            "S_amp": None if Sa is None else Sa.T.copy(),
            "S_phase": None if Sp is None else Sp.T.copy()}


def basis_jz(n, nup):
    """BasisJz order: the integers below 2^n with nup set bits, increasing."""
    from itertools import combinations
    out = sorted(sum(1 << i for i in c) for c in combinations(range(n), nup))
    return np.array(out, dtype=np.uint32)


def basis_to_sigma(basis, n):
    b = np.asarray(basis, dtype=np.uint64)[:, None]
    bits = (b >> np.arange(n, dtype=np.uint64)[None, :]) & 1
    return (1 - 2 * bits).astype(np.int8)
