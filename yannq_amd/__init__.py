"""yannq_amd: the VMC / SR step of yannq on one AMD Instinct MI355X (gfx950).

The compute path is the in-tree HIP library ``libyannq_amd.so`` behind the C
ABI declared in ``include/yannq_amd.h``; this module is a thin ctypes binding
of that ABI (used by the tests and the benchmark).  There is no CPU fallback:
if the library is missing or no GPU is visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyannq_amd.so")
HEADER = os.path.join(HERE, "..", "include", "yannq_amd.h")

# enums of include/yannq_amd.h
OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_NAN, ERR_NOT_SPD, ERR_NOT_CONVERGED, ERR_OOM, \
    ERR_UNSUPPORTED = range(9)
TFI, XXZ, J1J2 = 0, 1, 2
LOCAL, EXCHANGE = 0, 1
FP64, MIXED = 0, 1
SOLVE_EXPLICIT, SOLVE_CG, SOLVE_CG_FACTORED = 0, 1, 2
SR_F, SR_OMEAN, SR_O, SR_S, SR_ELOC = range(5)
DTYPE_F64, DTYPE_C128 = 1, 2

STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_STATE", 3: "ERR_HIP", 4: "ERR_NAN",
                5: "ERR_NOT_SPD", 6: "ERR_NOT_CONVERGED", 7: "ERR_OOM", 8: "ERR_UNSUPPORTED"}


class YqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class HamDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("J", C.c_double), ("h", C.c_double),
                ("delta", C.c_double), ("J2", C.c_double), ("sign_rule", C.c_int)]


class MultiDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("use_bias", C.c_int), ("chains", C.c_int),
                ("sweep_factor", C.c_int), ("seed", C.c_uint64), ("ham", HamDesc)]


class FFDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("channels", C.c_int), ("kernel", C.c_int),
                ("hidden", C.c_int), ("dtype", C.c_int)]


class FFLayer(C.Structure):
    """yq_ff_layer: kind CONV1D(a in, b out, c kernel, d stride), FC(a in,
    b out), ACT (a = activation, param), MEAN({a, b}, axis c)."""
    _fields_ = [("kind", C.c_int), ("a", C.c_int), ("b", C.c_int), ("c", C.c_int),
                ("d", C.c_int), ("use_bias", C.c_int), ("param", C.c_double)]


LAYER_CONV1D, LAYER_FC, LAYER_ACT, LAYER_MEAN = 1, 2, 3, 4
(ACT_IDENTITY, ACT_LNCOSH, ACT_TANH, ACT_WEAKTANH, ACT_SIGMOID, ACT_RELU, ACT_LEAKYRELU,
 ACT_HARDTANH, ACT_SOFTSHRINK, ACT_LEAKYHARDTANH, ACT_SOFTSIGN, ACT_COS) = range(12)


class StepCfg(C.Structure):
    _fields_ = [("n_sweeps", C.c_int), ("n_discard", C.c_int), ("randomize", C.c_int),
                ("nup", C.c_int), ("solver", C.c_int), ("precision", C.c_int),
                ("lambda_", C.c_double), ("cg_tol", C.c_double), ("cg_max_iter", C.c_int),
                ("lr", C.c_double), ("grad_clip", C.c_double), ("precond", C.c_int),
                ("warm_start", C.c_int)]


class StepOut(C.Structure):
    _fields_ = [("energy", C.c_double), ("energy_var", C.c_double),
                ("update_norm", C.c_double), ("cg_err", C.c_double), ("cg_iters", C.c_int),
                ("n_samples", C.c_int), ("sample_ms", C.c_double), ("solve_ms", C.c_double)]


_lib = None


def header_symbols():
    """Names of every entry point declared in include/yannq_amd.h."""
    import re
    txt = open(HEADER).read()
    return re.findall(r"YQ_API\s+[^;(]*?\b(yq_\w+)\s*\(", txt)


def lib():
    """Load libyannq_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m yannq_amd.build`")
    L = C.CDLL(LIB_PATH)
    i, d, p, u64 = C.c_int, C.c_double, C.c_void_p, C.c_uint64
    sig = {
        "yq_ctx_create": [i, C.POINTER(p)], "yq_ctx_destroy": [p],
        "yq_rbm_create": [p, i, i, i, i], "yq_rbm_dims": [p, p, p, p],
        "yq_rbm_set_params": [p, p], "yq_rbm_get_params": [p, p],
        "yq_rbm_update_params": [p, p], "yq_rbm_init_random": [p, u64, d],
        "yq_rbm_has_nan": [p, p], "yq_ham_set": [p, p],
        "yq_sampler_config": [p, i, i, i, i, i, u64], "yq_sampler_randomize": [p, i],
        "yq_sampler_set_sigma": [p, p], "yq_sampler_get_state": [p, p, p, p],
        "yq_sample": [p, i, i, p], "yq_samples_get": [p, p, p],
        "yq_samples_set": [p, i, p, i], "yq_eloc": [p, p, p, p],
        "yq_sr_build": [p, i, i], "yq_sr_get": [p, i, p], "yq_sr_apply": [p, d, p, p],
        "yq_sr_solve": [p, d, d, i, p, p, p, p], "yq_sr_step": [p, p, p],
        "yq_prof_enable": [p, C.c_char_p], "yq_prof_read": [p, C.c_char_p, p, p],
        "yq_sr_set_options": [p, i, i],
        "yq_set_graph_mode": [p, i], "yq_set_autotune": [p, i], "yq_set_tiles": [p, i, i],
        "yq_multi_create": [p, i, p, i, i], "yq_multi_set_params": [p, i, p],
        "yq_multi_get_params": [p, i, p], "yq_multi_init_random": [p, C.c_uint64, d],
        "yq_multi_sample": [p, i, i], "yq_multi_samples_get": [p, i, p, p],
        "yq_multi_sampler_state": [p, i, p, p, p], "yq_multi_eloc": [p, p, p],
        "yq_multi_sr_step": [p, p, p, i], "yq_multi_set_grouped": [p, i],
        "yq_ff_create": [p, p], "yq_ff_npar": [p, p], "yq_ff_set_params": [p, p],
        "yq_ff_get_params": [p, p], "yq_ff_update_params": [p, p],
        "yq_ff_init_random": [p, u64, i], "yq_ff_forward": [p, i, p, p],
        "yq_ff_jacobian": [p, i, p, p, p], "yq_ff_forward_dev": [p, i, p, p, p],
        "yq_ff_jacobian_dev": [p, i, p, p, p], "yq_ff_vjp_dev": [p, i, p, p, p, p],
        "yq_rbm_log_ratio": [p, i, p, i, p, p],
        "yq_sr_fisher_average": [p, d, i],
        "yq_ff_vjp_dev_noalloc": [p, i, p, p, p, p],
        "yq_ff_create_stack": [p, i, i, i, p], "yq_ff_act_len": [p, p],
        "yq_ckpt_rbm_encode": [i, i, i, i, p, i, p, C.c_size_t],
        "yq_ckpt_rbm_decode": [i, p, C.c_size_t, i, p, p, p, p, C.c_size_t],
        "yq_rbm_save": [p, C.c_char_p, i, i], "yq_rbm_load": [p, C.c_char_p, i, i],
        "yq_exact_batch": [p, i, p], "yq_exact_probs": [p, p],
        "yq_exact_sample": [p, i, i, u64, p],
        "yq_ap_enable": [p, i], "yq_ap_dims": [p, p, p], "yq_ap_get_params": [p, p],
        "yq_ap_set_params": [p, p], "yq_ap_update_params": [p, p],
        "yq_ap_log_psi": [p, i, p, p], "yq_chain_lanes": [i],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i
    L.yq_ckpt_rbm_bytes.argtypes = [i, i, i, i, i]
    L.yq_ckpt_rbm_bytes.restype = C.c_size_t
    L.yq_last_error.argtypes = [p]
    L.yq_last_error.restype = C.c_char_p
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# checkpoint element types (yq_ckpt_elem): RBM<float>, RBM<double>,
# RBM<complex<float>>, RBM<complex<double>>
CKPT_F32, CKPT_F64, CKPT_C64, CKPT_C128 = 1, 2, 3, 4


def ckpt_encode(params, n, m, use_bias=True, elem=CKPT_F64, pointer=True) -> bytes:
    """The reference's cereal binary encoding of an RBM (host only).  params:
    the [W | a | b] layout; complex element types take a complex array."""
    L = lib()
    cplx = elem in (CKPT_C64, CKPT_C128)
    v = np.ascontiguousarray(params, dtype=np.complex128 if cplx else np.float64)
    nb = L.yq_ckpt_rbm_bytes(elem, n, m, int(use_bias), int(pointer))
    dim = n * m + (n + m if use_bias else 0)
    if nb == 0 or v.size != dim:
        raise YqError(ERR_ARG, "ckpt_encode: bad sizes")
    out = np.empty(nb, dtype=np.uint8)
    rc = L.yq_ckpt_rbm_encode(elem, n, m, int(use_bias), _ptr(v), int(pointer), _ptr(out), nb)
    if rc:
        raise YqError(rc, "ckpt_encode failed")
    return out.tobytes()


def ckpt_decode(buf: bytes, elem=CKPT_F64, pointer=True):
    """Inverse of ckpt_encode: (n, m, use_bias, params)."""
    L = lib()
    b = np.frombuffer(buf, dtype=np.uint8)
    n, m, ub = C.c_int(), C.c_int(), C.c_int()
    rc = L.yq_ckpt_rbm_decode(elem, _ptr(b), b.size, int(pointer), C.byref(n), C.byref(m),
                              C.byref(ub), None, 0)
    if rc:
        raise YqError(rc, "not an RBM checkpoint of this element type and form")
    cplx = elem in (CKPT_C64, CKPT_C128)
    dim = n.value * m.value + (n.value + m.value if ub.value else 0)
    v = np.empty(dim, dtype=np.complex128 if cplx else np.float64)
    rc = L.yq_ckpt_rbm_decode(elem, _ptr(b), b.size, int(pointer), C.byref(n), C.byref(m),
                              C.byref(ub), _ptr(v), v.size * (2 if cplx else 1))
    if rc:
        raise YqError(rc, "malformed RBM checkpoint")
    return n.value, m.value, bool(ub.value), v


class Context:
    """One device context (include/yannq_amd.h); methods mirror the ABI."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.yq_ctx_create(device, C.byref(h))
        if rc:
            raise YqError(rc, "yq_ctx_create failed (is a gfx950 GPU visible?)")
        self._h = h
        self.n = self.m = self.dim = 0
        self.chains = 0
        self.ns = 0
        self.ap = False

    def close(self):
        if getattr(self, "_h", None):
            self._L.yq_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc, allow=()):
        if rc and rc not in allow:
            raise YqError(rc, self._L.yq_last_error(self._h).decode())
        return rc

    # machine
    def rbm_create(self, n, m, use_bias=True, dtype=np.float64):
        """dtype float64 (RBM<double>) or complex128 (RBM<complex<double>>);
        a complex machine takes and returns complex128 arrays."""
        self.dt = np.dtype(dtype)
        code = DTYPE_C128 if self.dt == np.complex128 else DTYPE_F64
        self._chk(self._L.yq_rbm_create(self._h, code, n, m, int(use_bias)))
        self.ap = False
        self.n, self.m = n, m
        self.dim = n * m + (n + m if use_bias else 0)

    @property
    def cplx(self):
        return getattr(self, "dt", np.dtype(np.float64)) == np.complex128

    def _vec(self, v):
        return np.ascontiguousarray(v, dtype=np.complex128 if self.cplx else np.float64)

    def _out(self, *shape):
        return np.empty(shape, np.complex128 if self.cplx else np.float64)

    def set_params(self, v):
        v = self._vec(v)
        assert v.shape == (self.dim,)
        self._chk(self._L.yq_rbm_set_params(self._h, _ptr(v)))

    def get_params(self):
        out = self._out(self.dim)
        self._chk(self._L.yq_rbm_get_params(self._h, _ptr(out)))
        return out

    def update_params(self, d):
        d = self._vec(d)
        self._chk(self._L.yq_rbm_update_params(self._h, _ptr(d)))

    def init_random(self, seed, sigma):
        self._chk(self._L.yq_rbm_init_random(self._h, seed, sigma))

    def has_nan(self):
        out = np.zeros(1, np.int32)
        self._chk(self._L.yq_rbm_has_nan(self._h, _ptr(out)))
        return bool(out[0])

    def ham_set(self, kind, n, J=-1.0, h=-1.0, delta=1.0, J2=0.0, sign_rule=0):
        d = HamDesc(kind, n, J, h, delta, J2, sign_rule)
        self._chk(self._L.yq_ham_set(self._h, C.byref(d)))

    # sampler
    def sampler_config(self, sweeper, chains, precision=FP64, seed=1, sweep_factor=1, n_temps=1):
        self._chk(self._L.yq_sampler_config(self._h, sweeper, n_temps, chains, sweep_factor,
                                            precision, seed))
        self.chains = chains * n_temps   # chain states (n_temps per chain group)
        self.n_temps = n_temps

    def randomize(self, nup=-1):
        self._chk(self._L.yq_sampler_randomize(self._h, nup))

    def set_sigma(self, sigma):
        sigma = np.ascontiguousarray(sigma, dtype=np.int8)
        assert sigma.shape == (self.chains, self.n)
        self._chk(self._L.yq_sampler_set_sigma(self._h, _ptr(sigma)))

    def sampler_state(self):
        s = np.empty((self.chains, self.n), np.int8)
        ctr = np.empty(self.chains, np.uint64)
        acc = np.empty(self.chains, np.uint64)
        self._chk(self._L.yq_sampler_get_state(self._h, _ptr(s), _ptr(ctr), _ptr(acc)))
        return s, ctr, acc

    def sample(self, n_sweeps, n_discard):
        ns = C.c_int()
        self._chk(self._L.yq_sample(self._h, n_sweeps, n_discard, C.byref(ns)))
        self.ns = ns.value
        return ns.value

    def samples(self):
        s = np.empty((self.ns, self.n), np.int8)
        t = self._out(self.ns, self.m)
        self._chk(self._L.yq_samples_get(self._h, _ptr(s), _ptr(t)))
        return s, t

    def log_ratio(self, sigma, sites):
        """logRatio(vector) per row: sigma (ns x n), sites (ns x k, entries < 0
        skipped); real or complex per the machine."""
        sigma = np.ascontiguousarray(sigma, dtype=np.int8)
        sites = np.ascontiguousarray(sites, dtype=np.int32)
        ns, k = sites.shape
        out = np.empty(ns, dtype=np.complex128 if self.cplx else np.float64)
        self._chk(self._L.yq_rbm_log_ratio(self._h, ns, _ptr(sigma), k, _ptr(sites), _ptr(out)))
        return out

    def samples_set(self, sigma, precision=FP64):
        sigma = np.ascontiguousarray(sigma, dtype=np.int8)
        self._chk(self._L.yq_samples_set(self._h, sigma.shape[0], _ptr(sigma), precision))
        self.ns = sigma.shape[0]

    def exact_batch(self, nup=-1):
        """Every configuration of the full basis (nup < 0) or of the sector
        with nup set bits, weighted by |psi|^2 (SRMatExact); returns the
        basis size."""
        B = C.c_int()
        self._chk(self._L.yq_exact_batch(self._h, nup, C.byref(B)))
        self.ns = B.value
        return B.value

    def exact_probs(self):
        out = np.empty(self.ns)
        self._chk(self._L.yq_exact_probs(self._h, _ptr(out)))
        return out

    def exact_sample(self, n_samples, seed, nup=-1):
        """ExactSampler: n_samples draws from |psi|^2; returns their basis
        positions."""
        idx = np.empty(n_samples, np.int32)
        self._chk(self._L.yq_exact_sample(self._h, nup, n_samples, seed, _ptr(idx)))
        self.ns = n_samples
        return idx

    @property
    def sr_dim(self):
        """Length of the SR system's vectors ([RBM | FF] for AmplitudePhase)."""
        return self.dim + self.ff_npar if self.ap else self.dim

    def _eout(self, count):
        return np.empty(count, np.complex128 if (self.cplx or self.ap) else np.float64)

    def eloc(self):
        e = self._eout(self.ns)
        mean, var = C.c_double(), C.c_double()
        self._chk(self._L.yq_eloc(self._h, _ptr(e), C.byref(mean), C.byref(var)))
        return e, mean.value, var.value

    # SR
    def sr_build(self, solver, precision=FP64):
        self._chk(self._L.yq_sr_build(self._h, solver, precision))

    def sr_fisher_average(self, beta2, reset=False):
        self._chk(self._L.yq_sr_fisher_average(self._h, beta2, int(reset)))

    def sr_get(self, item):
        D = self.sr_dim
        if item == SR_ELOC:
            out = self._eout(self.ns)
            self._chk(self._L.yq_sr_get(self._h, item, _ptr(out)))
            return out
        size = {SR_F: D, SR_OMEAN: D, SR_O: self.ns * D, SR_S: D * D}[item]
        out = self._out(size)
        self._chk(self._L.yq_sr_get(self._h, item, _ptr(out)))
        if item == SR_O:
            return out.reshape(self.ns, D)
        if item == SR_S:  # column-major on the ABI
            return out.reshape(D, D).T.copy()
        return out

    def sr_apply(self, shift, v):
        v = self._vec(v)
        out = self._out(self.sr_dim)
        self._chk(self._L.yq_sr_apply(self._h, shift, _ptr(v), _ptr(out)))
        return out

    def sr_solve(self, shift, tol=1e-10, max_iter=10000, rhs=None, allow_unconverged=False):
        x = self._out(self.sr_dim)
        res, it = C.c_double(), C.c_int()
        r = None if rhs is None else self._vec(rhs)
        self._chk(self._L.yq_sr_solve(self._h, shift, tol, max_iter, _ptr(r), _ptr(x),
                                      C.byref(res), C.byref(it)),
                  allow=(ERR_NOT_CONVERGED,) if allow_unconverged else ())
        return x, res.value, it.value

    # FeedForward machine
    def ff_create(self, n, channels, kernel, hidden, dtype=np.float32):
        d = FFDesc(n, channels, kernel, hidden, 1 if np.dtype(dtype) == np.float64 else 0)
        self._chk(self._L.yq_ff_create(self._h, C.byref(d)))
        self.ap = False
        k = C.c_int()
        self._chk(self._L.yq_ff_npar(self._h, C.byref(k)))
        self.ff_n, self.ff_npar = n, k.value
        self.ff_dtype = np.dtype(np.float64 if d.dtype else np.float32)

    def ff_create_stack(self, n, layers, dtype=np.float64):
        """A general layer stack: layers are tuples (kind, a, b, c, d,
        use_bias, param) in addLayer order."""
        arr = (FFLayer * len(layers))(*[FFLayer(*[int(v) for v in L[:6]], float(L[6]))
                                        for L in layers])
        dt = 1 if np.dtype(dtype) == np.float64 else 0
        self._chk(self._L.yq_ff_create_stack(self._h, n, dt, len(layers), C.byref(arr)))
        self.ap = False
        k = C.c_int()
        self._chk(self._L.yq_ff_npar(self._h, C.byref(k)))
        self.ff_n, self.ff_npar = n, k.value
        self.ff_dtype = np.dtype(np.float64 if dt else np.float32)

    def ff_act_len(self):
        k = C.c_int()
        self._chk(self._L.yq_ff_act_len(self._h, C.byref(k)))
        return k.value

    def ff_set_params(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.ff_npar,)
        self._chk(self._L.yq_ff_set_params(self._h, _ptr(v)))

    def ff_get_params(self):
        out = np.empty(self.ff_npar)
        self._chk(self._L.yq_ff_get_params(self._h, _ptr(out)))
        return out

    def ff_init_random(self, seed, mode=3):
        self._chk(self._L.yq_ff_init_random(self._h, seed, mode))

    def ff_forward(self, sigma):
        sigma = np.ascontiguousarray(sigma, dtype=np.int8)
        out = np.empty(sigma.shape[0])
        self._chk(self._L.yq_ff_forward(self._h, sigma.shape[0], _ptr(sigma), _ptr(out)))
        return out

    def ff_jacobian(self, sigma):
        sigma = np.ascontiguousarray(sigma, dtype=np.int8)
        J = np.empty((sigma.shape[0], self.ff_npar))
        out = np.empty(sigma.shape[0])
        self._chk(self._L.yq_ff_jacobian(self._h, sigma.shape[0], _ptr(sigma), _ptr(J),
                                         _ptr(out)))
        return J, out

    # device-pointer forms (integers from e.g. torch.Tensor.data_ptr())
    def ff_forward_dev(self, ns, d_sigma, d_out, d_act=None):
        self._chk(self._L.yq_ff_forward_dev(self._h, ns, d_sigma, d_out, d_act))

    def ff_jacobian_dev(self, ns, d_sigma, d_jac, d_out=None):
        self._chk(self._L.yq_ff_jacobian_dev(self._h, ns, d_sigma, d_jac, d_out))

    def ff_vjp_dev(self, ns, d_sigma, d_act, d_r, d_grad):
        self._chk(self._L.yq_ff_vjp_dev(self._h, ns, d_sigma, d_act, d_r, d_grad))

    def ff_vjp_dev_noalloc(self, ns, d_sigma, d_act, d_r, d_grad):
        """yq_ff_vjp_dev with its partial sums in the library's fixed device
        array (no device allocation; calls overlapping on one GPU race)."""
        self._chk(self._L.yq_ff_vjp_dev_noalloc(self._h, ns, d_sigma, d_act, d_r, d_grad))

    # AmplitudePhase machine: this context's real RBM (amplitude) and its
    # FeedForward network (phase, float64)
    def ap_enable(self, on=True):
        self._chk(self._L.yq_ap_enable(self._h, int(on)))
        self.ap = bool(on)

    def ap_dims(self):
        a, p = C.c_int(), C.c_int()
        self._chk(self._L.yq_ap_dims(self._h, C.byref(a), C.byref(p)))
        return a.value, p.value

    def ap_get_params(self):
        out = np.empty(self.sr_dim)
        self._chk(self._L.yq_ap_get_params(self._h, _ptr(out)))
        return out

    def ap_set_params(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.sr_dim,)
        self._chk(self._L.yq_ap_set_params(self._h, _ptr(v)))

    def ap_update_params(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        assert d.shape == (self.sr_dim,)
        self._chk(self._L.yq_ap_update_params(self._h, _ptr(d)))

    def ap_log_psi(self, sigma):
        sigma = np.ascontiguousarray(np.atleast_2d(sigma), dtype=np.int8)
        out = np.empty(sigma.shape[0], np.complex128)
        self._chk(self._L.yq_ap_log_psi(self._h, sigma.shape[0], _ptr(sigma), _ptr(out)))
        return out

    # many machines in one context
    def multi_create(self, machines, sweeper=LOCAL, precision=MIXED):
        """machines: list of dicts n, m, chains, seed, ham=(kind, kwargs) and
        optional use_bias (default True) and sweep_factor (default 1)."""
        arr = (MultiDesc * len(machines))()
        for q, mc in enumerate(machines):
            kind, hk = mc["ham"]
            hd = HamDesc(kind, mc["n"], hk.get("J", -1.0), hk.get("h", -1.0),
                         hk.get("delta", 1.0), hk.get("J2", 0.0), hk.get("sign_rule", 0))
            arr[q] = MultiDesc(mc["n"], mc["m"], int(mc.get("use_bias", True)), mc["chains"],
                               mc.get("sweep_factor", 1), mc["seed"], hd)
        self._chk(self._L.yq_multi_create(self._h, len(machines), arr, sweeper, precision))
        self.multi = [dict(mc) for mc in machines]
        for mc in self.multi:
            mc["dim"] = mc["n"] * mc["m"] + (mc["n"] + mc["m"] if mc.get("use_bias", True) else 0)

    def multi_set_params(self, idx, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == (self.multi[idx]["dim"],)
        self._chk(self._L.yq_multi_set_params(self._h, idx, _ptr(v)))

    def multi_get_params(self, idx):
        out = np.empty(self.multi[idx]["dim"])
        self._chk(self._L.yq_multi_get_params(self._h, idx, _ptr(out)))
        return out

    def multi_init_random(self, seed, sigma):
        self._chk(self._L.yq_multi_init_random(self._h, seed, sigma))

    def multi_sample(self, n_sweeps, n_discard):
        self._chk(self._L.yq_multi_sample(self._h, n_sweeps, n_discard))
        self.multi_sweeps = n_sweeps

    def multi_samples(self, idx):
        mc = self.multi[idx]
        ns = mc["chains"] * self.multi_sweeps
        s = np.empty((ns, mc["n"]), np.int8)
        t = np.empty((ns, mc["m"]))
        self._chk(self._L.yq_multi_samples_get(self._h, idx, _ptr(s), _ptr(t)))
        return s, t

    def multi_state(self, idx):
        mc = self.multi[idx]
        s = np.empty((mc["chains"], mc["n"]), np.int8)
        ctr = np.empty(mc["chains"], np.uint64)
        acc = np.empty(mc["chains"], np.uint64)
        self._chk(self._L.yq_multi_sampler_state(self._h, idx, _ptr(s), _ptr(ctr), _ptr(acc)))
        return s, ctr, acc

    def multi_eloc(self):
        k = len(self.multi)
        mean, var = np.empty(k), np.empty(k)
        self._chk(self._L.yq_multi_eloc(self._h, _ptr(mean), _ptr(var)))
        return mean, var

    def multi_set_grouped(self, on=True):
        """Grouped SR solves in multi_sr_step (one launch per kernel and CG
        iteration for all machines) on or off."""
        self._chk(self._L.yq_multi_set_grouped(self._h, 1 if on else 0))

    def multi_sr_step(self, n_threads=0, **kw):
        cfg = StepCfg(kw.get("n_sweeps", 1), kw.get("n_discard", 0), int(kw.get("randomize", 0)),
                      kw.get("nup", -1), kw.get("solver", SOLVE_CG_FACTORED),
                      kw.get("precision", MIXED), kw.get("lam", 1e-3), kw.get("cg_tol", 1e-8),
                      kw.get("cg_max_iter", 20000), kw.get("lr", 0.01), kw.get("grad_clip", 0.0),
                      int(kw.get("precond", 0)), int(kw.get("warm_start", 0)))
        outs = (StepOut * len(self.multi))()
        self._chk(self._L.yq_multi_sr_step(self._h, C.byref(cfg), outs, n_threads),
                  allow=(ERR_NOT_CONVERGED,) if kw.get("allow_unconverged") else ())
        self.multi_sweeps = kw.get("n_sweeps", 1)
        return [{f: getattr(o, f) for f, _ in StepOut._fields_} for o in outs]

    def sr_options(self, precond=0, warm_start=0):
        self._chk(self._L.yq_sr_set_options(self._h, precond, warm_start))

    def rbm_save(self, path, elem=None, pointer=True):
        """Write the machine as the reference's w%04d.dat files are written
        (elem defaults to RBM<double> / RBM<complex<double>>)."""
        if elem is None:
            elem = CKPT_C128 if self.cplx else CKPT_F64
        self._chk(self._L.yq_rbm_save(self._h, os.fsencode(path), elem, int(pointer)))

    def rbm_load(self, path, elem=CKPT_F64, pointer=True):
        """Make the machine stored in a reference checkpoint the context's."""
        self._chk(self._L.yq_rbm_load(self._h, os.fsencode(path), elem, int(pointer)))
        n, m, dim = C.c_int(), C.c_int(), C.c_int()
        self._chk(self._L.yq_rbm_dims(self._h, C.byref(n), C.byref(m), C.byref(dim)))
        self.dt = np.dtype(np.complex128 if elem in (CKPT_C64, CKPT_C128) else np.float64)
        self.n, self.m, self.dim = n.value, m.value, dim.value

    def graph_mode(self, use_graphs=True):
        """True / 1: captured CG chunks of 8 iterations; an even int: chunks
        of that many iterations; False / 0: eager launches."""
        self._chk(self._L.yq_set_graph_mode(self._h, int(use_graphs)))

    def autotune(self, on=True):
        self._chk(self._L.yq_set_autotune(self._h, int(on)))

    def tiles(self, col_tiles=0, chunk=0):
        self._chk(self._L.yq_set_tiles(self._h, col_tiles, chunk))

    def sr_step(self, **kw):
        cfg = StepCfg(kw.get("n_sweeps", 1), kw.get("n_discard", 0), int(kw.get("randomize", 0)),
                      kw.get("nup", -1), kw.get("solver", SOLVE_CG_FACTORED),
                      kw.get("precision", FP64), kw.get("lam", 1e-3), kw.get("cg_tol", 1e-8),
                      kw.get("cg_max_iter", 20000), kw.get("lr", 0.01), kw.get("grad_clip", 0.0),
                      int(kw.get("precond", 0)), int(kw.get("warm_start", 0)))
        out = StepOut()
        self._chk(self._L.yq_sr_step(self._h, C.byref(cfg), C.byref(out)),
                  allow=(ERR_NOT_CONVERGED,) if kw.get("allow_unconverged") else ())
        self.ns = out.n_samples
        return {f: getattr(out, f) for f, _ in StepOut._fields_}
