"""Checkpoint wire format (host only, no GPU): the library's encoder and
decoder against byte-exact fixtures of the reference's weight files (cereal
BinaryArchive of RBM<T>: Serializers/SerializeRBM.hpp:21-79,
SerializeEigen.hpp:10-35; written through a unique_ptr by
Runners/RunRBM.hpp:155-165, read by AbstractRunner.hpp:83-94).  The fixtures
(tests/golden/ckpt_rbm_n2_m3.json) are built field by field by
tests/golden/make_ckpt_golden.py and list every field's offset."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ELEM = {"f32": 1, "f64": 2, "c64": 3, "c128": 4}


def _cases():
    with open(os.path.join(HERE, "golden", "ckpt_rbm_n2_m3.json")) as f:
        return json.load(f)["cases"]


def _params(c):
    re = np.array(c["params_re"])
    return re + 1j * np.array(c["params_im"]) if c["params_im"] is not None else re


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"{c['elem']}-bias{int(c['use_bias'])}-ptr{int(c['pointer'])}")
def test_encode_decode_fixture(yq, case):
    e = ELEM[case["elem"]]
    want = bytes.fromhex(case["hex"])
    got = yq.ckpt_encode(_params(case), case["n"], case["m"], case["use_bias"], e, case["pointer"])
    assert got == want
    n, m, ub, p = yq.ckpt_decode(want, e, case["pointer"])
    assert (n, m, ub) == (case["n"], case["m"], case["use_bias"])
    np.testing.assert_array_equal(p, _params(case))
    # every field is accounted for and the lengths add up
    lay = case["layout"]
    assert lay[0][0] == 0 and sum(l[1] for l in lay) == len(want)
    assert all(a[0] + a[1] == b[0] for a, b in zip(lay, lay[1:]))


def test_sizes_by_hand(yq):
    L = yq.lib()
    # N = 100, M = 400, RBM<double> with bias through a unique_ptr:
    # 1 + 4 + 1 + 4 + 4 + (8 + 320000) + (8 + 800) + (8 + 3200)
    assert L.yq_ckpt_rbm_bytes(yq.CKPT_F64, 100, 400, 1, 1) == 324038
    # RBM<complex<double>>, N = 64, M = 128, no bias, bare object
    assert L.yq_ckpt_rbm_bytes(yq.CKPT_C128, 64, 128, 0, 0) == 4 + 1 + 4 + 4 + 8 + 64 * 128 * 16
    assert L.yq_ckpt_rbm_bytes(7, 4, 4, 1, 1) == 0


@pytest.mark.parametrize("elem", [1, 2, 3, 4])
@pytest.mark.parametrize("use_bias", [True, False])
def test_round_trip(yq, elem, use_bias):
    rng = np.random.default_rng(elem + 10 * use_bias)
    n, m = 7, 13
    dim = n * m + (n + m if use_bias else 0)
    cplx = elem in (3, 4)
    p = rng.normal(size=dim) + (1j * rng.normal(size=dim) if cplx else 0)
    buf = yq.ckpt_encode(p, n, m, use_bias, elem, True)
    n2, m2, ub2, q = yq.ckpt_decode(buf, elem, True)
    assert (n2, m2, ub2) == (n, m, use_bias)
    if elem in (1, 3):   # float files hold the parameters rounded to float
        ref = p.astype(np.complex64 if cplx else np.float32)
        np.testing.assert_array_equal(q, ref.astype(q.dtype))
    else:
        np.testing.assert_array_equal(q, p)


def test_malformed_files_are_rejected(yq):
    p = np.arange(2 * 3 + 2 + 3, dtype=float)
    good = yq.ckpt_encode(p, 2, 3, True, yq.CKPT_F64, True)
    yq.ckpt_decode(good, yq.CKPT_F64, True)
    bad = [
        good[:-1],                                 # truncated
        good + b"\0",                              # trailing byte
        b"\0" + good[1:],                          # null unique_ptr
        good[:5] + b"\x02" + good[6:],             # useBias not a bool
    ]
    for b in bad:
        with pytest.raises(yq.YqError):
            yq.ckpt_decode(b, yq.CKPT_F64, True)
    with pytest.raises(yq.YqError):                # another scalar type: length differs
        yq.ckpt_decode(good, yq.CKPT_C128, True)
    with pytest.raises(yq.YqError):                # bare object read as unique_ptr form
        yq.ckpt_decode(yq.ckpt_encode(p, 2, 3, True, yq.CKPT_F64, False), yq.CKPT_F64, True)
    with pytest.raises(yq.YqError):                # W shape disagreeing with N, M
        yq.ckpt_decode(good[:14] + (4).to_bytes(4, "little") + good[18:], yq.CKPT_F64, True)
