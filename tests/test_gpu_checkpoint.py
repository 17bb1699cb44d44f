"""Checkpoints on the GPU: a machine written with yq_rbm_save and read back
with yq_rbm_load (the reference's w%04d.dat format) holds the same
parameters, and samples bit-identically to the original from the same seed;
float and complex element types round-trip as the format stores them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_bias", [True, False])
def test_save_load_round_trip(yq, ctx, tmp_path, use_bias):
    n, m = 24, 48
    ctx.rbm_create(n, m, use_bias=use_bias)
    ctx.init_random(5, 0.3)
    p = ctx.get_params()
    path = tmp_path / "w0000.dat"
    ctx.rbm_save(path)
    raw = path.read_bytes()
    assert len(raw) == yq.lib().yq_ckpt_rbm_bytes(yq.CKPT_F64, n, m, int(use_bias), 1)
    n2, m2, ub2, q = yq.ckpt_decode(raw, yq.CKPT_F64, True)
    assert (n2, m2, ub2) == (n, m, use_bias)
    np.testing.assert_array_equal(q, p)
    with yq.Context(0) as c2:
        c2.rbm_load(path)
        assert (c2.n, c2.m, c2.dim) == (n, m, ctx.dim)
        np.testing.assert_array_equal(c2.get_params(), p)
        # the reloaded machine samples exactly like the original
        for c in (ctx, c2):
            c.ham_set(yq.TFI, n, J=-1.0, h=-1.0)
            c.sampler_config(yq.LOCAL, 32, precision=yq.FP64, seed=11)
            c.sample(2, 3)
        s1, t1 = ctx.samples()
        s2, t2 = c2.samples()
        np.testing.assert_array_equal(s1, s2)
        np.testing.assert_array_equal(t1, t2)


def test_float_and_complex_files(yq, ctx, tmp_path):
    n, m = 8, 16
    ctx.rbm_create(n, m, dtype=np.complex128)
    ctx.init_random(9, 0.2)
    p = ctx.get_params()
    for elem, dt in ((yq.CKPT_C128, np.complex128), (yq.CKPT_C64, np.complex64)):
        path = tmp_path / f"c{elem}.dat"
        ctx.rbm_save(path, elem=elem)
        with yq.Context(0) as c2:
            c2.rbm_load(path, elem=elem)
            np.testing.assert_array_equal(c2.get_params(), p.astype(dt).astype(np.complex128))
    with pytest.raises(yq.YqError):   # a complex machine cannot be written as RBM<double>
        ctx.rbm_save(tmp_path / "x.dat", elem=yq.CKPT_F64)
    with pytest.raises(yq.YqError):   # nor read as another scalar type
        ctx.rbm_load(tmp_path / f"c{yq.CKPT_C128}.dat", elem=yq.CKPT_F64)


def _golden_cases():
    import json
    import os
    p = os.path.join(os.path.dirname(__file__), "golden", "ckpt_rbm_n2_m3.json")
    with open(p) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden_cases(),
                         ids=lambda c: f"{c['elem']}-bias{int(c['use_bias'])}-ptr{int(c['pointer'])}")
def test_device_machine_writes_the_golden_bytes(yq, ctx, tmp_path, case):
    """A machine whose parameters are set on the GPU and written by
    yq_rbm_save produces exactly the bytes of the committed fixture (every
    field at its stated offset, built by tests/golden/make_ckpt_golden.py from
    cereal's published binary format), for the four element types, with and
    without bias, through a unique_ptr and as the bare object; loading the
    fixture gives the machine back."""
    elem = {"f32": yq.CKPT_F32, "f64": yq.CKPT_F64, "c64": yq.CKPT_C64,
            "c128": yq.CKPT_C128}[case["elem"]]
    cplx = case["params_im"] is not None
    p = np.array(case["params_re"], dtype=np.float64)
    if cplx:
        p = p + 1j * np.array(case["params_im"], dtype=np.float64)
    ctx.rbm_create(case["n"], case["m"], use_bias=case["use_bias"],
                   dtype=np.complex128 if cplx else np.float64)
    ctx.set_params(p)
    path = tmp_path / "w.dat"
    ctx.rbm_save(path, elem=elem, pointer=case["pointer"])
    assert path.read_bytes().hex() == case["hex"]
    gold = tmp_path / "gold.dat"
    gold.write_bytes(bytes.fromhex(case["hex"]))
    with yq.Context(0) as c2:
        c2.rbm_load(gold, elem=elem, pointer=case["pointer"])
        assert (c2.n, c2.m) == (case["n"], case["m"])
        np.testing.assert_array_equal(c2.get_params(), p)
