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
