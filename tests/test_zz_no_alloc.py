"""yq_ff_vjp_dev_noalloc, the allocation-free FeedForward VJP, against
yq_ff_vjp_dev through the C ABI on device buffers: two calls in a row, the
second with a smaller batch, in float at the C4 network shape and in double;
with the same block count the two entry points are bit-identical, with fewer
blocks (the fixed array is the limit) they agree to float rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sig(rng, ns, n):
    s = np.ones((ns, n), np.int8)
    for q in range(ns):
        s[q, rng.choice(n, n // 2, replace=False)] = -1
    return s


@pytest.mark.parametrize("dtype,shape,sizes", [
    (np.float32, (40, 8, 5, 64), (8192, 1000)),
    (np.float64, (16, 3, 3, 12), (3000, 97)),
])
def test_noalloc_vjp_matches_vjp(ctx, dtype, shape, sizes):
    torch = pytest.importorskip("torch")
    n, C_, K, H = shape
    ctx.ff_create(n, C_, K, H, dtype=dtype)
    ctx.ff_init_random(11, 3)
    npar = ctx.ff_npar
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    rng = np.random.default_rng(5)
    row = npar * np.dtype(dtype).itemsize
    rows = (16 << 20) // row
    for ns in sizes:
        sig = torch.tensor(_sig(rng, ns, n), device="cuda:0")
        r = torch.tensor(rng.normal(size=ns), dtype=tdt, device="cuda:0")
        logpsi = torch.empty(ns, dtype=tdt, device="cuda:0")
        act = torch.empty((ns, 2 * C_ * n + H), dtype=tdt, device="cuda:0")
        torch.cuda.synchronize()
        ctx.ff_forward_dev(ns, sig.data_ptr(), logpsi.data_ptr(), act.data_ptr())
        for a in (act, None):
            g_ref = torch.full((npar,), float("nan"), dtype=tdt, device="cuda:0")
            g = torch.full((npar,), float("nan"), dtype=tdt, device="cuda:0")
            torch.cuda.synchronize()
            ap = a.data_ptr() if a is not None else None
            ctx.ff_vjp_dev(ns, sig.data_ptr(), ap, r.data_ptr(), g_ref.data_ptr())
            ctx.ff_vjp_dev_noalloc(ns, sig.data_ptr(), ap, r.data_ptr(), g.data_ptr())
            torch.cuda.synchronize()
            g_ref, g = g_ref.cpu().numpy(), g.cpu().numpy()
            assert np.isfinite(g).all()
            groups = -(-ns // (8 if dtype == np.float32 else 4))
            if min(groups, 512) <= rows:   # same grid: the same sums in the same order
                np.testing.assert_array_equal(g, g_ref)
            else:
                tol = 2e-5 if dtype == np.float32 else 1e-12
                assert np.linalg.norm(g - g_ref) <= tol * np.linalg.norm(g_ref)
