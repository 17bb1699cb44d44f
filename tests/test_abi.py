"""CPU-side checks of the C ABI: the library loads, exports every entry point
that include/yannq_amd.h declares, and reports errors without a GPU."""
import ctypes as C
import subprocess

import pytest


def test_exports_every_declared_symbol(yq):
    declared = set(yq.header_symbols())
    assert len(declared) >= 24
    out = subprocess.run(["nm", "-D", "--defined-only", yq.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = declared - exported
    assert not missing, missing
    # nothing but the ABI is exported
    assert {s for s in exported if s.startswith("yq_")} == declared


def test_null_context_is_rejected(yq):
    L = yq.lib()
    assert L.yq_ctx_create(0, None) == yq.ERR_ARG
    assert L.yq_ctx_destroy(None) == yq.OK
    assert L.yq_rbm_create(None, 1, 4, 4, 1) == yq.ERR_ARG
    assert L.yq_last_error(None) == b"null context"


def test_no_gpu_is_an_error_not_a_crash(yq):
    from tests.conftest import _gpu_available
    if _gpu_available():
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert yq.lib().yq_ctx_create(0, C.byref(h)) != yq.OK


def test_chain_lanes_rule(yq):
    """One wave up to 512 hidden units, then the smallest power-of-two group
    with at most 8 units per thread (a host-side rule, no GPU call)."""
    L = yq.lib()
    want = {1: 64, 400: 64, 512: 64, 513: 128, 1024: 128, 1025: 256, 2048: 256,
            2100: 512, 4096: 512, 4097: 1024, 8192: 1024}
    for m, g in want.items():
        assert L.yq_chain_lanes(m) == g, m
        assert g * 8 >= m and (g == 64 or g * 4 < m)
