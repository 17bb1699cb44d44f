"""The replayed kernels contain only instructions whose results the CPU replay
reproduces.

The sampler's accept/reject decisions and the real E_loc are bit-exact
against a CPU replay of the same IEEE operations.  That holds only if the
compiled kernels use no instruction with a context-dependent or approximate
result: packed FP32 arithmetic (disabled at build time), the hardware
transcendental approximations (v_exp/v_log/v_sin/v_cos/v_rsq/v_sqrt), mixed-
precision FMA forms, or denormal / rounding-mode switches; and every float
reciprocal must belong to the correctly rounded divide sequence
(v_div_scale / v_div_fmas / v_div_fixup) or to an integer division.  This
test compiles the sources to gfx950 assembly with the library's flags and
checks every sampler and E_loc ratio kernel.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "yannq_amd", "csrc")
BANNED = re.compile(r"^\s+(v_pk_(?:fma|mul|add)_f32|v_exp_f\d+|v_log_f\d+|v_sin_f\d+|v_cos_f\d+|"
                    r"v_rsq_f\d+|v_sqrt_f\d+|v_fma_mix\w*|v_mad_mix\w*|s_denorm_mode|"
                    r"s_round_mode)(?:_e32|_e64|_dpp|_sdwa)?\s", re.M)

pytestmark = [pytest.mark.slow,
              pytest.mark.skipif(shutil.which("hipcc") is None and
                                 not os.path.exists("/opt/rocm/bin/hipcc"),
                                 reason="hipcc not available")]


def _asm(src: str, out_dir: str) -> str:
    from yannq_amd import build
    flags = [f for f in build.CFLAGS if f not in ("-fPIC", "-fvisibility=hidden")]
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    subprocess.run([build.hipcc(), *flags, "--cuda-device-only", "-S", src, "-o", out],
                   check=True, capture_output=True, text=True)
    return open(out).read()


def _kernels(asm: str):
    lines = asm.split("\n")
    starts = [(i, ln.split(":")[0]) for i, ln in enumerate(lines)
              if re.match(r"^_Z[A-Za-z0-9_]*:", ln)]
    starts.append((len(lines), None))
    for (i, name), (j, _) in zip(starts, starts[1:]):
        yield name, "\n".join(lines[i:j])


def test_replayed_kernels_use_only_replayable_instructions():
    with tempfile.TemporaryDirectory() as d:
        with cf.ThreadPoolExecutor(2) as ex:
            asms = list(ex.map(lambda f: _asm(os.path.join(CSRC, f), d),
                               ["sampler.hip", "eloc.hip"]))
    checked, bad = 0, []
    for asm in asms:
        for name, body in _kernels(asm):
            if not re.search(r"k_sample|k_eloc_ratio", name):
                continue
            checked += 1
            hits = sorted(set(BANNED.findall(body)))
            if hits:
                bad.append(f"{name}: {hits}")
            # every float reciprocal is part of a correctly rounded divide or
            # of an integer division (v_cvt_f32_u32 / v_rcp_iflag_f32)
            rcp = len(re.findall(r"^\s+v_rcp_f32", body, re.M))
            fix = len(re.findall(r"^\s+v_div_fixup_f32", body, re.M))
            cvt = len(re.findall(r"^\s+v_cvt_f32_u32", body, re.M))
            if rcp > fix + cvt:
                bad.append(f"{name}: {rcp} v_rcp_f32 against {fix} divides, {cvt} int divisions")
    assert checked >= 40, checked
    assert not bad, "\n".join(bad)
