"""The header-only concept layer (include/yannq_amd/yannq_amd.hpp) compiles
against the C ABI, and (on a GPU) drives config C1 through RunRBM the way
Examples/RBMTFI.cpp drives the reference, and a complex RBM on the XXZ chain
the way Examples/RBMXXZ.cpp does."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _build(yq, tmp_path, name="run_tfi"):
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "cpp", name + ".cpp"), "-o", exe,
           "-L", os.path.dirname(yq.LIB_PATH), "-lyannq_amd",
           "-Wl,-rpath," + os.path.dirname(yq.LIB_PATH)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.parametrize("name", ["run_tfi", "run_xxz"])
def test_concept_layer_compiles(yq, tmp_path, name):
    assert os.path.exists(_build(yq, tmp_path, name))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "cg"])
def test_runner_reaches_ground_state(yq, tmp_path, mode):
    from tests.conftest import _gpu_available
    if not _gpu_available():
        pytest.skip("no GPU visible")
    exe = _build(yq, tmp_path)
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_complex_runner_reaches_ground_state(yq, tmp_path):
    from tests.conftest import _gpu_available
    if not _gpu_available():
        pytest.skip("no GPU visible")
    exe = _build(yq, tmp_path, "run_xxz")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
