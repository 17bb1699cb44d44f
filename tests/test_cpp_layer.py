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


@pytest.mark.parametrize("name", ["run_tfi", "run_xxz", "run_ckpt", "run_exact", "run_ap"])
def test_concept_layer_compiles(yq, tmp_path, name):
    assert os.path.exists(_build(yq, tmp_path, name))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "cg", "fisher"])
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


@pytest.mark.gpu
def test_runner_checkpoints(yq, tmp_path):
    """w%04d.dat files written by RunRBM decode as the reference's format and
    read back through initializeFrom."""
    import numpy as np
    from tests.conftest import _gpu_available
    if not _gpu_available():
        pytest.skip("no GPU visible")
    exe = _build(yq, tmp_path, "run_ckpt")
    work = tmp_path / "run"
    work.mkdir()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, cwd=str(work))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    n, m = 20, 20
    sizes = {f: (work / f).stat().st_size for f in ("w0000.dat", "w0010.dat", "w0020.dat")}
    assert set(sizes.values()) == {yq.lib().yq_ckpt_rbm_bytes(yq.CKPT_F64, n, m, 1, 1)}
    decoded = [yq.ckpt_decode((work / f).read_bytes(), yq.CKPT_F64, True) for f in sorted(sizes)]
    assert all(d[:3] == (n, m, True) for d in decoded)
    # initializeRandom(0.01) at epoch 0, then SGD steps move the parameters
    assert np.std(decoded[0][3]) < 0.05 and not np.array_equal(decoded[0][3], decoded[2][3])


@pytest.mark.gpu
def test_exact_runner_reaches_ground_state(yq, tmp_path):
    """RunRBMExact<complex<double>> over BasisFull as Examples/RBMTFIExact.cpp."""
    from tests.conftest import _gpu_available
    if not _gpu_available():
        pytest.skip("no GPU visible")
    exe = _build(yq, tmp_path, "run_exact")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["ngd", "srex", "example"])
def test_amplitude_phase_runners(yq, orc, tmp_path, mode):
    """RunAmplitudePhaseExact and the SRMatExact<AmplitudePhase> loop on the
    J1-J2 chain (N = 10, J2 = 0.2, sign rule) in the Sz = 0 sector, against
    the sector's ground energy from exact diagonalisation; "example" runs
    RunAmplitudePhaseExact with Examples/AmplitudePhase.cpp's phase network."""
    import numpy as np
    import scipy.sparse.linalg as sla
    from tests.conftest import _gpu_available
    if not _gpu_available():
        pytest.skip("no GPU visible")
    n = 10
    H = orc.ham_matrix(orc.ham(orc.J1J2, n, J=1.0, J2=0.2, sign_rule=1))
    basis = orc.basis_jz(n, n // 2)
    e0 = float(sla.eigsh(H[basis][:, basis].astype(float), k=1, which="SA")[0][0])
    exe = _build(yq, tmp_path, "run_ap")
    r = subprocess.run([exe, mode, repr(e0)], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
