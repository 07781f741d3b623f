import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_LIB = os.path.join(ROOT, "oracle", "libfs_oracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfs_hip.so)")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    from fluid_simulator_backup_amd import _abi
    if not os.path.exists(ORACLE_LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = _abi.load(ORACLE_LIB, "fso_")
    # the checker's sums are per plane, added in plane order: its results do
    # not depend on the thread count (at most the host's 16-core share)
    lib.fso_set_threads(min(os.cpu_count() or 1, 16))
    return lib


@pytest.fixture(scope="session")
def oracle_ad():
    """The CPU checker's reverse-mode build (oracle/fs_oracle_ad.cpp)."""
    import ctypes as C
    path = os.path.join(ROOT, "oracle", "libfs_oracle_ad.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = C.CDLL(path)
    vp, pd, i = C.c_void_p, C.POINTER(C.c_double), C.c_int
    lib.fsa_rollout_gradient.argtypes = [vp, vp, pd, pd, pd, pd, pd, i, i, C.c_double, pd, pd, pd, pd, pd, pd,
                                         C.POINTER(C.c_long)]
    lib.fsa_rollout_gradient.restype = i
    lib.fsa_rollout_values.argtypes = [vp, vp, pd, pd, pd, pd, pd, i, i, C.c_double]
    lib.fsa_rollout_values.restype = i
    return lib


@pytest.fixture(scope="session")
def hip_lib():
    if not _gpu_available():
        pytest.skip("no GPU visible")
    import fluid_simulator_backup_amd as fsa
    return fsa.load_library()  # raises if the HIP build is missing: no fallback


@pytest.fixture(scope="session")
def host_lib():
    """The HIP library for its host-only entry points (scripts, scene
    parsing): loading it touches no GPU."""
    import fluid_simulator_backup_amd as fsa
    return fsa.load_library()
