"""The C-ABI library loads and exports every function include/*.h declares
(no GPU work: loading only).  The CPU checker exports its own
header's functions too."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_timestep_surface():
    names = declared("include/fs_abi.h", "fs_")
    for required in ("fs_create", "fs_destroy", "fs_set_fields", "fs_get_fields", "fs_advance_frame", "fs_step",
                     "fs_cfl_dt", "fs_projection", "fs_reinitialization", "fs_levelset_advection",
                     "fs_velocity_advection", "fs_sourcing", "fs_apply_gravity", "fs_last_error"):
        assert required in names


def test_hip_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "fluid_simulator_backup_amd", "libfs_hip.so")
    if not os.path.exists(path):
        pytest.fail("libfs_hip.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert "fs_abi.h" in headers and "fs_world.h" in headers
    for h in headers:
        missing = [n for n in declared("include/" + h, "fs_") if not hasattr(lib, n)]
        assert not missing, (h, missing)


def test_oracle_exports_its_header(oracle_lib):
    missing = [n for n in declared("oracle/fs_oracle.h", "fso_") if not hasattr(oracle_lib, n)]
    assert not missing, missing


def test_structs_match_header_sizes():
    """The ctypes mirror has the same field list as the C structs."""
    from fluid_simulator_backup_amd import _abi
    text = open(os.path.join(ROOT, "include/fs_abi.h")).read()
    body = re.search(r"typedef struct fs_params \{(.*?)\} fs_params;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\b([a-z_0-9]+)(?:\[3\])?\s*[;,]", body)
    assert fields == [f[0] for f in _abi.FsParams._fields_]
