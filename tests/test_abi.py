"""C-ABI checks that need no GPU: the library loads, exports every function
include/swinmd_hip.h declares, and the Python declarations cover them."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.helpers import ROOT

HEADER = os.path.join(ROOT, "include", "swinmd_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smd_[a-z0-9_]+)\s*\(", text)))


def _lib_path():
    from swinmd_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libswinmd_hip.so not built (python -m swinmd_amd.build)")
    return _lib.LIB_PATH


def test_library_exports_every_declared_symbol():
    path = _lib_path()
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(smd_[a-z0-9_]+)", out))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing


def test_library_loads_and_python_signatures_match_header():
    from swinmd_amd import _lib
    _lib_path()
    lib = _lib.load()
    assert lib.smd_abi_version() == 1
    assert sorted(_lib.SIGNATURES) == _declared()
    assert lib.smd_strerror(-5) == b"pair list capacity exceeded"


def test_ctx_create_rejects_bad_topology_without_touching_the_gpu():
    """Topology validation runs before any device call."""
    import numpy as np

    from swinmd_amd import _lib
    _lib_path()
    lib = _lib.load()
    t = np.zeros(3, dtype=np.int32)
    mol = np.array([0, 2, 1], dtype=np.int32)      # not contiguous/non-decreasing
    q = np.zeros(3)
    m = np.ones(3)
    eo = np.zeros(4, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    topo = _lib.Topology(3, 3, 1, p(t), p(mol), p(q), p(m), p(eo), None)
    h = ctypes.c_void_p()
    assert lib.smd_ctx_create(ctypes.byref(h), ctypes.byref(topo), 0) == _lib.SMD_EINVAL
    assert not h.value


def test_header_entry_points_cite_the_reference():
    text = open(HEADER).read()
    assert "src/Force.h:30-72" in text and "src/forces/EwaldForce.cpp" in text
