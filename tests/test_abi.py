"""The C-ABI library loads and exports every entry point include/nnpx.h declares
(CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "nnpx.h")).read()
    return sorted(set(re.findall(r"\b(nnpx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("nnpx_create", "nnpx_load_model", "nnpx_set_structures", "nnpx_calculate_symmetry_function_groups",
              "nnpx_calculate_atomic_neural_networks", "nnpx_calculate_energy", "nnpx_calculate_forces",
              "nnpx_evaluate", "nnpx_get_forces"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import n2p2_amd
    if not os.path.exists(n2p2_amd.LIB_PATH):
        pytest.fail("libnnpx.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(n2p2_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_package_fails_loudly_without_library(tmp_path, monkeypatch):
    import n2p2_amd
    monkeypatch.setattr(n2p2_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(n2p2_amd, "_lib", None)
    with pytest.raises(n2p2_amd.NNPError):
        n2p2_amd.library()


def test_remap_matches_reference_rules():
    """Structure::remap: fractional coordinates wrapped into [0, 1)."""
    import numpy as np
    from n2p2_amd import remap
    box = np.array([[10.0, 0, 0], [1.0, 9.0, 0], [0.5, 0.5, 8.0]])
    rng = np.random.default_rng(1)
    pos = rng.uniform(-25, 25, (200, 3))
    r = remap(pos, box)
    f = r @ np.linalg.inv(box)
    assert (f >= -1e-12).all() and (f < 1 + 1e-12).all()
    d = (pos - r) @ np.linalg.inv(box)
    assert np.abs(d - np.round(d)).max() < 1e-9
