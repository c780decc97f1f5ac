"""The C-ABI library loads on a machine without a GPU and exports every entry point the header declares."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sbp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sbp_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "sbp_partition" in names and "sbp_mdl" in names and "sbp_merge_blocks" in names


def test_library_exports_every_declared_symbol():
    from sbp_amd import native
    if not os.path.exists(native.LIB_PATH):
        pytest.fail("libsbp_amd.so not built (run __graft_entry__.build())")
    L = native.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(native.EXPORTS) == declared_functions()


def test_native_has_no_fallback_when_library_missing(monkeypatch, tmp_path):
    from sbp_amd import native
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(native.SbpError):
        native.lib()
