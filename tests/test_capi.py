"""The C-ABI library loads and exports every entry point include/xmatch.h declares
(no GPU needed; nothing here launches device work)."""
import ctypes
import os
import re

import pytest

import xmatch_amd as xm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(xm_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_matches_binding_list():
    assert _declared() == set(xm.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = xm.lib()
    for name in sorted(_declared()):
        assert hasattr(L, name), name
    assert L.xm_abi_version() == 1


def test_nm_dynamic_symbols():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", xm.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = set(re.findall(r"\b(xm_[a-z0-9_]+)\b", out))
    assert _declared() <= syms


def test_null_arguments_are_rejected():
    L = xm.lib()
    assert L.xm_ctx_create(0, None) == xm.XM_EINVAL
    assert L.xm_match(None, None, None, 1.0, 0, 0, None, None) == xm.XM_EINVAL
    assert L.xm_catalog_info(None, None, None, None, None) == xm.XM_EINVAL
    assert L.xm_result_count(None, None) == xm.XM_EINVAL
    assert L.xm_last_error_message(None) is not None


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(xm, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(xm, "_lib", None)
    with pytest.raises(ImportError):
        xm.lib()
