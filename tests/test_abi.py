"""CPU checks of the C-ABI library: it builds, loads, and exports every symbol
include/orbit_hip.h declares (no compute calls; this container has no GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(orbit_[A-Za-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_the_binding_list():
    from orbit_amd import _lib
    assert _declared() == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from orbit_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liborbit_hip.so not built")
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (orbit_[A-Za-z0-9_]+)", out))
    missing = _declared() - exported
    assert not missing, missing
    h = _lib.lib()
    for name in _lib.EXPORTS:
        assert hasattr(h, name)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from orbit_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.OrbitLibraryMissing):
        _lib.lib()


def test_oracle_is_not_imported_by_the_product():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "orbit_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith(".py"):
                txt = open(os.path.join(dirpath, fn)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle\b", txt, flags=re.M), fn


def _param_counts():
    """Parameter count of every function include/orbit_hip.h declares."""
    out = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(orbit_[A-Za-z0-9_]+)\s*\(([^;{]*?)\)\s*;", txt, flags=re.S):
            params = m.group(2).strip()
            out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_bindings_pass_as_many_arguments_as_the_header_declares():
    from orbit_amd import _lib
    counts = _param_counts()
    sig = _lib.signatures()
    assert set(sig) == set(counts)
    bad = {k: (len(sig[k][1]), counts[k]) for k in sig if len(sig[k][1]) != counts[k]}
    assert not bad, bad
