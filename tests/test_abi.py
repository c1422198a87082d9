"""The C-ABI library loads on a CPU-only host and exports every entry point
that include/dcm_hip.h declares; the ctypes binding declares the same set.
No compute call is made here (there is no GPU)."""
import ctypes
import os
import re

import pytest

from dcm_amd import hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dcm_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dcm_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    assert declared() == H.declared_symbols()


@pytest.mark.skipif(not os.path.exists(H.LIB_PATH), reason="libdcm_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(H.LIB_PATH)
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(H.LIB_PATH), reason="libdcm_hip.so not built")
def test_version_string():
    assert H.load().dcm_version().decode().startswith("dcm_hip")


def test_fortran_bindings_name_declared_entry_points():
    """Every bind(c, name=...) of the Fortran binding module names a function
    that the header declares (a typo would only show at link time of a
    model build)."""
    src = open(os.path.join(ROOT, "dynamical-cloud-model_amd", "fortran", "dcm_hip.f90")).read()
    bound = set(re.findall(r'bind\(c,\s*name="(dcm_[a-z0-9_]+)"\)', src, flags=re.I))
    assert bound, "no bindings found"
    assert not bound - set(declared()), sorted(bound - set(declared()))


def test_fortran_bindings_are_public():
    """Every C function the Fortran binding module declares is in its public
    list, so that a model can call it after `use dcm_hip`."""
    src = open(os.path.join(ROOT, "dynamical-cloud-model_amd", "fortran", "dcm_hip.f90")).read()
    bound = set(re.findall(r"function\s+(dcm_[a-z0-9_]+)\s*\([^)]*\)\s*&?\s*(?:\n\s*)?[^\n]*bind\(c",
                           src, flags=re.I))
    bound |= set(re.findall(r'bind\(c,\s*name="(dcm_[a-z0-9_]+)"\)', src, flags=re.I))
    public = set()
    for stmt in re.findall(r"public\s*::(.*?)(?<!&)\n", src.replace("&\n", " "), flags=re.I):
        public |= {n.strip() for n in stmt.split(",")}
    assert not bound - public, sorted(bound - public)
