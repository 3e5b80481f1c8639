"""The C-ABI library loads and exports every symbol include/svdag.h declares."""
import ctypes
import os
import re

import svdag_py

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "svdag.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(svdag_[a-z0-9_]+)\s*\(", text)))


def test_header_parses():
    names = declared_functions()
    assert "svdag_edit" in names and "svdag_hashset_find_dev" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = svdag_py.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_declared_symbol():
    assert sorted(svdag_py.EXPORTED) == declared_functions()


def test_abi_version_and_error_string():
    lib = svdag_py.lib()
    assert lib.svdag_abi_version() == 2
    assert isinstance(lib.svdag_last_error(), bytes)


def test_invalid_arguments_rejected_without_gpu():
    lib = svdag_py.lib()
    # NULL out pointers are rejected before any device call.
    assert lib.svdag_device_count(None) == -1
    assert lib.svdag_hashset_create(0, 0, None, None, None) == -1
    assert lib.svdag_create(None, None) == -1
