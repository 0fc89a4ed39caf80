"""The Fortran module fe-project_amd/fortran/fedg_c_binding.F90 declares a
bind(C) interface for every entry point of include/fedg.h, and the module, the
cal_tend_ex wrapper and the harness compile with the image's amdflang (the
build step compiles them; this checks what it produced)."""
import os
import re
import shutil

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fedg.h")
MODULE = os.path.join(ROOT, "fe-project_amd", "fortran", "fedg_c_binding.F90")
HARNESS = os.path.join(ROOT, "fe-project_amd", "fortran", "fedg_dropin_harness")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fedg_[a-z0-9_]+)\s*\(", text))


def module_bindings():
    text = open(MODULE).read()
    return re.findall(r'bind\(C,\s*name="(fedg_[a-z0-9_]+)"\)', text)


def test_every_exported_function_has_a_fortran_interface():
    h = header_functions()
    b = module_bindings()
    assert len(h) > 70
    assert len(b) == len(set(b)), "an entry point is declared twice"
    assert set(b) == h, f"missing in Fortran: {sorted(h - set(b))}; unknown: {sorted(set(b) - h)}"


def test_public_list_names_every_interface():
    text = open(MODULE).read()
    public = set(re.findall(r"\bfedg_[a-z0-9_]+", " ".join(re.findall(r"^\s*public ::(.*)$", text, flags=re.M))))
    assert set(module_bindings()) <= public


@pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"),
                    reason="no Fortran compiler in this image")
def test_harness_was_built():
    assert os.access(HARNESS, os.X_OK)
