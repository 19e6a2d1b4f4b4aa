"""C-ABI checks that need no GPU: the libraries load, export every function the
public headers declare, the headers compile as C and C++, and an external
GRAPE-6 caller links against the library unchanged."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

import sapporo2_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = {
    "g5lib.h": sa.LIB_PATH,
    "g6lib.h": sa.LIB_PATH,
    "sapporo_amd.h": sa.LIB_PATH,
    "sapporo6thlib.h": sa.LIB6_PATH,
}


def declared_functions(header):
    txt = open(os.path.join(INC, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt))
    keywords = {"if", "defined", "sizeof", "return"}
    # keep only names that appear as declarations: "<type> name(" at line start
    decl = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                         txt, flags=re.M):
        decl.add(m.group(1))
    return sorted((names & decl) - keywords)


def exported(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib], text=True)
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_header_symbols_exported(header):
    lib = HEADERS[header]
    assert os.path.exists(lib), f"{lib} not built"
    names = declared_functions(header)
    assert len(names) >= 5, names
    syms = exported(lib)
    missing = []
    for n in names:
        # get_j_part_data / pick_up_predictor_2 keep C linkage in the reference too
        if n not in syms:
            missing.append(n)
    assert not missing, f"{os.path.basename(lib)} lacks {missing}"


def test_engine_symbol_list_matches_header():
    names = set(declared_functions("sapporo_amd.h"))
    assert set(sa.ENGINE_SYMBOLS) <= names


def test_libraries_load_without_gpu():
    for p in (sa.LIB_PATH, sa.LIB6_PATH):
        ctypes.CDLL(p)   # loading must not need a device


def test_headers_compile_as_c_and_cxx(tmp_path):
    src_c = tmp_path / "t.c"
    src_c.write_text('#include "g5lib.h"\n#include "g6lib.h"\n#include "sapporo_amd.h"\n#include "sapporo6thlib.h"\n'
                     "int main(void){ return 0; }\n")
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", INC, str(src_c)])
    src_cc = tmp_path / "t.cpp"
    src_cc.write_text(src_c.read_text())
    subprocess.check_call(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I", INC, str(src_cc)])


def test_external_g6_caller_links(tmp_path):
    """A GRAPE-6 Hermite integrator (tests/integrator/hermite4_g6.cpp) links unchanged."""
    exe = tmp_path / "h4"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", INC, "-o", str(exe),
                           os.path.join(ROOT, "tests", "integrator", "hermite4_g6.cpp"),
                           "-L", os.path.dirname(sa.LIB_PATH), "-lsapporo_amd"])
    assert exe.exists()


def test_open_without_device_fails_cleanly():
    """No device: open reports an error instead of falling back to the CPU."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        sa.Engine(sa.FOURTH, sa.DOUBLESINGLE)
