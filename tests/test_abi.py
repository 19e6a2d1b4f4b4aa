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
    """Functions the header declares for the library to export (static inline
    caller-side definitions excluded: they live in the header)."""
    txt = open(os.path.join(INC, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = "\n".join(line for line in txt.splitlines() if not line.lstrip().startswith("static inline"))
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


# The caller side of the reference header (testCodes/src/g6lib.h:52-90).
G6_CALLER_SIDE = ["g6_close", "g6_open", "g6_npipes", "g6_set_ti", "g6_reset", "g6_reset_fofpga",
                  "g6_set_j_particle", "g6calc_firsthalf", "g6calc_lasthalf", "g6_set_tunit", "g6_set_xunit",
                  "g6_set_j_particle_multisend_mxfast_", "g6_reinitialize", "g6_print_chip_status",
                  "g6_flush_jp_buffer_and_multisend", "g6_initialize_jp_buffer",
                  "g6_set_overflow_flag_test_mode_"]


def test_g6_header_has_reference_caller_side():
    txt = open(os.path.join(INC, "g6lib.h")).read()
    defined = set(re.findall(r"static inline [A-Za-z_ ]+?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt))
    missing = [n for n in G6_CALLER_SIDE if n not in defined]
    assert not missing, f"g6lib.h lacks the caller-side definitions {missing}"


def test_caller_side_wrappers_link_from_two_units(tmp_path):
    """A C and a C++ translation unit of one program both include g6lib.h and call the
    by-value wrappers; the program links against the library (the header's definitions
    are static inline, so they do not collide)."""
    (tmp_path / "a.c").write_text(
        '#include "g6lib.h"\n'
        "int unit_a(void) {\n"
        "  double x[1][3] = {{0, 0, 0}}, v[1][3] = {{0, 0, 0}}, a[1][3], j[1][3], pot[1], h2[1] = {0};\n"
        "  int id[1] = {0};\n"
        "  g6_set_ti(0, 0.0);\n"
        "  g6calc_firsthalf(0, 1, 1, id, x, v, a, j, pot, 1e-4, h2);\n"
        "  return g6calc_lasthalf(0, 1, 1, id, x, v, 1e-4, h2, a, j, pot) + g6_npipes();\n"
        "}\n")
    (tmp_path / "b.cpp").write_text(
        '#include "g6lib.h"\n'
        'extern "C" int unit_a(void);\n'
        "int main() {\n"
        "  double m = 1, x[3] = {0, 0, 0}, z[3] = {0, 0, 0};\n"
        "  int cl = 0, n = 1, addr = 0, id = 0;\n"
        "  if (g6_open(0) != 0) return 1;\n"
        "  g6_set_j_particle(0, 0, 0, 0.0, 0.0, 1.0, z, z, z, z, x);\n"
        "  g6_set_j_particle_multisend_mxfast_(&cl, &n, &addr, &id, &m, x);\n"
        "  int r = unit_a() + g6_reset(0) + g6_initialize_jp_buffer(0, 1);\n"
        "  return g6_close(0) + (r < 0);\n"
        "}\n")
    objs = []
    for src, cc, std in (("a.c", "gcc", "-std=c99"), ("b.cpp", "g++", "-std=c++11")):
        obj = tmp_path / (src + ".o")
        subprocess.check_call([cc, std, "-Wall", "-Werror", "-c", "-I", INC, str(tmp_path / src), "-o", str(obj)])
        objs.append(str(obj))
    exe = tmp_path / "two_units"
    subprocess.check_call(["g++", "-o", str(exe)] + objs + ["-L", os.path.dirname(sa.LIB_PATH), "-lsapporo_amd"])
    assert exe.exists()


def test_wrapper_integrator_builds(tmp_path):
    """The shared-step integrator written against the caller-side names only compiles
    warning-free and links unchanged (tests/integrator/hermite4_wrap_g6.cpp)."""
    exe = tmp_path / "h4w"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", INC, "-o", str(exe),
                           os.path.join(ROOT, "tests", "integrator", "hermite4_wrap_g6.cpp"),
                           "-L", os.path.dirname(sa.LIB_PATH), "-lsapporo_amd"])
    assert exe.exists()


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
