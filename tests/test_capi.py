"""CPU checks of the drop-in boundary: the library loads, every function the
headers in include/ declare is exported, and the C++ templates resolve to the
reference's mangled names (no GPU calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

import xss_testlib as T

LIB = os.path.join(T.REPO, "x86-simd-sort_amd", "libxss_amd.so")
INC = os.path.join(T.REPO, "include")


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_library_loads():
    L = ctypes.CDLL(LIB)
    L.xss_gpu_version.restype = ctypes.c_char_p
    assert L.xss_gpu_version().startswith(b"xss-amd")


def test_every_c_declaration_is_exported():
    pre = subprocess.run(["gcc", "-E", "-P", "-x", "c", os.path.join(INC, "xss_gpu.h")], capture_output=True,
                         text=True, check=True).stdout
    names = set(re.findall(r"\b((?:xss_gpu_|xss_host_|keyvalue_qsort_)\w+)\s*\(", pre))
    assert len(names) >= 210, len(names)
    assert {f"xss_host_argsort_arg_{t}" for t in T.TAGS} <= names
    missing = sorted(names - exported())
    assert not missing, missing


def test_cpp_templates_resolve_to_library_symbols(tmp_path):
    """Instantiate every template the reference exports (lib/x86simdsort.cpp)
    from a translation unit compiled by g++ against include/x86simdsort.h and
    check that each symbol it needs is defined by the library."""
    types = ["uint16_t", "int16_t", "float", "double", "uint32_t", "int32_t", "uint64_t", "int64_t"]
    kv = ["uint32_t", "int32_t", "float", "uint64_t", "int64_t", "double"]
    src = ['#include "x86simdsort.h"', "void use() {"]
    for t in types:
        src += [f"{{ {t} *a = nullptr; x86simdsort::qsort(a, 0); x86simdsort::qselect(a, 0, 0);",
                f"  x86simdsort::partial_qsort(a, 0, 0); x86simdsort::argsort(a, 0); x86simdsort::argselect(a, 0, 0); }}"]
    for k in kv + ["uint16_t", "int16_t"]:  # 16-bit keys: beyond the reference's 36 pairs
        for v in kv:
            src += [f"{{ {k} *a = nullptr; {v} *b = nullptr; x86simdsort::keyvalue_qsort(a, b, 0);",
                    f"  x86simdsort::keyvalue_select(a, b, 0, 0); x86simdsort::keyvalue_partial_sort(a, b, 0, 0); }}"]
    src += ["}"]
    f = tmp_path / "use.cpp"
    f.write_text("\n".join(src))
    obj = tmp_path / "use.o"
    subprocess.run(["g++", "-std=c++17", "-c", "-I", INC, str(f), "-o", str(obj)], check=True)
    und = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout
    need = {ln.split()[-1] for ln in und.splitlines() if "x86simdsort" in ln}
    assert len(need) == 8 * 5 + (36 + 12) * 3
    missing = sorted(need - exported())
    assert not missing, missing


@pytest.mark.skipif(T.ref() is None, reason="oracle/_ref not built")
def test_exports_cover_the_reference_library():
    """Every x86simdsort symbol the reference's own build exports."""
    out = subprocess.run(["nm", "-D", "--defined-only", T.REF_SO], capture_output=True, text=True, check=True).stdout
    ref_syms = {ln.split()[-1] for ln in out.splitlines() if "x86simdsort" in ln}
    assert len(ref_syms) >= 150
    assert not sorted(ref_syms - exported())


def test_static_api_resolves_to_library_symbols(tmp_path):
    """include/x86simdsort-static-incl.h (the reference's x86simdsortStatic::
    names, src/x86simdsort-static-incl.h:7-71): every function, including the
    NumPy arg forms, compiles with g++ and links to symbols the library
    defines."""
    types = ["uint16_t", "int16_t", "float", "double", "uint32_t", "int32_t", "uint64_t", "int64_t"]
    src = ['#include "x86simdsort-static-incl.h"', "void use() {"]
    for t in types:
        src += [f"{{ {t} *a = nullptr; size_t *g = nullptr; x86simdsortStatic::qsort(a, 0);",
                f"  x86simdsortStatic::qselect(a, 0, 0); x86simdsortStatic::partial_qsort(a, 0, 0);",
                f"  x86simdsortStatic::argsort(a, 0); x86simdsortStatic::argsort(a, g, 0, true, true);",
                f"  x86simdsortStatic::argselect(a, 0, 0); x86simdsortStatic::argselect(a, g, 0, 0, true); }}"]
    src += ["{ float *k = nullptr; uint64_t *v = nullptr; x86simdsortStatic::keyvalue_qsort(k, v, 0);",
            "  x86simdsortStatic::keyvalue_select(k, v, 0, 0); x86simdsortStatic::keyvalue_partial_sort(k, v, 0, 0); }",
            "}"]
    f = tmp_path / "use_static.cpp"
    f.write_text("\n".join(src))
    obj = tmp_path / "use_static.o"
    subprocess.run(["g++", "-std=c++17", "-c", "-I", INC, str(f), "-o", str(obj)], check=True)
    und = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout
    need = {ln.split()[-1] for ln in und.splitlines() if "x86simdsort" in ln or "xss_" in ln}
    assert {f"xss_host_argsort_arg_{t}" for t in ("u16", "i16", "f32", "f64", "u32", "i32", "u64", "i64")} <= need
    missing = sorted(need - exported())
    assert not missing, missing


def test_cmake_package_links_the_dropin_program(tmp_path):
    """x86-simd-sort_amd/cmake/xss_amdConfig.cmake (the counterpart of the
    reference's pkg-config file, meson.build:72-83): a CMake project finds
    the package, and the C++ drop-in program links against its target (it is
    not run: no GPU calls here)."""
    import shutil
    if shutil.which("cmake") is None:
        pytest.skip("cmake not installed")
    proj = tmp_path / "proj"
    proj.mkdir()
    (proj / "CMakeLists.txt").write_text(
        "cmake_minimum_required(VERSION 3.16)\n"
        "project(use_xss CXX)\n"
        "find_package(xss_amd 7.0 CONFIG REQUIRED)\n"
        f"add_executable(dropin {os.path.join(T.REPO, 'tests', 'cpp', 'dropin.cpp')})\n"
        "target_link_libraries(dropin PRIVATE xss_amd::x86simdsortcpp)\n")
    build = tmp_path / "build"
    subprocess.run(["cmake", "-S", str(proj), "-B", str(build), "-DCMAKE_CXX_COMPILER=g++",
                    f"-Dxss_amd_DIR={os.path.join(T.REPO, 'x86-simd-sort_amd', 'cmake')}"],
                   check=True, capture_output=True, text=True)
    subprocess.run(["cmake", "--build", str(build)], check=True, capture_output=True, text=True)
    exe = build / "dropin"
    assert exe.exists()
    needed = subprocess.run(["readelf", "-d", str(exe)], capture_output=True, text=True, check=True).stdout
    assert "libxss_amd.so" in needed
