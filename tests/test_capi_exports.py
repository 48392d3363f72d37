"""The C-ABI library loads without a GPU and exports every entry point the
public headers declare; the Python binding covers exactly that set.  No call
here touches a device."""
import ctypes as C
import os
import re
import subprocess

from opendeformer_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "od.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(od_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    assert "od_pcg_solve" in names and "od_newmark_create" in names and len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (od_[a-z0-9_]+)$", out.stdout, flags=re.M))
    assert set(declared()) <= exported


def test_python_binding_covers_the_header():
    assert sorted(set(capi.EXPORTS)) == declared()


def test_version_and_error_text_without_device():
    lib = capi.lib()
    assert lib.od_version() > 0
    assert isinstance(lib.od_last_error(), bytes)


def test_host_header_compiles_against_the_abi():
    src = os.path.join(ROOT, "tests", "cpp", "host_api_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_server_rejects_bad_parameters_before_touching_a_device():
    import ctypes
    lib = capi.lib()
    h = ctypes.c_void_p()
    # zero budget, too many slots, unknown precision: argument errors, no device call made
    assert lib.od_server_create(0, 0, 4, capi.OD_MIXED, capi.OD_PRE_IC0, 1e-6, ctypes.byref(h)) == capi.OD_ERR_ARG
    assert lib.od_server_create(0, 1 << 30, 64, capi.OD_MIXED, capi.OD_PRE_IC0, 1e-6, ctypes.byref(h)) == capi.OD_ERR_ARG
    assert lib.od_server_create(0, 1 << 30, 4, 7, capi.OD_PRE_IC0, 1e-6, ctypes.byref(h)) == capi.OD_ERR_UNSUPPORTED
    assert b"server" in lib.od_last_error() or b"precision" in lib.od_last_error()
