"""The C-ABI library loads and exports every entry point include/hifde_amd.h
declares (no compute: this runs without a GPU)."""
import ctypes
import os
import re

import hifde3d_amd

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                   "hifde_amd.h")


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hifde_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_surface():
    names = declared_functions()
    assert len(names) >= 20
    assert set(names) == set(hifde3d_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hifde3d_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_sets_signatures():
    L = hifde3d_amd.lib()
    for n in hifde3d_amd.EXPORTS:
        assert getattr(L, n).argtypes is not None or n in ("hifde_last_error",)


def test_bad_arguments_return_error_codes_without_gpu():
    L = hifde3d_amd.lib()
    h = ctypes.c_void_p()
    # (N+1)/P not a power of two: rejected before any device call
    assert L.hifde_create(14, 4, 1e-6, 0, 0, ctypes.byref(h)) == -1
    assert L.hifde_create(15, 4, -1.0, 0, 0, ctypes.byref(h)) == -1
    assert L.hifde_create(15, 4, 1e-6, 7, 0, ctypes.byref(h)) == -1
    assert b"power of two" in L.hifde_last_error() or b"precision" in L.hifde_last_error()
    assert L.hifde_num_levels(None) == -1


def test_library_has_no_unresolved_internal_symbols():
    import subprocess
    out = subprocess.run(["nm", "-D", "--undefined-only", hifde3d_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    assert "hifde" not in out, out
    # load with immediate binding, as the runtime loader on the GPU host does
    ctypes.CDLL(hifde3d_amd.LIB_PATH, mode=os.RTLD_NOW)


def test_cpp_dropin_header_compiles_and_links():
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for prog in ("hif_driver", "free_funcs"):
        src = os.path.join(root, "tests", "cpp", prog + ".cpp")
        subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(root, "include"),
                        src], check=True)
        drv = os.path.join(root, "tests", "cpp", prog)
        assert os.path.exists(drv), f"build() compiles tests/cpp/{prog}"
        out = subprocess.run(["ldd", drv], capture_output=True, text=True).stdout
        assert "libhifde_amd.so" in out
        assert "not found" not in out.split("libhifde_amd.so")[1].split("\n")[0]
