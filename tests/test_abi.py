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


def test_shipped_gemm_table_is_well_formed():
    """hifde3d_amd/gemm_tune.txt (written by hifde3d_amd/tune.py): every entry
    is a valid class and kind, each class appears once, and the library
    accepts the file (parsing needs no GPU)."""
    path = os.path.join(os.path.dirname(hifde3d_amd.LIB_PATH), "gemm_tune.txt")
    assert os.path.exists(path)
    seen = set()
    for line in open(path):
        line = line.split("#")[0].split()
        if not line:
            continue
        es, ta, lo, kb, mb, nb, kind = map(int, line)
        assert es in (4, 8) and ta in (0, 1) and lo in (0, 1)
        assert 0 <= kb < 8 and 0 <= mb < 6 and 0 <= nb < 5 and kind in (0, 1, 2)
        assert (es, ta, lo, kb, mb, nb) not in seen
        seen.add((es, ta, lo, kb, mb, nb))
    assert len(seen) > 50
    L = hifde3d_amd.lib()
    assert L.hifde_gemm_table_load(path.encode()) == len(seen)
    assert L.hifde_gemm_table_load(b"/nonexistent/table.txt") == -1
