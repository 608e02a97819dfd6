"""The reference's plugin entry point run(void*, void*, void*, const char*,
const char*) (rhic/include/HydroPlugin.h:11), exported by libvah.so with the
reference's C++ linkage, and its C-linkage twin vah_plugin_run.  CPU checks:
the mangled symbol is exported, a caller written as the reference's Run.cpp
(tests/plugin_caller.cpp) compiles and links against the library, and the
argument checks return before any device call.  The run itself is compared
with the gpu-vah command line in tests/test_gpu_driver.py."""
import ctypes
import os
import shutil
import subprocess

import pytest

from tests.parity import vah

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "gpu-vah_amd")
CALLER = os.path.join(ROOT, "tests", "plugin_caller.cpp")
MANGLED = "_Z3runPvS_S_PKcS1_"


def build_caller(out_dir):
    """Compile and link the Run.cpp-shaped caller against libvah.so."""
    exe = os.path.join(str(out_dir), "plugin_caller")
    subprocess.run(["g++", "-O1", "-o", exe, CALLER, f"-L{LIBDIR}", "-lvah", f"-Wl,-rpath,{LIBDIR}"], check=True,
                   capture_output=True, text=True)
    return exe


def test_run_symbol_has_the_reference_signature():
    V = vah()
    V.lib()
    syms = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, "libvah.so")], check=True,
                          capture_output=True, text=True).stdout.split()
    assert MANGLED in syms and "vah_plugin_run" in syms
    if shutil.which("c++filt"):
        dem = subprocess.run(["c++filt", MANGLED], check=True, capture_output=True, text=True).stdout.strip()
        assert dem == "run(void*, void*, void*, char const*, char const*)"


def test_reference_shaped_caller_links(tmp_path):
    exe = build_caller(tmp_path)
    assert os.path.exists(exe)
    undef = subprocess.run(["nm", "-u", exe], check=True, capture_output=True, text=True).stdout
    assert MANGLED in undef  # the caller binds the C++ symbol, resolved from libvah.so


def test_plugin_run_argument_checks():
    V = vah()
    L = V.lib()
    f = L.vah_plugin_run
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_char_p] * 2
    f.restype = ctypes.c_int
    blk = ctypes.create_string_buffer(64)
    assert f(None, blk, blk, b".", b"/tmp") == V.VAH_ERR_ARG
    assert f(blk, None, blk, b".", b"/tmp") == V.VAH_ERR_ARG
    assert f(blk, blk, None, b".", b"/tmp") == V.VAH_ERR_ARG
    assert f(blk, blk, blk, b".", None) == V.VAH_ERR_ARG
