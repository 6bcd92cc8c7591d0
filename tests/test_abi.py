"""The C ABI: the built library loads and exports every entry point include/dtdiff.h declares
(no device call is made, so this runs without a GPU)."""
import ctypes
import os
import re

import dtdiff

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(REPO, "include", "dtdiff.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dtd_[a-z0-9_]+)\s*\(", text)))


def test_header_and_python_agree_on_the_entry_points():
    assert declared() == sorted(dtdiff.SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dtdiff.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_default_opts_are_the_reference_constants():
    # multigrid.rs:45-56 (omega 0.8, nu 2, scale 0.5), diffusion.rs:497 (400 iterations).
    o = dtdiff.default_opts()
    assert (o.device, o.nu, o.max_iter) == (0, 2, 400)
    assert o.omega == ctypes.c_float(0.8).value and o.scale == 0.5
    assert o.cycle_precision == dtdiff.CYCLE_F32 and o.guess == dtdiff.GUESS_LOD


def test_null_arguments_are_rejected_without_a_device():
    lib = dtdiff.load()
    assert lib.dtd_create(None, None) == dtdiff.DTD_EINVAL
    assert lib.dtd_last_error(None) == b"null context"
    bad = dtdiff.Opts()
    lib.dtd_default_opts(ctypes.byref(bad))
    bad.nu = 0
    out = ctypes.c_void_p()
    assert lib.dtd_create(ctypes.byref(bad), ctypes.byref(out)) == dtdiff.DTD_EINVAL
    assert not out.value
