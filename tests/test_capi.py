"""The C ABI library itself, without a GPU: it loads, exports every function that
include/marfa.h declares (and nothing the Python binding needs is missing), and refuses
to run without a HIP device instead of computing on the CPU."""
import ctypes as C
import os
import re

import pytest

from marfa_amd import lbl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "marfa.h")


def declared():
    txt = open(HEADER).read()
    return set(re.findall(r"^(?:int|const char \*)\s*(marfa_\w+)\(", txt, re.M))


def test_header_declares_binding_exports():
    assert declared() == set(lbl.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = lbl.library()
    for name in sorted(declared()):
        assert hasattr(L, name), name
    # the symbols are exported with C linkage (no C++ mangling)
    nm = os.popen(f"nm -D --defined-only {lbl.LIB_PATH}").read()
    for name in declared():
        assert re.search(rf"\bT {name}$", nm, re.M), name


def test_version():
    assert lbl.library().marfa_version() >= 1


def test_null_arguments_are_errors():
    L = lbl.library()
    assert L.marfa_ctx_create(0, None) == 20           # MARFA_E_ARG
    assert L.marfa_ctx_destroy(None) == 20
    n = C.c_int32(-1)
    assert L.marfa_windows(None, None, 0, C.byref(n), None, None, None, None) == 20


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_no_device_is_an_error():
    L = lbl.library()
    h = C.c_void_p()
    assert L.marfa_ctx_create(0, C.byref(h)) == 23     # MARFA_E_NODEVICE
    with pytest.raises(lbl.MarfaError):
        lbl.LBLContext(0)
